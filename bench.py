#!/usr/bin/env python3
"""Benchmark of the WCSPH substep engine on one MI355X.

Headline workload: the 1,048,576-particle obstacle scene (128 x 64 x 128
lattice, icosphere mesh obstacle with 128^3 SDF/volume map, h = 0.012,
dt = 1e-4) in its contact phase: settled 1,500 substeps from rest so that
pressure acts on most particles and thousands of keys change bucket per
substep.  A "step" is one substep: hash, stable sort, cell ranges, reorder,
boundary + density, pressure + viscosity + integrate, over every particle.
Inputs are resident in HBM before the timed region; the timed region replays
the captured substep graph.

Beside it: the same scene from rest as configured (viscosity 0.01, free fall
until it diverges at the first contact), the 8,388,608-particle sort stress,
256 batched scenes against per-scene loops, forward + adjoint against
forward, the Lague dam break, and the CPU oracle on a bounded sample.

Prints one JSON line: particle-substeps/s, the HBM roofline of the dominant
kernel (algorithmic bytes per launch / mean HIP-event launch time on the
simulation stream), its measured HBM traffic per launch (child runs of this
script under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, scaled by a copy of
known size), its VALU issue rate against the peak, and the CPU oracle.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "particle-steps/s @1M WCSPH; batched vs looped scenes/s; adjoint/forward cost"
HBM_PEAK = 8.0e12
# Algorithmic bytes per particle-substep, per phase: every array read or
# written once, neighbour re-reads as cache hits, grids L2-resident.
ALG_BYTES = {"hash": 16, "sort": 28, "cell_ranges": 8, "reorder": 52, "density": 16, "forces_integrate": 52}
B_ALG = sum(ALG_BYTES.values())  # 172
# the memory-bound phases timed as one class: sort (with the reorder in its
# merge) and cell ranges; the hash is fused into the force pass
MEM_CLASS_BYTES = ALG_BYTES["sort"] + ALG_BYTES["cell_ranges"] + ALG_BYTES["reorder"]  # 88


def stage_bytes(list_mean: float, reorder_in_sort: bool = False) -> dict:
    """Bytes per particle that this engine's stages must move, with the
    neighbour list (4 bytes per entry, list_mean entries per particle) that the
    density pass writes and the force pass reads (reported beside the
    algorithmic bytes, never as them).  A re-sorting substep reorders the
    state in the sort's merge kernel (reorder_in_sort): the reorder's bytes
    less the permutation's re-read count to the sort."""
    L = 4.0 * list_mean
    return {
        # keys in, sorted keys + permutation out, table fill
        "sort": 28.0 + (48.0 if reorder_in_sort else 0.0),
        "cell_ranges": 8.0,        # offsets table in, cell starts out
        "reorder": 52.0,           # permutation 4, gather x v 32, write x v 32 (float4) less the w lanes
        # positions 16 in; rho 4, boundary (grad, psi) 16, list count 4, EOS terms 8 out; list out
        "density": 48.0 + L,
        # x v rho boundary count 56 in + list; x v key 36 out
        "forces_integrate": 92.0 + L,
    }


# kernel-name prefixes of the stages of a re-sorting substep (every substep
# after the first) in the rocprofv3 databases; the merge also reorders x and v
STAGE_KERNELS = {"density": "k_density", "forces_integrate": "k_forces_integrate",
                 "sort_merge": "k_resort_merge", "cell_ranges": "k_resort_cell_starts",
                 "sort_count": "k_resort_count", "sort_compact": "k_resort_compact"}
CALIBRATION_MIB = 256


def describe_shape(shape) -> str:
    kind, sp, res = shape
    grid = "x".join(str(int(r)) for r in res)
    if kind == "mesh":
        return f"triangle mesh, {len(sp[0])} vertices / {len(sp[1])} triangles, {grid} grids"
    names = {0: "analytic box", 1: "analytic sphere", 2: "analytic box walls"}
    return f"{names.get(kind, str(kind))}, {grid} grids"


def kernel_base(name: str) -> str:
    """'void vfs::(anonymous namespace)::k_density<128, true>(...)' -> 'k_density'."""
    name = name.replace("(anonymous namespace)::", "")
    return name.split("(")[0].split("<")[0].split("::")[-1].split(" ")[-1]


def measure_traffic(workload: str, state: str | None, timeout: float = 240.0) -> dict:
    """HBM bytes per launch of each stage kernel: two child runs of this script
    under rocprofv3 --pmc (FETCH_SIZE, then WRITE_SIZE), each counter scaled by
    what it reads for a device-to-device copy of CALIBRATION_MIB MiB (which
    moves that many bytes each way).  Returns {} when rocprofv3 is unavailable
    or fails; runs before this process touches the GPU."""
    import shutil
    import sqlite3
    import subprocess
    import tempfile

    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not Path(prof).exists():
        return {"error": "rocprofv3 not found"}
    out = {}
    env = dict(os.environ, TMPDIR="/tmp")
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        per_counter = {}
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            d = Path(tmp) / counter
            cmd = [prof, "--pmc", counter, "-d", str(d), "-o", "run", "--", sys.executable, str(ROOT / "bench.py"),
                   "--traffic-probe", "--workload", workload] + (["--state", state] if state else [])
            try:
                r = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, text=True, timeout=timeout)
            except subprocess.TimeoutExpired:
                return {"error": f"rocprofv3 --pmc {counter} timed out"}
            dbs = list(d.rglob("*results.db"))
            if r.returncode != 0 and not dbs:
                return {"error": f"rocprofv3 --pmc {counter} failed: {r.stderr[-300:]}"}
            if not dbs:
                return {"error": f"rocprofv3 --pmc {counter}: no database"}
            rows = sqlite3.connect(str(dbs[0])).execute(
                "select kernel_name, value, grid_size from counters_collection where counter_name = ?",
                (counter,)).fetchall()
            per_counter[counter] = rows
        calib = {}
        for counter, rows in per_counter.items():
            copies = [(g, v) for name, v, g in rows if "copyBuffer" in name]
            if not copies:
                return {"error": "calibration copy not found"}
            calib[counter] = (CALIBRATION_MIB << 20) / (max(copies)[1] * 1024.0)
        for stage, prefix in STAGE_KERNELS.items():
            tot = 0.0
            for counter, rows in per_counter.items():
                vals = [v for name, v, g in rows if kernel_base(name) == prefix]
                if not vals:
                    tot = None
                    break
                tot += sum(vals) / len(vals) * 1024.0 * calib[counter]
            if tot is not None:
                out[stage] = tot
        out["calibration"] = {k: round(v, 4) for k, v in calib.items()}
    return out


# VALU issue of the neighbour kernels: one more child run under rocprofv3
# --pmc with these SQ counters (totals over the device per dispatch)
VALU_COUNTERS = ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_THREAD_CYCLES_VALU")
SIMDS = 256 * 4          # MI355X: 256 CUs of 4 SIMD16 units; a wave64 VALU instruction issues over 4 cycles
PEAK_CLOCK_GHZ = 2.4     # MI355X peak engine clock
VALU_PEAK = SIMDS * PEAK_CLOCK_GHZ * 1e9 / 4.0  # wave64 VALU instructions per second


def measure_valu(workload: str, state: str | None, timeout: float = 240.0) -> dict:
    """Per launch of the density and force kernels: VALU instructions per wave,
    the fraction of lanes active in them, and the VALU instructions as cycles
    per SIMD (4 per wave64 instruction) for bench-side timing to compare with
    the kernel's duration.  {} when rocprofv3 is unavailable or fails."""
    import shutil
    import sqlite3
    import subprocess
    import tempfile

    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not Path(prof).exists():
        return {}
    env = dict(os.environ, TMPDIR="/tmp")
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        d = Path(tmp) / "valu"
        cmd = [prof, "--pmc", *VALU_COUNTERS, "-d", str(d), "-o", "run", "--", sys.executable,
               str(ROOT / "bench.py"), "--traffic-probe", "--workload", workload] + (["--state", state] if state else [])
        try:
            subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, text=True, timeout=timeout)
        except subprocess.TimeoutExpired:
            return {}
        dbs = list(d.rglob("*results.db"))
        if not dbs:
            return {}
        rows = sqlite3.connect(str(dbs[0])).execute(
            "select kernel_name, counter_name, value from counters_collection").fetchall()
    out = {}
    for stage in ("density", "forces_integrate"):
        prefix = STAGE_KERNELS[stage]
        per = {}
        for name, counter, value in rows:
            if kernel_base(name) == prefix:
                per.setdefault(counter, []).append(value)
        if not all(c in per for c in VALU_COUNTERS):
            continue
        mean = {c: sum(v) / len(v) for c, v in per.items()}
        out[stage] = {
            "valu_insts": mean["SQ_INSTS_VALU"],
            "valu_insts_per_wave": mean["SQ_INSTS_VALU"] / mean["SQ_WAVES"],
            "lane_utilization": mean["SQ_THREAD_CYCLES_VALU"] / (64.0 * mean["SQ_ACTIVE_INST_VALU"]),
            "valu_cycles_per_simd": 4.0 * mean["SQ_INSTS_VALU"] / SIMDS,
        }
    return out


def candidates_per_particle(x: np.ndarray, cell: float, n: int) -> float:
    """Mean candidates that a particle's cell walk tests on the state x: the
    sizes of the 27 buckets of its stencil, summed (a bucket is visited again
    when two stencil cells share a key), with the substep's cell hash
    (HashCell3D mod n) restated in numpy."""
    c = np.floor(x / np.float32(cell)).astype(np.int64)

    def key(cc):
        u = (cc + 25) & 0xFFFFFFFF
        lo, hi = u % 50, u // 50
        h = lo[:, 0] + 50 * (lo[:, 1] + 50 * lo[:, 2]) + hi[:, 0] * 15823 + hi[:, 1] * 9737333 + hi[:, 2] * 440817757
        return (h & 0xFFFFFFFF) % n

    counts = np.bincount(key(c), minlength=n)
    total = 0
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                total += int(counts[key(c + np.array([dx, dy, dz], np.int64))].sum())
    return total / max(1, len(x))


def workload_params(V, w, flags: int = 0):
    return V.make_params(w.n, w.h, dt=w.dt, box_size=w.box_size, box_pos=w.box_pos, flags=flags, **w.params_kw)


def settled_state(w, grids, path: str | None = None):
    """(x, v) of workload w after its settling substeps: read from `path` (an
    npz that a --make-state child of this run wrote) or computed here."""
    import vkfluidsim_amd as V

    if path and Path(path).exists():
        d = np.load(path)
        return d["x"], d["v"]
    if w.settle == 0:
        return w.x, w.v
    sim = V.Sim(workload_params(V, w), grids)
    sim.set_state(w.x, w.v)
    sim.substeps(w.settle)
    x, v, _, _ = sim.get_state()
    sim.close()
    return x, v


def make_state(workload: str, path: str) -> int:
    """Child: settle the workload and write its state for the probes and the parent."""
    import vkfluidsim_amd as V
    from vkfluidsim_amd import configs as CF

    w = CF.WORKLOADS[workload]()
    x, v = settled_state(w, V.boundary_grids(w.shapes, w.h))
    np.savez(path, x=x, v=v)
    return 0


def traffic_probe(workload: str, state: str | None) -> int:
    """Child of measure_traffic / measure_valu: one calibration copy, then
    three substeps from the workload's (settled) state."""
    import vkfluidsim_amd as V
    from vkfluidsim_amd import configs as CF

    nbytes = CALIBRATION_MIB << 20
    src, dst = V.DeviceArray(nbytes), V.DeviceArray(nbytes)
    V._lib.call("vfs_memcpy_d2d", dst.ptr, src.ptr, nbytes)
    src.free()
    dst.free()
    w = CF.WORKLOADS[workload]()
    grids = V.boundary_grids(w.shapes, w.h)
    x, v = settled_state(w, grids, state)
    sim = V.Sim(workload_params(V, w), grids)
    sim.set_state(x, v)
    sim.substeps(3)
    sim.synchronize()
    sim.close()
    return 0


def cpu_baseline(w, grids, x, v, steps: int) -> dict:
    """The FP32 CPU oracle (OpenMP) on the same workload from the same state."""
    sys.path.insert(0, str(ROOT / "tests"))
    import harness as H
    import vkfluidsim_amd as V

    threads = H.threads()
    p = workload_params(V, w)
    o = H.OracleSim(p, grids, nthreads=threads)
    o.set_state(x, v)
    o.substeps(1)  # first touch / page-in
    t0 = time.perf_counter()
    o.substeps(steps)
    dt = time.perf_counter() - t0
    del o
    return {"value": w.n * steps / dt, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"{steps} substeps of the {w.n}-particle workload ({w.name}, from the timed state) after "
                      f"its first substep (FP32 oracle, OpenMP, {dt:.1f} s)"}


def bench_batched(n_scenes: int, steps: int, looped_steps: int) -> dict:
    """256 independent dam breaks of uneven sizes (log-uniform 2,048-65,536
    particles, own boxes and 32^3 wall maps): one vfs_batch (autotuned launch
    configuration, cell size included) against a loop over one vfs_sim per
    scene, each replaying its own substep graph.  Throughput in
    particle-substeps/s and scene-substeps/s."""
    import ctypes as C

    import vkfluidsim_amd as V
    from vkfluidsim_amd import configs as CF

    ws = CF.batched_dam_breaks(n_scenes=n_scenes)
    grids = [V.boundary_grids(w.shapes, w.h) for w in ws]
    params = [V.make_params(w.n, w.h, dt=w.dt, box_size=w.box_size, box_pos=w.box_pos) for w in ws]
    total = sum(w.n for w in ws)

    batch = V.Batch(params, grids)
    for k, w in enumerate(ws):
        batch.set_state(k, w.x, w.v)
    batch.substeps(2)
    batch.synchronize()
    tuning, tuned_ms = batch.autotune(3, allow_cell_size=True)
    batch.substeps(5)
    batch.synchronize()
    t0 = time.perf_counter()
    batch.substeps(steps)
    batch.synchronize()
    tb = (time.perf_counter() - t0) / steps
    finite = all(bool(np.isfinite(batch.get_state(k)[0]).all()) for k in range(n_scenes))
    list_mean, list_max, list_over = batch.neighbour_stats()
    stages = batch.profile(3)
    batch.close()
    out = {
        "workload": f"{n_scenes} dam breaks, {total} particles (log-uniform 2048-65536 per scene), 32^3 wall maps",
        "batched": {"particle_steps_per_s": total / tb, "scene_steps_per_s": n_scenes / tb,
                    "ms_per_substep": tb * 1e3, "substeps": steps,
                    "tuning": list(tuning.as_tuple()), "tuning_keeps_reference_hash": tuning.cell_scale == 1.0,
                    "stages_ms": {name: round(ms, 4) for name, ms, _ in stages}, "finite": finite,
                    "neighbour_list": {"mean": round(list_mean, 3), "max": list_max, "overflowed": list_over}},
    }
    if looped_steps == 0:
        return out

    sims = []
    for k, w in enumerate(ws):
        sim = V.Sim(params[k], grids[k])
        sim.set_state(w.x, w.v)
        sim.substeps(1)
        sims.append(sim)
    for sim in sims:
        sim.synchronize()
    # (a) every scene on its own stream: graphs of different scenes may overlap
    t0 = time.perf_counter()
    for _ in range(looped_steps):
        for sim in sims:
            sim.substeps(1)
    for sim in sims:
        sim.synchronize()
    tl = (time.perf_counter() - t0) / looped_steps
    # (b) one stream for the whole loop: scene after scene
    stream = C.c_void_p()
    V._lib.call("vfs_stream_create", C.byref(stream))
    for sim in sims:
        sim.substeps(1, stream)
    V._lib.call("vfs_stream_synchronize", stream)
    t0 = time.perf_counter()
    for _ in range(looped_steps):
        for sim in sims:
            sim.substeps(1, stream)
    V._lib.call("vfs_stream_synchronize", stream)
    ts = (time.perf_counter() - t0) / looped_steps
    V._lib.call("vfs_stream_destroy", stream)
    for sim in sims:
        sim.close()
    out.update({
        "looped": {"particle_steps_per_s": total / ts, "scene_steps_per_s": n_scenes / ts,
                   "ms_per_substep": ts * 1e3, "substeps": looped_steps,
                   "how": "one vfs_sim per scene, each replaying its substep graph, scene after scene on one stream"},
        "looped_own_streams": {"particle_steps_per_s": total / tl, "scene_steps_per_s": n_scenes / tl,
                               "ms_per_substep": tl * 1e3, "substeps": looped_steps,
                               "how": "as looped, each scene on its own stream (graphs of different scenes "
                                      "may overlap on the GPU)"},
        "speedup": ts / tb,
        "speedup_vs_own_streams": tl / tb,
    })
    return out


def bench_adjoint(steps: int, every: int, repeats: int = 3) -> dict:
    """Differentiable dam break: 65,536 particles (64 x 32 x 32 lattice, spacing
    0.02, the 1.6 x 0.8 x 0.8 walled box), `steps` substeps, loss = mean squared
    difference of the density on a 32^3 grid against the field of a run with
    initial velocities perturbed by N(0, 0.01); gradients with respect to the
    initial velocities (and positions) and the viscosity, checkpoints every
    `every` substeps.  Reports forward + adjoint time over plain forward time."""
    import vkfluidsim_amd as V
    from vkfluidsim_amd import configs as CF

    w = CF.dam_break(side=(64, 32, 32), seed=CF.SEED_BASE + 4, steps=steps)
    grids = V.boundary_grids(w.shapes, w.h)
    p = V.make_params(w.n, w.h, dt=w.dt, box_size=w.box_size, box_pos=w.box_pos)
    sim = V.Sim(p, grids)
    dims = (32, 32, 32)
    spacing = tuple(b / (d - 1) for b, d in zip(w.box_size, dims))
    adj = V.Adjoint(sim, steps, every, (0.0, 0.0, 0.0), spacing, dims)
    v_t = (0.01 * CF.normal(CF.SEED_BASE + 4, 3 * w.n)).astype(np.float32).reshape(w.n, 3)
    _, target = adj.forward(w.x, v_t)

    def plain():
        sim.set_state(w.x, w.v)
        sim.substeps(1)
        sim.synchronize()
        t0 = time.perf_counter()
        sim.substeps(steps)
        sim.synchronize()
        return time.perf_counter() - t0

    def fwd_adj():
        t0 = time.perf_counter()
        loss, _ = adj.forward(w.x, w.v, target)
        t1 = time.perf_counter()
        gx, gv, gnu = adj.backward()
        t2 = time.perf_counter()
        return t1 - t0, t2 - t1, loss, gv, gnu

    plain()
    fwd_adj()
    t_plain = min(plain() for _ in range(repeats))
    runs = [fwd_adj() for _ in range(repeats)]
    t_f = min(r[0] for r in runs)
    t_b = min(r[1] for r in runs)
    loss, gv, gnu = runs[-1][2], runs[-1][3], runs[-1][4]
    # one more reverse pass with an event pair around every launch group
    # (outside the timed runs): total ms per stage over the pass
    stages = adj.profile_backward()
    adj.close()
    sim.close()
    return {
        "workload": f"dam break {w.n} particles, {steps} substeps, checkpoints every {every}, 32^3 density grid",
        "forward_ms": t_plain * 1e3,
        "adjoint_forward_ms": t_f * 1e3,
        "adjoint_backward_ms": t_b * 1e3,
        "fwd_plus_adj_over_fwd": (t_f + t_b) / t_plain,
        "loss": loss,
        "grad_v0_norm": float(np.linalg.norm(gv)),
        "grad_viscosity": gnu,
        "backward_stages_ms": {name: round(ms, 3) for name, ms, _ in stages},
        "backward_stage_groups": {name: k for name, _, k in stages},
    }


def bench_reference_scenes(frames: int) -> dict:
    """The reference's own scenes, as a user of it runs them: assets/scenes/
    dam_break.json (75,000 particles) and obstacle.json (40,000 particles,
    box.obj rotated and scaled, suzanne.obj), loaded from their committed
    copies (tests/golden/reference_scene) and stepped frame by frame as
    GenericScene::Step does (one frame = `iterations` substeps).  Reports the
    set-up time (JSON, seeding, every boundary object's SDF and volume map on
    the GPU) and the time per frame."""
    import vkfluidsim_amd as V
    from vkfluidsim_amd._lib import call

    root = ROOT / "tests" / "golden" / "reference_scene"
    out = {}
    for name in ("dam_break", "obstacle"):
        t0 = time.perf_counter()
        sc = V.Scene(str(root / f"{name}.json"), str(root))
        call("vfs_sim_synchronize", sc.sim_handle)
        t1 = time.perf_counter()
        sc.step()  # first frame: full sort, graph capture
        call("vfs_sim_synchronize", sc.sim_handle)
        t2 = time.perf_counter()
        for _ in range(frames):
            sc.step()
        call("vfs_sim_synchronize", sc.sim_handle)
        t3 = time.perf_counter()
        it = int(sc.params.iterations)
        out[name] = {
            "particles": sc.n,
            "iterations_per_frame": it,
            "frames": frames,
            "setup_ms": round((t1 - t0) * 1e3, 1),
            "ms_per_frame": (t3 - t2) * 1e3 / frames,
            "frames_per_s": frames / (t3 - t2),
            "particle_steps_per_s": sc.n * it * frames / (t3 - t2),
        }
        sc.close()
    return out


def bench_lague(steps: int, cpu_steps: int, with_cpu: bool) -> dict:
    """The reference's Lague dam break (dam_break_scene.cpp:8-50): 150,000
    particles, target density 630, 23 x 10 x 10 box, smooth radius 0.2, three
    substeps of 1/360 per frame.  Particle-substeps/s of the graph-replayed
    substep, per-stage times, and the FP32 oracle (OpenMP) on `cpu_steps`
    substeps of the same state."""
    import vkfluidsim_amd as V
    from vkfluidsim_amd import configs as CF

    w = CF.lague_dam_break()
    p = V.make_params(w.n, w.h, dt=w.dt, box_size=w.box_size, box_pos=w.box_pos, rho0=w.params_kw["rho0"],
                      iterations=w.params_kw["iterations"])
    sim = V.Lague(p)
    sim.set_state(w.x, w.v)
    sim.substeps(20)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.substeps(steps)
    sim.synchronize()
    dt = (time.perf_counter() - t0) / steps
    stages = sim.profile(3)
    x, v, rho, _ = sim.get_state()
    out = {
        "workload": f"Lague dam break, {w.n} particles (50 x 60 x 50 at density 630), substeps 20-{20 + steps}",
        "particle_steps_per_s": w.n / dt,
        "ms_per_substep": dt * 1e3,
        "substeps": steps,
        "stages_ms": {name: round(ms, 4) for name, ms, _ in stages},
        "end_state": {"finite": bool(np.isfinite(x).all() and np.isfinite(v).all()),
                      "rho_mean": float(rho.mean()), "max_speed": float(np.linalg.norm(v, axis=1).max())},
    }
    sim.close()
    if with_cpu:
        sys.path.insert(0, str(ROOT / "tests"))
        import harness as H

        o = H.OracleLague(p, V.make_lague_params(), nthreads=H.threads())
        o.set_state(w.x, w.v)
        o.substeps(1)
        t0 = time.perf_counter()
        o.substeps(cpu_steps)
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": w.n * cpu_steps / tc, "unit": "particle-steps/s", "cores": H.threads(),
                               "kind": "port", "sample": f"{cpu_steps} substeps after the first (FP32 oracle, "
                                                         f"OpenMP, {tc:.1f} s)"}
    return out


def summarize_stats(db: str) -> str:
    """Text table of a rocprofv3 --kernel-trace --stats database (top_kernels,
    whose durations read in microseconds; printed in milliseconds)."""
    import sqlite3

    rows = sqlite3.connect(db).execute(
        "select name, total_calls, total_duration, average, percentage from top_kernels").fetchall()
    lines = [f"{'kernel':34s} {'calls':>6s} {'total_ms':>11s} {'avg_ms':>9s} {'pct':>6s}"]
    for name, calls, total, avg, pct in rows:
        short = name.replace("(anonymous namespace)::", "").replace("void ", "").split("(")[0].replace("vfs::", "")
        lines.append(f"{short:34s} {calls:6d} {total / 1e3:11.1f} {avg / 1e3:9.3f} {pct:6.2f}")
    return "\n".join(lines)


def summarize_stats_by_grid(db: str) -> str:
    """Per kernel and grid size (threads) of a rocprofv3 --kernel-trace
    database: calls, total and mean duration (the kernels view, in ns; printed
    in ms), so that runs of several scene sizes in one process separate."""
    import sqlite3

    rows = sqlite3.connect(db).execute(
        "select name, grid_x, count(*), sum(duration), avg(duration) from kernels group by name, grid_x "
        "order by sum(duration) desc").fetchall()
    lines = [f"{'kernel':34s} {'threads':>9s} {'calls':>6s} {'total_ms':>11s} {'avg_ms':>9s}"]
    for name, grid, calls, total, avg in rows:
        short = name.replace("(anonymous namespace)::", "").replace("void ", "").split("(")[0].replace("vfs::", "")
        lines.append(f"{short:34s} {grid:9d} {calls:6d} {total / 1e6:11.1f} {avg / 1e6:9.3f}")
    return "\n".join(lines)


SQ_COUNTER_PASSES = (
    "SQ_WAVES SQ_WAVE_CYCLES SQ_BUSY_CYCLES SQ_WAIT_ANY SQ_WAIT_INST_ANY SQ_ACTIVE_INST_ANY SQ_ACTIVE_INST_VALU "
    "SQ_ACTIVE_INST_LDS",
    "SQ_INSTS_VALU SQ_INSTS_LDS SQ_INSTS_SALU SQ_INSTS_VMEM_RD SQ_INSTS_VMEM_WR SQ_LDS_BANK_CONFLICT "
    "SQ_WAIT_INST_LDS SQ_INSTS_BRANCH",
    "SQ_ACTIVE_INST_SCA SQ_ACTIVE_INST_VMEM SQ_ACTIVE_INST_MISC SQ_INST_CYCLES_VMEM_RD SQ_LDS_IDX_ACTIVE "
    "SQ_LDS_ADDR_CONFLICT SQ_IFETCH SQ_THREAD_CYCLES_VALU",
)
# The vector memory path of the same kernels: texture address (TA) and data
# (TD) units and the L1 cache (TCP) of every CU, summed over the 256 CUs, with
# the kernel's cycles (GRBM_GUI_ACTIVE counts each of the 8 XCDs).
MEMORY_COUNTER_PASS = ("GRBM_GUI_ACTIVE TA_TA_BUSY_sum TA_BUFFER_READ_WAVEFRONTS_sum TD_TD_BUSY_sum TD_TC_STALL_sum "
                       "TCP_TOTAL_CACHE_ACCESSES_sum TCP_TCC_READ_REQ_sum TCP_PENDING_STALL_CYCLES_sum "
                       "TCP_TCR_TCP_STALL_CYCLES_sum")
CUS, XCDS = 256, 8


def summarize_counters(dbs: list) -> str:
    """Per-wave SQ counters of the density and force kernels from rocprofv3 --pmc
    databases of --traffic-probe runs (one database per pass of
    SQ_COUNTER_PASSES), with the VALU issue fraction against the kernel's
    duration from the same runs."""
    import sqlite3

    per = {}
    for db in dbs:
        con = sqlite3.connect(db)
        for name, counter, value in con.execute("select kernel_name, counter_name, value from counters_collection"):
            per.setdefault(kernel_base(name), {}).setdefault(counter, []).append(value)
    lines = []
    for k in ("k_density", "k_forces_integrate"):
        if k not in per or "SQ_WAVES" not in per[k]:
            continue
        mean = {c: sum(v) / len(v) for c, v in per[k].items()}
        waves = mean["SQ_WAVES"]
        lines.append(f"{k}: {waves:.0f} waves per launch (mean of {len(per[k]['SQ_WAVES'])} launches)")
        for c in sorted(mean):
            if c == "SQ_WAVES":
                continue
            lines.append(f"  {c:26s} {mean[c]:14.4g}   per wave {mean[c] / waves:10.1f}")
        if "SQ_THREAD_CYCLES_VALU" in mean and "SQ_ACTIVE_INST_VALU" in mean:
            lines.append(f"  lanes active per VALU instruction: "
                         f"{mean['SQ_THREAD_CYCLES_VALU'] / (64.0 * mean['SQ_ACTIVE_INST_VALU']):.3f}")
        if "GRBM_GUI_ACTIVE" in mean:
            cyc = mean["GRBM_GUI_ACTIVE"] / XCDS
            lines.append(f"  kernel cycles (GRBM_GUI_ACTIVE / {XCDS} XCDs): {cyc:.0f}")
            if "SQ_ACTIVE_INST_VALU" in mean:
                lines.append(f"  VALU busy (4 x SQ_ACTIVE_INST_VALU / {SIMDS} SIMDs / kernel cycles): "
                             f"{4.0 * mean['SQ_ACTIVE_INST_VALU'] / SIMDS / cyc:.3f}")
            for c in ("TA_TA_BUSY_sum", "TD_TD_BUSY_sum", "TD_TC_STALL_sum", "TCP_PENDING_STALL_CYCLES_sum"):
                if c in mean:
                    lines.append(f"  {c} per CU / kernel cycles: {mean[c] / CUS / cyc:.3f}")
            if "SQ_INSTS_VMEM_RD" in mean:
                lines.append(f"  TD data cycles (16 per wave64 16-byte load, SQ_INSTS_VMEM_RD) per CU / kernel cycles: "
                             f"{16.0 * mean['SQ_INSTS_VMEM_RD'] / CUS / cyc:.3f}")
            if "TCP_TOTAL_CACHE_ACCESSES_sum" in mean and "TCP_TCC_READ_REQ_sum" in mean:
                lines.append(f"  L1 hit rate (1 - TCP_TCC_READ_REQ / TCP_TOTAL_CACHE_ACCESSES): "
                             f"{1.0 - mean['TCP_TCC_READ_REQ_sum'] / mean['TCP_TOTAL_CACHE_ACCESSES_sum']:.3f}")
    return "\n".join(lines)


def summarize_pmc(fetch_db: str, write_db: str) -> str:
    """Per-kernel FETCH_SIZE / WRITE_SIZE (KiB per dispatch) of two rocprofv3
    --pmc databases of the traffic probe, and the HBM bytes per launch after
    scaling by the calibration copy."""
    import sqlite3

    per = {}
    for counter, db in (("FETCH_SIZE", fetch_db), ("WRITE_SIZE", write_db)):
        per[counter] = sqlite3.connect(db).execute(
            "select kernel_name, value, grid_size from counters_collection where counter_name = ?",
            (counter,)).fetchall()
    calib = {}
    for counter, rows in per.items():
        copies = [(g, v) for name, v, g in rows if "copyBuffer" in name]
        calib[counter] = (CALIBRATION_MIB << 20) / (max(copies)[1] * 1024.0)
    names = sorted({kernel_base(n) for rows in per.values() for n, _, _ in rows})
    lines = [f"calibration: {CALIBRATION_MIB} MiB device-to-device copy reads FETCH_SIZE x "
             f"{calib['FETCH_SIZE']:.4f}, writes WRITE_SIZE x {calib['WRITE_SIZE']:.4f}",
             f"{'kernel':30s} {'n':>3s} {'FETCH_KiB':>11s} {'WRITE_KiB':>11s} {'HBM_MB/launch':>14s}"]
    for k in names:
        if "copyBuffer" in k or "fillBuffer" in k:
            continue
        f = [v for n, v, _ in per["FETCH_SIZE"] if kernel_base(n) == k]
        w = [v for n, v, _ in per["WRITE_SIZE"] if kernel_base(n) == k]
        fa = sum(f) / len(f) if f else 0.0
        wa = sum(w) / len(w) if w else 0.0
        mb = (fa * calib["FETCH_SIZE"] + wa * calib["WRITE_SIZE"]) * 1024.0 / 1e6
        lines.append(f"{k:30s} {len(f):3d} {fa:11.1f} {wa:11.1f} {mb:14.1f}")
    return "\n".join(lines)


def time_substeps(sim, warmup: int, steps: int):
    """Seconds of `steps` graph-replayed substeps after `warmup` untimed ones,
    bracketed by stream synchronisation."""
    sim.substeps(warmup)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.substeps(steps)
    sim.synchronize()
    return time.perf_counter() - t0


def state_summary(sim, w) -> dict:
    """Sanity of a state and the work it gives the substep: finite, in the box,
    densities, pressure share, boundary share, list lengths."""
    list_mean, list_max, list_over = sim.neighbour_stats()
    x, v, rho, _ = sim.get_state()
    psi, _ = sim.get_boundary()
    lo = np.asarray(w.box_pos, np.float32)
    hi = lo + np.asarray(w.box_size, np.float32)
    speed = np.linalg.norm(v, axis=1) if w.n else np.zeros(1)
    return {"finite": bool(np.isfinite(x).all() and np.isfinite(v).all() and np.isfinite(rho).all()),
            "max_speed": float(speed.max()), "speed_p99": float(np.percentile(speed, 99)),
            "outside_box": int(np.any((x < lo) | (x > hi), axis=1).sum()),
            "rho_mean": float(rho.mean()) if w.n else 0.0,
            "above_rest_density": round(float(np.mean(rho > w.params_kw.get("rho0", 1000.0))), 3),
            "near_boundary_object": round(float(np.mean(psi > 0)), 4),
            "neighbour_list": {"mean": round(list_mean, 3), "max": list_max, "overflowed": list_over}}


def changed_keys_per_substep(sim, k: int = 5) -> float:
    """Mean keys that changed bucket over k more substeps (each read after its substep)."""
    tot = 0
    for _ in range(k):
        sim.substeps(1)
        tot += sim.changed_keys()
    return tot / k


def stage_table(stages):
    stage_ms = {name: ms for name, ms, _ in stages}
    launches = {name: la for name, _, la in stages}
    sort_ms = sum(ms for name, ms in stage_ms.items() if name.startswith("sort"))
    return stage_ms, launches, sort_ms


def mem_class_frac(n: int, stage_ms: dict, sort_ms: float):
    """The memory-bound class (sort with the reorder in its merge, cell ranges)
    against its algorithmic bytes (sort 28 + ranges 8 + reorder 52 per particle)."""
    ms = sort_ms + stage_ms.get("cell_ranges", 0.0) + stage_ms.get("reorder", 0.0)
    return (MEM_CLASS_BYTES * n / (ms * 1e-3)) / HBM_PEAK if ms > 0 else None


def bench_free_fall(steps: int) -> dict:
    """The 1M obstacle scene as configured (viscosity 0.01) from rest: the lattice
    falls freely (no pressure, no relative motion) until it reaches the floor at
    ~320 substeps, where the viscous term diverges; timed over substeps 20-220."""
    import vkfluidsim_amd as V
    from vkfluidsim_amd import configs as CF

    w = CF.obstacle_1m()
    grids = V.boundary_grids(w.shapes, w.h)
    sim = V.Sim(workload_params(V, w), grids)
    sim.set_state(w.x, w.v)
    el = time_substeps(sim, 20, steps)
    ck = changed_keys_per_substep(sim)
    out = {"workload": f"{w.name}: {w.n} particles from rest, viscosity 0.01, substeps 20-{20 + steps}",
           "particle_steps_per_s": w.n * steps / el, "ms_per_substep": 1e3 * el / steps,
           "changed_keys_per_substep": ck, "state": state_summary(sim, w)}
    sim.close()
    return out


def bench_sort_stress(warmup: int, steps: int) -> dict:
    """8,388,608 particles uniform in 4 x 2 x 2 with N(0, 0.5) velocities,
    h = 0.02, 256^3 wall map: every substep re-sorts 150,000-290,000 changed
    keys through the multi-tile radix passes.  Timed over its 200 substeps
    after the first (full-sort) one; the random placement throws a few close
    pairs apart (~0.2% of the particles end outside the box, the state stays
    finite)."""
    import vkfluidsim_amd as V
    from vkfluidsim_amd import configs as CF

    w = CF.uniform_stress()
    grids = V.boundary_grids(w.shapes, w.h)
    sim = V.Sim(workload_params(V, w), grids)
    sim.set_state(w.x, w.v)
    el = time_substeps(sim, warmup, steps)
    state = state_summary(sim, w)
    ck = changed_keys_per_substep(sim, 2)
    stage_ms, launches, sort_ms = stage_table(sim.profile(3))
    sim.close()
    return {"workload": f"{w.name}: {w.n} particles, substeps {warmup}-{warmup + steps} from the seeded state",
            "particle_steps_per_s": w.n * steps / el, "ms_per_substep": 1e3 * el / steps,
            "changed_keys_per_substep": ck,
            "stages_ms": {k: round(v, 5) for k, v in stage_ms.items()}, "stage_launches": launches,
            "memory_class_hbm_frac": mem_class_frac(w.n, stage_ms, sort_ms),
            "whole_step_hbm_frac": B_ALG * w.n * steps / el / HBM_PEAK, "state": state}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="obstacle_1m_contact")
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=48)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 --pmc child runs")
    ap.add_argument("--no-batched", action="store_true", help="skip the 256-scene batched vs looped measurement")
    ap.add_argument("--no-looped", action="store_true", help="time the batch alone, without the per-scene loops")
    ap.add_argument("--batched-scenes", type=int, default=256)
    ap.add_argument("--batched-steps", type=int, default=200)
    ap.add_argument("--looped-steps", type=int, default=20)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the forward + adjoint measurement")
    ap.add_argument("--no-lague", action="store_true", help="skip the Lague model measurement")
    ap.add_argument("--no-reference-scenes", action="store_true",
                    help="skip the reference's own dam_break.json and obstacle.json scenes")
    ap.add_argument("--no-extra-legs", action="store_true", help="skip the free-fall and sort-stress legs")
    ap.add_argument("--no-free-fall", action="store_true", help="skip the free-fall leg")
    ap.add_argument("--no-sort-stress", action="store_true", help="skip the 8M sort-stress leg")
    ap.add_argument("--lague-steps", type=int, default=300, help="timed substeps of the Lague dam break")
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--adjoint-every", type=int, default=20)
    ap.add_argument("--stress-steps", type=int, default=200)
    ap.add_argument("--traffic-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--make-state", metavar="NPZ", help="write the workload's settled state and exit")
    ap.add_argument("--state", metavar="NPZ", help="settled state of the workload written by --make-state "
                                                   "(default: made by a child of this run)")
    ap.add_argument("--summarize-stats", metavar="DB", help="print a rocprofv3 --stats database as text")
    ap.add_argument("--summarize-stats-by-grid", metavar="DB",
                    help="print a rocprofv3 --kernel-trace database per kernel and grid size")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_DB", "WRITE_DB"),
                    help="print the HBM traffic of two --pmc databases of --traffic-probe runs")
    ap.add_argument("--summarize-counters", nargs="+", metavar="DB",
                    help="print the per-wave SQ counters of --pmc databases of --traffic-probe runs")
    args = ap.parse_args()
    if args.summarize_stats_by_grid:
        print(summarize_stats_by_grid(args.summarize_stats_by_grid))
        return 0
    if args.summarize_stats:
        print(summarize_stats(args.summarize_stats))
        return 0
    if args.summarize_pmc:
        print(summarize_pmc(*args.summarize_pmc))
        return 0
    if args.summarize_counters:
        print(summarize_counters(args.summarize_counters))
        return 0

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the substep engine is a "
              "single-GPU kernel path; multi-GPU sharding is a design note, not a measured path.",
              file=sys.stderr)
        return 2
    if args.make_state:
        return make_state(args.workload, args.make_state)
    if args.traffic_probe:
        return traffic_probe(args.workload, args.state)

    import subprocess
    import tempfile

    from vkfluidsim_amd import configs as CF

    tmpdir = tempfile.TemporaryDirectory(dir="/tmp")
    state_path = args.state
    if state_path is None and CF.WORKLOADS[args.workload]().settle:
        # the settled state, made once by a child before this process touches the GPU
        state_path = str(Path(tmpdir.name) / "state.npz")
        subprocess.run([sys.executable, str(ROOT / "bench.py"), "--workload", args.workload, "--make-state",
                        state_path], check=True, timeout=600)
    traffic = {} if args.no_traffic else measure_traffic(args.workload, state_path)
    valu = {} if args.no_traffic else measure_valu(args.workload, state_path)

    import vkfluidsim_amd as V

    if V.device_count() < 1:
        print("bench.py needs a GPU", file=sys.stderr)
        return 1
    w = CF.WORKLOADS[args.workload]()
    t_grid = time.perf_counter()
    grids = V.boundary_grids(w.shapes, w.h)
    grid_ms = 1e3 * (time.perf_counter() - t_grid)
    x0, v0 = settled_state(w, grids, state_path)
    sim = V.Sim(workload_params(V, w), grids)
    sim.set_state(x0, v0)
    elapsed = time_substeps(sim, args.warmup, args.steps)
    ms_per_step = 1e3 * elapsed / args.steps
    value = w.n * args.steps / elapsed

    end_state = state_summary(sim, w)
    end_state["substeps_run"] = w.settle + args.warmup + args.steps
    list_mean = end_state["neighbour_list"]["mean"]
    x_end, _, _, _ = sim.get_state()
    cand = candidates_per_particle(x_end, w.h, w.n) if valu and w.n else None
    changed = changed_keys_per_substep(sim)
    stage_ms, launches, sort_ms = stage_table(sim.profile(args.profile_steps))
    sim.close()

    per = dict((k, v) for k, v in stage_ms.items() if not k.startswith("sort"))
    per["sort"] = sort_ms
    dom = max(per, key=per.get)
    dom_launch_ms = per[dom] / max(1, launches.get(dom, 1)) if dom != "sort" else sort_ms
    alg = ALG_BYTES.get(dom, 0) + (ALG_BYTES["reorder"] if dom == "sort" else 0)
    achieved = alg * w.n / (dom_launch_ms * 1e-3)
    sb = stage_bytes(list_mean, reorder_in_sort="reorder" not in per)
    dom_valu = valu.get(dom)

    def valu_obj(k, v):
        t = per[k] / max(1, launches.get(k, 1)) * 1e-3
        item = (cand or 1.0) if k == "density" else max(list_mean, 1.0)
        return {"valu_insts_per_wave": round(v["valu_insts_per_wave"]),
                "work_per_particle": ({"candidates": round(cand, 1)} if k == "density" and cand
                                      else {"listed_neighbours": round(list_mean, 1)}),
                "valu_insts_per_item": round(v["valu_insts_per_wave"] / item, 1),
                "lane_utilization": round(v["lane_utilization"], 3),
                "valu_busy_frac": round(v["valu_cycles_per_simd"] / (t * PEAK_CLOCK_GHZ * 1e9), 3),
                "clock_ghz_assumed": PEAK_CLOCK_GHZ}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": w.name, "n_particles": w.n, "smooth_radius": w.h, "dt": w.dt,
                   "viscosity": w.params_kw.get("viscosity", 0.01), "settling_substeps": w.settle,
                   "boundary_grids": [list(g.resolution) for g in grids],
                   "boundary_objects": [describe_shape(sh) for sh in w.shapes], "substep_graph": True},
        "setup": {"boundary_grids_ms": round(grid_ms, 1),
                  "what": "SDF + volume map of every boundary object generated on the GPU (FP64, 16^3 "
                          "Gauss-Legendre per node), host copies included"},
        # HBM roofline of the dominant kernel against its algorithmic bytes
        # (density: read x 12 + write rho 4 per particle); the bytes this
        # engine must move (its neighbour lists included) and the measured
        # traffic beside them
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": traffic.get(dom) if dom in traffic else None,
                     "algorithmic_bytes": alg * w.n, "bytes_per_particle": alg, "launch_ms": dom_launch_ms,
                     "engine_bytes_per_particle": round(sb.get(dom, 0.0), 2),
                     "engine_bytes_frac": sb.get(dom, 0.0) * w.n / (dom_launch_ms * 1e-3) / HBM_PEAK},
        # the neighbour kernels are bound by VALU issue: wave64 VALU
        # instructions per second against the peak issue rate (4 cycles per
        # instruction on each of the 1,024 SIMDs at the peak clock)
        "valu_roofline": ({"bound": "valu", "kernel": dom,
                           "achieved": dom_valu["valu_insts"] / (dom_launch_ms * 1e-3), "peak": VALU_PEAK,
                           "unit": "wave64 VALU instructions/s",
                           "frac": dom_valu["valu_insts"] / (dom_launch_ms * 1e-3) / VALU_PEAK}
                          if dom_valu else None),
        "valu_issue": {k: valu_obj(k, v) for k, v in valu.items() if k in per} or None,
        "whole_step_hbm_frac": B_ALG * value / HBM_PEAK,
        "memory_class_hbm_frac": mem_class_frac(w.n, stage_ms, sort_ms),
        "changed_keys_per_substep": changed,
        "traffic_per_launch": {k: round(v) for k, v in traffic.items() if k in STAGE_KERNELS} or traffic or None,
        "stages_ms": {k: round(v, 5) for k, v in stage_ms.items()},
        "stage_launches": launches,
        "event_sum_ms_per_step": sum(stage_ms.values()),
        "end_state": end_state,
    }
    out["free_fall"] = None if args.no_extra_legs or args.no_free_fall else bench_free_fall(200)
    out["sort_stress"] = (None if args.no_extra_legs or args.no_sort_stress
                          else bench_sort_stress(1, args.stress_steps))
    out["batched"] = None if args.no_batched else bench_batched(args.batched_scenes, args.batched_steps,
                                                                  0 if args.no_looped else args.looped_steps)
    out["adjoint"] = None if args.no_adjoint else bench_adjoint(args.adjoint_steps, args.adjoint_every)
    out["lague"] = None if args.no_lague else bench_lague(args.lague_steps, 100, not args.no_cpu_baseline)
    out["reference_scenes"] = None if args.no_reference_scenes else bench_reference_scenes(100)
    out["cpu_baseline"] = None if args.no_cpu_baseline else cpu_baseline(w, grids, x0, v0, args.cpu_steps)
    tmpdir.cleanup()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
