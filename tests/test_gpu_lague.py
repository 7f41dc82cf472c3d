"""LagueModel on the GPU (vfs_lague_*) against the oracle's FP32 restatement:
hash tables, positions, velocities and densities bit for bit after 1 and 60
substeps of a dam break that hits the walls, graph against eager launches,
parameter changes, and FP32 drift against the FP64 twin."""
import numpy as np
import pytest

import harness as H
import vkfluidsim_amd as V
from vkfluidsim_amd import configs as CF

pytestmark = pytest.mark.gpu


def _params(w, flags=V.VFS_FLAG_TRACK_IDS):
    return V.make_params(w.n, w.h, dt=w.dt, box_size=w.box_size, box_pos=w.box_pos, rho0=w.params_kw["rho0"],
                         iterations=w.params_kw["iterations"], flags=flags)


def _small():
    # the reference block shape at 1/5 scale, in a box it reaches within the run
    w = CF.lague_dam_break(counts=(10, 12, 10))
    w.box_size = (3.0, 3.0, 3.0)
    w.x = w.x - np.array([0.0, 0.0, w.x[:, 2].min() - 0.5], np.float32)
    return w


def _pair(w, lague=None, flags=V.VFS_FLAG_TRACK_IDS):
    p = _params(w, flags)
    lp = lague or V.make_lague_params()
    sim = V.Lague(p, lp)
    sim.set_state(w.x, w.v)
    orc = H.OracleLague(p, lp)
    orc.set_state(w.x, w.v)
    return sim, orc


def _same(sim, orc):
    x, v, rho, pid = sim.get_state(ids=True)
    ox, ov, orho, _, opid = orc.get_state()
    np.testing.assert_array_equal(pid, opid)
    np.testing.assert_array_equal(x, ox.astype(np.float32))
    np.testing.assert_array_equal(v, ov.astype(np.float32))
    np.testing.assert_array_equal(rho, orho.astype(np.float32))


@pytest.mark.parametrize("steps", [1, 60])
def test_lague_bit_exact(steps):
    w = _small()
    print(f"case lague {w.n} particles, {steps} substeps", flush=True)
    sim, orc = _pair(w)
    sim.substeps(steps)
    orc.substeps(steps)
    _same(sim, orc)
    k, p, o = sim.get_hash()
    assert np.all(np.diff(k.astype(np.int64)) >= 0)
    x, v, _, _ = sim.get_state()
    if steps > 1:
        # the block spreads over the floor and touches the walls
        assert x[:, 1].min() == 0.0 and np.abs(v).max() > 0.1


def test_lague_graph_and_eager_agree():
    w = _small()
    a = V.Lague(_params(w, V.VFS_FLAG_NO_GRAPH))
    b = V.Lague(_params(w, 0))
    for s in (a, b):
        s.set_state(w.x, w.v)
        s.substeps(25)
    for fa, fb in zip(a.get_state()[:3], b.get_state()[:3]):
        np.testing.assert_array_equal(fa, fb)


def test_lague_params_change():
    w = _small()
    sim, orc = _pair(w)
    sim.substeps(10)
    orc.substeps(10)
    lp = V.make_lague_params(wall_damping_factor=0.5, pressure_multiplier=150.0, near_pressure_multiplier=4.0,
                             viscosity=0.1)
    sim.set_lague_params(lp)
    got = sim.get_lague_params()
    assert (got.pressure_multiplier, got.viscosity_strenght) == (150.0, pytest.approx(0.1))
    # the oracle continues from the same state with the new parameters
    x, v, _, _ = sim.get_state()
    p = _params(w)
    o2 = H.OracleLague(p, lp)
    o2.set_state(x, v)
    sim2 = V.Lague(p, lp)
    sim2.set_state(x, v)
    sim.substeps(10)
    sim2.substeps(10)
    o2.substeps(10)
    for fa, fb in zip(sim.get_state()[:3], sim2.get_state()[:3]):
        np.testing.assert_array_equal(fa, fb)
    _same(sim2, o2)


def test_lague_fp64_drift():
    w = _small()
    p = _params(w)
    sim = V.Lague(p)
    sim.set_state(w.x, w.v)
    o64 = H.OracleLague(p, V.make_lague_params(), fp64=True)
    o64.set_state(w.x, w.v)
    sim.substeps(30)
    o64.substeps(30)
    x, v, rho, pid = sim.get_state(ids=True)
    ox, ov, orho, _, opid = o64.get_state()
    x, v, rho = x[np.argsort(pid)], v[np.argsort(pid)], rho[np.argsort(pid)]
    ox, ov, orho = ox[np.argsort(opid)], ov[np.argsort(opid)], orho[np.argsort(opid)]
    ex, ev, er = H.rel_err(x, ox), H.rel_err(v, ov), H.rel_err(rho, orho)
    print(f"Lague FP32 vs FP64 after 30 substeps: x {ex:.2e}, v {ev:.2e}, rho {er:.2e}")
    assert ex < 1e-4 and er < 1e-3 and ev < 1e-2


def test_lague_profile_stages():
    w = _small()
    sim = V.Lague(_params(w))
    sim.set_state(w.x, w.v)
    names = [s[0] for s in sim.profile(3)]
    assert names == ["predict", "sort_prepare", "sort_scan", "sort_histogram", "sort_small", "sort_pass",
                     "sort_merge", "reorder", "cell_ranges", "density", "forces_integrate"]
    with pytest.raises(V.VfsError, match="no boundary objects"):
        sim.get_boundary()


@pytest.mark.parametrize("part", ["all", "half"])
def test_lague_neighbour_list_overflow_bit_exact(part):
    """Both passes walk the density pass's neighbour list; a particle with more
    neighbours than the list holds walks the cells again, with the same
    results.  A compressed block overflows every list ("all") or the lists of
    its compressed half only ("half")."""
    w = _small()
    c = w.x.mean(axis=0)
    sel = np.ones(w.n, bool) if part == "all" else (w.x[:, 0] < c[0])
    w.x = w.x.copy()
    w.x[sel] = c + (w.x[sel] - c) * np.float32(0.45)
    print(f"case lague overflow {part}: {w.n} particles", flush=True)
    sim, orc = _pair(w)
    sim.substeps(1)
    orc.substeps(1)
    _same(sim, orc)
    mean, mx, over = sim.neighbour_stats()
    assert mx > 64 and over > 0
    if part == "half":
        assert over < w.n
    sim.substeps(4)
    orc.substeps(4)
    _same(sim, orc)


def test_cpp_dam_break_scene_matches_the_c_abi(tmp_path):
    """tests/cpp/facade_lague (DamBreakScene with LagueModel in the C++ facade,
    dam_break_scene.cpp) gives the same state as the reference-sized Lague dam
    break through ctypes (configs.lague_dam_break), bit for bit."""
    import subprocess
    from pathlib import Path
    exe = Path(__file__).resolve().parent / "cpp" / "facade_lague"
    out = tmp_path / "lague.bin"
    print("case facade DamBreakScene: 150,000 particles, 4 frames", flush=True)
    r = subprocess.run([str(exe), "4", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = out.read_bytes()
    n = int(np.frombuffer(raw[:4], np.uint32)[0])
    f = np.frombuffer(raw[4:], np.float32)
    x, v, rho = f[:3 * n].reshape(n, 3), f[3 * n:6 * n].reshape(n, 3), f[6 * n:7 * n]
    w = CF.lague_dam_break()
    assert n == w.n
    sim = V.Lague(_params(w, flags=0))
    sim.set_state(w.x, w.v)
    for _ in range(4):
        sim.step()
    ex, ev, er, _ = sim.get_state()
    np.testing.assert_array_equal(x, ex)
    np.testing.assert_array_equal(v, ev)
    np.testing.assert_array_equal(rho, er)
    sim.close()


@pytest.mark.parametrize("seed", range(12))
def test_lague_random_scenes_bit_exact(seed):
    """Seeded random Lague scenes: ragged particle counts (30-20,000), a
    jittered block plus a cloud reaching past the box (offset into negative
    coordinates), random velocities and gravity, smoothing radius, rest
    density, time step and the four model parameters; 20 substeps, bit-exact
    against the oracle after every 5."""
    r = np.random.default_rng(9100 + seed)
    n = int(np.exp(r.uniform(np.log(30.0), np.log(20000.0))))
    h = float(r.uniform(0.1, 0.4))
    box = r.uniform(2.0, 8.0, 3)
    box_pos = r.uniform(-3.0, 0.5, 3)
    rho0 = float(r.uniform(300.0, 1000.0))
    spacing = h / r.uniform(1.5, 2.5)
    n_block = int(n * r.uniform(0.5, 0.9))
    side = int(np.ceil(n_block ** (1.0 / 3.0)))
    block = CF.lattice((side, side, side), spacing, box_pos + spacing, 0.1 * spacing, 300 + seed)[:n_block]
    cloud = (box_pos + r.uniform(-0.1, 1.1, (n - n_block, 3)) * box).astype(np.float32)
    x = np.concatenate([block, cloud]).astype(np.float32)
    v = r.standard_normal((n, 3)).astype(np.float32)
    g = r.standard_normal(3)
    p = V.make_params(n, h, dt=float(r.uniform(0.002, 0.01)), box_size=box, box_pos=box_pos, rho0=rho0,
                      gravity=9.81 * g / np.linalg.norm(g), iterations=int(r.integers(1, 4)),
                      flags=V.VFS_FLAG_TRACK_IDS)
    lp = V.make_lague_params(wall_damping_factor=float(r.uniform(0.5, 1.0)),
                             pressure_multiplier=float(r.uniform(50.0, 500.0)),
                             near_pressure_multiplier=float(r.uniform(0.5, 5.0)),
                             viscosity=float(r.uniform(0.0, 0.1)))
    print(f"case lague random {seed}: n={n} h={h:.3f} rho0={rho0:.0f}", flush=True)
    sim = V.Lague(p, lp)
    sim.set_state(x, v)
    orc = H.OracleLague(p, lp)
    orc.set_state(x, v)
    for _ in range(4):
        sim.substeps(5)
        orc.substeps(5)
        _same(sim, orc)
