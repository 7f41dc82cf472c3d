"""GPU tests of the differentiable substeps (vfs_adjoint_*): forward
consistency, and gradients against central finite differences of the GPU
forward and of the FP64 CPU restatement (directional derivatives)."""
import numpy as np
import pytest
from scipy.spatial import cKDTree

import harness as H
import vkfluidsim_amd as V
from vkfluidsim_amd import configs as CF

pytestmark = pytest.mark.gpu

EPS = 1e-3  # central finite differences


def _setup(side=(16, 16, 16), steps=20, every=5, dims=(32, 32, 32)):
    w = CF.dam_break(side=side, map_res=32)
    grids = [V.volume_map(k, sp, r, w.h) for k, sp, r in w.shapes]
    p = V.make_params(w.n, w.h, dt=w.dt, box_size=w.box_size, box_pos=w.box_pos)
    sim = V.Sim(p, grids)
    spacing = tuple(b / (d - 1) for b, d in zip(w.box_size, dims))
    adj = V.Adjoint(sim, steps, every, (0.0, 0.0, 0.0), spacing, dims)
    # target: the density field of a run whose initial velocities are perturbed by N(0, 0.01)
    v_t = (0.01 * CF.normal(CF.SEED_BASE + 4, 3 * w.n)).astype(np.float32).reshape(w.n, 3)
    _, target = adj.forward(w.x, v_t)
    return w, grids, p, sim, adj, target, spacing, dims


def _grid_points(spacing, dims):
    """Grid points in index order i + nx (j + ny k), as float32 products."""
    k, j, i = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]), indexing="ij")
    idx = np.stack([i.reshape(-1), j.reshape(-1), k.reshape(-1)], 1)
    return (idx.astype(np.float32) * np.asarray(spacing, np.float32)).astype(np.float64)


def _grid_density_f64(x, h, mass, spacing, dims):
    """rho(g) = sum_j m W(|g - x_j|) in float64 (cubic spline, sigma = 8/(pi h^3))."""
    pts = _grid_points(spacing, dims)
    x = np.asarray(x, np.float64)
    tree = cKDTree(x)
    rho = np.zeros(len(pts))
    sigma = 8.0 / (np.pi * h ** 3)
    for k, nb in enumerate(tree.query_ball_point(pts, h)):
        if not nb:
            continue
        q = np.linalg.norm(x[nb] - pts[k], axis=1) / h
        wq = np.where(q <= 0.5, sigma * (6 * q ** 3 - 6 * q ** 2 + 1), 2 * sigma * (1 - q) ** 3)
        rho[k] = mass * np.sum(np.where(q <= 1.0, wq, 0.0))
    return rho


def _mass(h):
    h = np.float32(h)
    return float(np.float32(1000.0) * ((np.float32(0.8) * (h * h) * h) / np.float32(8.0)))


def test_adjoint_forward_matches_plain_substeps():
    w, grids, p, sim, adj, target, spacing, dims = _setup()
    loss, rho = adj.forward(w.x, w.v, target)
    x_a, v_a, _, _ = sim.get_state()
    ref = V.Sim(p, grids)
    ref.set_state(w.x, w.v)
    ref.substeps(20)
    x_r, v_r, _, _ = ref.get_state()
    np.testing.assert_array_equal(x_a, x_r)
    np.testing.assert_array_equal(v_a, v_r)
    want = _grid_density_f64(x_r, float(np.float32(w.h)), _mass(w.h), spacing, dims)
    # the GPU evaluates the kernel in double: only the float copy of rho rounds
    assert np.abs(rho - want).max() <= 1e-7 * np.abs(want).max()
    # the loss is summed from the double-precision field; rho is its float copy
    assert loss == pytest.approx(np.mean((rho.astype(np.float64) - target) ** 2), rel=1e-4)
    assert loss > 0


def test_backward_profile_times_every_stage_and_changes_nothing():
    """vfs_adjoint_profile_backward runs the reverse pass once more with
    events around every launch group: the reverse kernels run once per
    substep (20), the recomputed forward kernels once per substep, the list
    check once per checkpoint interval (4), and the gradients of a reverse pass
    after it are those before it, bit for bit."""
    w, grids, p, sim, adj, target, spacing, dims = _setup()
    adj.forward(w.x, w.v, target)
    gx0, gv0, gnu0 = adj.backward()
    stages = {name: (ms, k) for name, ms, k in adj.profile_backward()}
    print(stages)
    for name in ("adj_integrate", "adj_forces", "adj_viscosity_sum", "adj_boundary", "adj_density", "density",
                 "forces_integrate"):
        assert stages[name][1] == 20, name
        assert stages[name][0] > 0.0, name
    assert stages["list_check"][1] == 4
    assert stages["loss_gradient"][1] == 1
    gx1, gv1, gnu1 = adj.backward()
    np.testing.assert_array_equal(gx0, gx1)
    np.testing.assert_array_equal(gv0, gv1)
    assert gnu0 == gnu1


def _directional(adj, x0, v0, target, dx, dv):
    lp, _ = adj.forward(x0 + EPS * dx, v0 + EPS * dv, target)
    lm, _ = adj.forward(x0 - EPS * dx, v0 - EPS * dv, target)
    return (lp - lm) / (2 * EPS)


def test_adjoint_velocity_gradient_matches_gpu_finite_differences():
    """Central differences (eps = 1e-3) of the GPU forward along a random
    direction of the initial velocities, 200 substeps.  The difference
    quotient of the FP32 forward carries the rounding of its trajectory: it
    sits 0.3% from the adjoint of that trajectory (measured: -9.406e-2 against
    -9.378e-2; 0.3-0.5% for eps 1e-3 ... 1e-2).  The adjoint itself equals the
    FP64 adjoint of the same trajectory to 1e-6
    (test_adjoint_matches_fp64_adjoint_of_the_same_trajectory), whose algorithm
    matches FP64 central differences to 1e-5 (test_adjoint_cpu.py)."""
    w, grids, p, sim, adj, target, spacing, dims = _setup(steps=200, every=20)
    x0, v0 = w.x.astype(np.float32), w.v.astype(np.float32)
    adj.forward(x0, v0, target)
    gx, gv, gnu = adj.backward()
    assert np.isfinite(gx).all() and np.isfinite(gv).all() and np.isfinite(gnu)
    dv = np.random.default_rng(7).standard_normal(v0.shape).astype(np.float32)
    fd = _directional(adj, x0, v0, target, np.zeros_like(x0), dv)
    an = float(np.sum(gv.astype(np.float64) * dv))
    print(f"d loss/d v0 along a random direction: adjoint {an:.6e}, GPU central difference {fd:.6e}")
    assert an == pytest.approx(fd, rel=1e-2)


@pytest.mark.parametrize("steps", [20, 200])
def test_adjoint_matches_fp64_host_adjoint(steps):
    """The GPU adjoint (FP32 trajectory) against the FP64 host adjoint of the
    oracle, 4,096 particles, checkpoints every 20 substeps."""
    w, grids, p, sim, adj, target, spacing, dims = _setup(steps=steps, every=min(steps, 20))
    adj.forward(w.x, w.v, target)
    gx, gv, gnu = adj.backward()
    L64, _, hx, hv, hnu = H.oracle_adjoint(p, grids, w.x.astype(np.float64), w.v.astype(np.float64), steps,
                                           (0.0, 0.0, 0.0), spacing, dims, target)
    ex = np.linalg.norm(gx - hx) / np.linalg.norm(hx)
    ev = np.linalg.norm(gv - hv) / np.linalg.norm(hv)
    print(f"{steps} substeps: |g_gpu - g_fp64| / |g_fp64|: x0 {ex:.2e}, v0 {ev:.2e}; "
          f"viscosity {gnu:.5e} vs {hnu:.5e}")
    assert ex < 3e-2 and ev < 2e-2
    assert np.sign(gnu) == np.sign(hnu) and abs(gnu - hnu) <= 0.3 * abs(hnu)


@pytest.mark.parametrize("steps", [20, 200])
def test_adjoint_matches_fp64_adjoint_of_the_same_trajectory(steps):
    """The GPU adjoint against the oracle's FP64 adjoint run on the FP32
    trajectory, which the GPU reproduces bit for bit: the adjoint arithmetic
    alone, without the divergence of FP32 and FP64 trajectories."""
    w, grids, p, sim, adj, target, spacing, dims = _setup(steps=steps, every=min(steps, 20))
    loss, _ = adj.forward(w.x, w.v, target)
    gx, gv, gnu = adj.backward()
    L, mx, mv, mnu = H.oracle_adjoint_f32_trajectory(p, grids, w.x, w.v, steps, (0.0, 0.0, 0.0), spacing, dims,
                                                     target.astype(np.float64))
    ex = np.linalg.norm(gx - mx) / np.linalg.norm(mx)
    ev = np.linalg.norm(gv - mv) / np.linalg.norm(mv)
    print(f"{steps} substeps, same trajectory: x0 {ex:.2e}, v0 {ev:.2e}, viscosity {gnu:.8e} vs {mnu:.8e}")
    # both evaluate the loss and the reverse substeps in double on the same
    # float tape (measured: 2.5e-8, the rounding of the float gradients)
    assert loss == pytest.approx(L, rel=1e-9)
    assert ex < 1e-6 and ev < 1e-6
    assert gnu == pytest.approx(mnu, rel=1e-6)


def test_torch_autograd_function_wrapper():
    """tests/torch_ext (a torch::autograd::Function over vfs_adjoint_*) gives
    the loss and gradients of the C ABI path.  Runs in a child process that
    initialises torch's HIP runtime before libvfs.so is loaded, so that both
    share one runtime."""
    import subprocess
    import sys
    from pathlib import Path

    script = Path(__file__).resolve().parent / "torch_ext" / "check_autograd.py"
    r = subprocess.run([sys.executable, str(script)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stderr[-4000:]


def test_backward_rejects_lists_past_the_capacity():
    """The reverse pass differentiates complete neighbour lists only: when a
    recomputed substep's list outgrows VFS_NEIGHBOUR_CAPACITY (particles at
    0.3 h spacing, ~150 neighbours each) the backward stops with a state error
    naming the capacity instead of returning gradients of truncated lists.
    The check is the maximum list length over every particle of every
    substep of a segment, so it also pins that reduction."""
    h = 0.05
    x = CF.lattice((10, 10, 10), 0.3 * h, (0.3, 0.3, 0.3), 0.03 * h, 5)
    n = x.shape[0]
    p = V.make_params(n, h, dt=1e-5, box_size=(1.0, 1.0, 1.0))
    sim = V.Sim(p, [])
    adj = V.Adjoint(sim, 4, 2, (0.0, 0.0, 0.0), (0.1, 0.1, 0.1), (8, 8, 8))
    adj.forward(x, np.zeros_like(x))
    _, mx, over = sim.neighbour_stats()
    assert mx > V.VFS_NEIGHBOUR_CAPACITY and over > 0
    with pytest.raises(V.VfsError, match="VFS_NEIGHBOUR_CAPACITY"):
        adj.backward()
    adj.close()
    sim.close()


def test_backward_accepts_lists_at_the_capacity_edge():
    """The same check passes when the longest list fits: a lattice at 0.5 h
    spacing (~33 neighbours) runs the backward to the end."""
    h = 0.05
    x = CF.lattice((12, 12, 12), 0.5 * h, (0.3, 0.3, 0.3), 0.03 * h, 6)
    n = x.shape[0]
    p = V.make_params(n, h, dt=1e-5, box_size=(1.0, 1.0, 1.0))
    sim = V.Sim(p, [])
    adj = V.Adjoint(sim, 4, 2, (0.0, 0.0, 0.0), (0.1, 0.1, 0.1), (8, 8, 8))
    adj.forward(x, np.zeros_like(x))
    _, mx, _ = sim.neighbour_stats()
    assert 0 < mx <= V.VFS_NEIGHBOUR_CAPACITY
    gx, gv, gnu = adj.backward()
    assert np.all(np.isfinite(gx)) and np.all(np.isfinite(gv)) and np.isfinite(gnu)
    adj.close()
    sim.close()


@pytest.mark.parametrize("seed", range(6))
def test_adjoint_random_scenes_match_fp64_adjoint_of_the_same_trajectory(seed):
    """Seeded random scenes for the reverse pass: a jittered block of 300-4,000
    particles at about half the smoothing radius apart, random velocities,
    box (offset into negative coordinates), gravity, stiffness, exponent and
    viscosity, box walls and sometimes a translated sphere, 8-30 substeps with
    a random checkpoint interval and grid.  The GPU gradients equal the FP64
    adjoint of the same FP32 trajectory to 1e-6, as in the fixed case."""
    r = np.random.default_rng(4400 + seed)
    h = float(r.uniform(0.03, 0.06))
    box = r.uniform(0.5, 1.2, 3)
    box_pos = r.uniform(-0.5, 0.2, 3)
    side = tuple(int(v) for v in r.integers(6, 16, 3))
    spacing_p = h / r.uniform(1.9, 2.2)
    x = CF.lattice(side, spacing_p, box_pos + 2 * spacing_p, 0.05 * spacing_p, 500 + seed)
    n = x.shape[0]
    v = (0.2 * r.standard_normal((n, 3))).astype(np.float32)
    g = r.standard_normal(3)
    steps = int(r.integers(8, 31))
    every = int(r.integers(1, steps + 1))
    p = V.make_params(n, h, dt=float(r.uniform(0.5, 2.0) * 1e-4), box_size=box, box_pos=box_pos,
                      stiffness=float(r.uniform(200.0, 2000.0)), gamma=float(r.choice([1.0, 7.0])),
                      viscosity=float(10.0 ** r.uniform(-3.0, -1.0)), gravity=9.81 * g / np.linalg.norm(g))
    c = box_pos + 0.5 * box
    grids = [V.volume_map(V.VFS_SHAPE_BOX_WALLS, (*c, *(0.5 * box)), (int(r.integers(16, 33)),) * 3, h)]
    if r.random() < 0.5:
        sphere = V.volume_map(V.VFS_SHAPE_SPHERE, (0.0, 0.0, 0.0, float(r.uniform(0.05, 0.15))), (20, 20, 20), h)
        t = np.eye(4, dtype=np.float32)
        t[3, :3] = -(box_pos + r.uniform(0.3, 0.7, 3) * box).astype(np.float32)
        sphere.transform = t.reshape(-1)
        grids.append(sphere)
    dims = tuple(int(d) for d in r.integers(8, 24, 3))
    spacing = tuple(float(b) / (d - 1) for b, d in zip(box, dims))
    origin = tuple(float(o) for o in box_pos)
    print(f"case adjoint random {seed}: n={n} steps={steps} every={every} objects={len(grids)}", flush=True)
    sim = V.Sim(p, grids)
    adj = V.Adjoint(sim, steps, every, origin, spacing, dims)
    _, target = adj.forward(x, v + (0.01 * r.standard_normal((n, 3))).astype(np.float32))
    loss, _ = adj.forward(x, v, target)
    gx, gv, gnu = adj.backward()
    L, mx, mv, mnu = H.oracle_adjoint_f32_trajectory(p, grids, x, v, steps, origin, spacing, dims,
                                                     target.astype(np.float64))
    ex = np.linalg.norm(gx - mx) / np.linalg.norm(mx)
    ev = np.linalg.norm(gv - mv) / np.linalg.norm(mv)
    print(f"  x0 {ex:.2e}, v0 {ev:.2e}, viscosity {gnu:.8e} vs {mnu:.8e}")
    assert loss == pytest.approx(L, rel=1e-9)
    assert ex < 1e-6 and ev < 1e-6
    assert gnu == pytest.approx(mnu, rel=1e-6, abs=1e-12 * max(1.0, abs(mnu)))
    adj.close()
    sim.close()
