"""The FP64 host adjoint (oracle) against central finite differences of the
FP64 forward: the adjoint's algorithm, which the GPU kernels restate, is
pinned by this check (4,096-particle dam break with box walls, 20 substeps)."""
import numpy as np
import pytest

import harness as H
import vkfluidsim_amd as V
from vkfluidsim_amd import configs as CF

STEPS = 20


def _case():
    w = CF.dam_break(side=(16, 16, 16), map_res=32)
    grids = [H.oracle_volume_map(k, sp, r, w.h) for k, sp, r in w.shapes]
    p = V.make_params(w.n, w.h, dt=w.dt, box_size=w.box_size, box_pos=w.box_pos)
    dims = (32, 32, 32)
    spacing = tuple(b / (d - 1) for b, d in zip(w.box_size, dims))
    x0 = w.x.astype(np.float64)
    v_t = 0.01 * CF.normal(CF.SEED_BASE + 4, 3 * w.n).reshape(w.n, 3)
    _, target, _, _, _ = H.oracle_adjoint(p, grids, x0, v_t, STEPS, (0, 0, 0), spacing, dims)
    return w, grids, p, dims, spacing, x0, target


def test_host_adjoint_matches_fp64_finite_differences():
    w, grids, p, dims, spacing, x0, target = _case()
    v0 = np.zeros_like(x0)
    L, _, gx, gv, gnu = H.oracle_adjoint(p, grids, x0, v0, STEPS, (0, 0, 0), spacing, dims, target)
    assert L > 0

    def loss(x, v, nu=None):
        q = p if nu is None else V.make_params(w.n, w.h, dt=w.dt, box_size=w.box_size, box_pos=w.box_pos,
                                               viscosity=nu)
        return H.oracle_adjoint(q, grids, x, v, STEPS, (0, 0, 0), spacing, dims, target)[0]

    # the largest components (particles next to the walls) and an interior one
    picks = np.argsort(-np.abs(gx).max(axis=1))[:2].tolist() + [2000]
    for i in picks:
        for a in range(3):
            e = np.zeros_like(x0)
            e[i, a] = 1e-7
            fd = (loss(x0 + e, v0) - loss(x0 - e, v0)) / 2e-7
            assert gx[i, a] == pytest.approx(fd, rel=1e-3, abs=1e-6 * np.abs(gx).max()), (i, a)
    dv = np.random.default_rng(7).standard_normal(v0.shape)
    fd = (loss(x0, v0 + 1e-6 * dv) - loss(x0, v0 - 1e-6 * dv)) / 2e-6
    assert float(np.sum(gv * dv)) == pytest.approx(fd, rel=1e-5)
    # the viscosity parameter is a float: difference over the float values taken
    nu = float(p.viscosity_strenght)
    hi, lo = float(np.float32(nu + 1e-7)), float(np.float32(nu - 1e-7))
    fd = (loss(x0, v0, hi) - loss(x0, v0, lo)) / (hi - lo)
    assert gnu == pytest.approx(fd, rel=1e-3)


def test_host_adjoint_matches_fp64_finite_differences_200_substeps():
    """The 4,096-particle variant over 200 substeps: directional derivative
    along random initial velocities and positions against FP64 central
    differences."""
    w = CF.dam_break(side=(16, 16, 16), map_res=32)
    grids = [H.oracle_volume_map(k, sp, r, w.h) for k, sp, r in w.shapes]
    p = V.make_params(w.n, w.h, dt=w.dt, box_size=w.box_size, box_pos=w.box_pos)
    dims = (32, 32, 32)
    spacing = tuple(b / (d - 1) for b, d in zip(w.box_size, dims))
    x0 = w.x.astype(np.float64)
    v_t = 0.01 * CF.normal(CF.SEED_BASE + 4, 3 * w.n).reshape(w.n, 3)
    steps = 200
    _, target, _, _, _ = H.oracle_adjoint(p, grids, x0, v_t, steps, (0, 0, 0), spacing, dims)
    v0 = np.zeros_like(x0)
    L, _, gx, gv, gnu = H.oracle_adjoint(p, grids, x0, v0, steps, (0, 0, 0), spacing, dims, target)

    def loss(x, v):
        return H.oracle_adjoint(p, grids, x, v, steps, (0, 0, 0), spacing, dims, target)[0]

    rng = np.random.default_rng(7)
    dv = rng.standard_normal(v0.shape)
    fd = (loss(x0, v0 + 1e-6 * dv) - loss(x0, v0 - 1e-6 * dv)) / 2e-6
    assert float(np.sum(gv * dv)) == pytest.approx(fd, rel=1e-5)
    dx = rng.standard_normal(x0.shape)
    fd = (loss(x0 + 1e-8 * dx, v0) - loss(x0 - 1e-8 * dx, v0)) / 2e-8
    assert float(np.sum(gx * dx)) == pytest.approx(fd, rel=1e-3)


def test_fp64_adjoint_of_fp32_trajectory_close_to_fp64_adjoint():
    """The widened-tape variant runs the same FP64 adjoint on the FP32 oracle's
    trajectory; over a short run the two trajectories, and so the gradients,
    stay within FP32 drift."""
    w = CF.dam_break(side=(12, 12, 12), map_res=24)
    grids = [H.oracle_volume_map(k, sp, r, w.h) for k, sp, r in w.shapes]
    p = V.make_params(w.n, w.h, dt=w.dt, box_size=w.box_size, box_pos=w.box_pos)
    dims = (24, 24, 24)
    spacing = tuple(b / (d - 1) for b, d in zip(w.box_size, dims))
    target = np.zeros(int(np.prod(dims)))
    L, gx, gv, gnu = H.oracle_adjoint_f32_trajectory(p, grids, w.x, w.v, 10, (0.0, 0.0, 0.0), spacing, dims, target)
    L64, _, hx, hv, hnu = H.oracle_adjoint(p, grids, w.x.astype(np.float64), w.v.astype(np.float64), 10,
                                           (0.0, 0.0, 0.0), spacing, dims, target)
    assert L == pytest.approx(L64, rel=1e-5)
    assert np.linalg.norm(gv - hv) / np.linalg.norm(hv) < 1e-2
    assert np.linalg.norm(gx - hx) / np.linalg.norm(hx) < 1e-2
    assert gnu == pytest.approx(hnu, rel=1e-2)
