"""The oracle's LagueModel restatement pinned on the CPU: kernel
normalisations and derivatives, free fall and wall reflections in closed form,
densities against brute force, momentum of an isolated pair, FP32 against the
FP64 twin."""
import numpy as np
import pytest

import harness as H
import vkfluidsim_amd as V
from vkfluidsim_amd import configs as CF


def _params(n, h=0.2, rho0=630.0, box=(23.0, 10.0, 10.0), gravity=(0.0, -9.81, 0.0), dt=1.0 / 120.0, iters=3):
    return V.make_params(n, h, dt=dt, box_size=box, rho0=rho0, gravity=gravity, iterations=iters)


def test_kernels_normalised_and_derivatives():
    h = 0.2
    o = H.OracleLague(_params(1, h=h), V.make_lague_params())
    d = np.linspace(0.0, h, 4001)
    k = np.array([o.kernels(x) for x in d])
    w = 4 * np.pi * d ** 2
    for col in (0, 1, 4):  # spiky2, spiky3, poly6 integrate to one over the ball
        assert np.trapezoid(k[:, col] * w, d) == pytest.approx(1.0, rel=2e-5)
    # gradients are the radial derivatives (central differences inside the support)
    mid = slice(200, 3800)
    for f, g in ((0, 2), (1, 3)):
        fd = np.gradient(k[:, f], d)
        np.testing.assert_allclose(k[mid, g], fd[mid], rtol=2e-3, atol=1e-3 * np.abs(k[:, g]).max())
    assert (o.kernels(h * 1.0001) == 0).all()
    assert o.kernels(h)[2] == 0.0 and o.kernels(h)[0] == 0.0  # gradients include d == h, kernels do not


def test_free_fall_and_wall_reflection_closed_form():
    """Particles farther apart than h feel only gravity (their own density terms
    cancel in the pressure loop); the floor reflects with the damping factor."""
    dt = 1.0 / 120.0
    x0 = np.array([[1.0, 5.0, 1.0], [3.0, 0.01, 3.0], [5.0, 9.0, 5.0]], np.float32)
    v0 = np.array([[0.0, 0.0, 0.0], [0.0, -2.0, 0.0], [0.5, 0.0, -1.0]], np.float32)
    p = _params(3, dt=dt, iters=1)
    o = H.OracleLague(p, V.make_lague_params(wall_damping_factor=0.5), fp64=True)
    o.set_state(x0, v0)
    o.substeps(1)
    x, v, rho, near, pid = o.get_state()
    x, v = x[np.argsort(pid)], v[np.argsort(pid)]
    g = np.array([0.0, -9.81, 0.0])
    v1 = v0.astype(np.float64) + g * np.float32(dt)
    x1 = x0.astype(np.float64) + v1 * np.float32(dt)
    # particle 1 crosses the floor y = 0: clamped, velocity reversed and damped
    x1[1, 1], v1[1, 1] = 0.0, -v1[1, 1] * 0.5
    np.testing.assert_allclose(x, x1, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(v, v1, rtol=1e-6, atol=1e-7)


def test_densities_match_brute_force():
    rng = np.random.default_rng(2)
    n = 600
    x0 = rng.uniform([1, 1, 1], [2.2, 2.2, 2.2], size=(n, 3)).astype(np.float32)
    v0 = rng.normal(0, 0.3, size=(n, 3)).astype(np.float32)
    p = _params(n)
    o = H.OracleLague(p, V.make_lague_params())
    o.set_state(x0, v0)
    o.substeps(1)
    _, _, rho, near, pid = o.get_state()
    # predicted positions x + (v + g dt) / 120 in float32, then brute force in float64
    dt = np.float32(p.time_scale * p.fixed_dt / p.iterations)
    vv = v0 + np.array([0, -9.81, 0], np.float32) * dt
    pred = (x0 + vv * np.float32(1.0 / 120.0)).astype(np.float64)
    h = 0.2
    d = np.linalg.norm(pred[:, None, :] - pred[None, :, :], axis=2)
    inside = d <= h
    # the stencil walks a bucket once per stencil cell whose key it holds, so
    # a pair counts as often as j's key appears among i's 27 stencil keys
    cells = np.floor(pred.astype(np.float32) / np.float32(h)).astype(np.int64)
    key = H.hash_cells(cells) % n
    offs = np.array([(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)])
    stencil = np.stack([H.hash_cells(cells + o) % n for o in offs], 1)  # (n, 27)
    mult = (stencil[:, :, None] == key[None, None, :]).sum(1)  # (n, n)
    assert mult[inside].min() >= 1 and mult.max() >= 2  # collisions do occur at this table size
    want = (np.where(inside, (h - d) ** 2 * 15 / (2 * np.pi * h ** 5), 0) * mult).sum(1)
    want_near = (np.where(inside, (h - d) ** 3 * 15 / (np.pi * h ** 6), 0) * mult).sum(1)
    np.testing.assert_allclose(rho, want[pid], rtol=2e-5)
    np.testing.assert_allclose(near, want_near[pid], rtol=2e-5)


def test_isolated_pair_conserves_momentum():
    x0 = np.array([[2.0, 2.0, 2.0], [2.08, 2.05, 1.97]], np.float32)
    v0 = np.array([[0.3, 0.0, -0.2], [-0.1, 0.4, 0.0]], np.float32)
    p = _params(2, gravity=(0.0, 0.0, 0.0), iters=1)
    o = H.OracleLague(p, V.make_lague_params(), fp64=True)
    o.set_state(x0, v0)
    o.substeps(1)
    _, v, _, _, _ = o.get_state()
    np.testing.assert_allclose(v.sum(0), v0.astype(np.float64).sum(0), atol=1e-12)
    assert np.abs(v - v0[np.argsort(np.argsort([0, 1]))]).max() > 1e-3  # the pair interacts


def test_fp32_and_fp64_twins_agree():
    w = CF.lague_dam_break(counts=(10, 12, 10))
    p = _params(w.n)
    a = H.OracleLague(p, V.make_lague_params())
    b = H.OracleLague(p, V.make_lague_params(), fp64=True)
    for s in (a, b):
        s.set_state(w.x, w.v)
        s.substeps(3)
    xa, va, ra, _, _ = a.get_state()
    xb, vb, rb, _, _ = b.get_state()
    assert H.rel_err(xa, xb) < 1e-5 and H.rel_err(ra, rb) < 1e-4
    assert H.rel_err(va, vb) < 1e-3


def test_lague_workload_matches_reference_scene():
    w = CF.lague_dam_break()
    assert w.n == 150_000
    step = np.float32(np.cbrt(np.float32(1.0) / np.float32(630.0)))
    assert np.isclose(w.x[1, 2] - w.x[0, 2], step)
    assert w.x[:, 0].min() == 0 and w.x[:, 1].min() == 0
    # centred in z in the 10-wide box
    assert abs((w.x[:, 2].min() + w.x[:, 2].max() + step) / 2 - 5.0) < step
