"""Parity at the benchmark sizes: the 1,048,576-particle obstacle workload (icosphere
mesh obstacle, 10 substeps) and the 8,388,608-particle uniform workload (3 substeps)
bit for bit against the FP32 oracle over whole substeps (the first sorts from
scratch, the others re-sort); the 256-scene batch against single-scene
simulations; size-independent properties after longer runs (sorted keys, a true
permutation, ids a permutation, offsets consistent, finite state); the 65,536-particle
differentiable dam break against the FP64 adjoint of the same trajectory."""
import numpy as np
import pytest

import harness as H
import vkfluidsim_amd as V
from vkfluidsim_amd import configs as CF

pytestmark = pytest.mark.gpu


def _params(w, flags=V.VFS_FLAG_TRACK_IDS):
    return V.make_params(w.n, w.h, dt=w.dt, box_size=w.box_size, box_pos=w.box_pos, flags=flags, **w.params_kw)


def _same(sim, orc):
    x, v, rho, pid = sim.get_state(ids=True)
    ox, ov, orho, opid, _ = orc.get_state()
    np.testing.assert_array_equal(pid, opid)
    np.testing.assert_array_equal(x.astype(np.float64), ox)
    np.testing.assert_array_equal(v.astype(np.float64), ov)
    np.testing.assert_array_equal(rho.astype(np.float64), orho)
    for a, b in zip(sim.get_hash(), orc.get_hash()):
        np.testing.assert_array_equal(a, b)


def _properties(sim, n):
    keys, perm, offsets = sim.get_hash()
    assert np.all(np.diff(keys.astype(np.int64)) >= 0)
    assert np.array_equal(np.sort(perm), np.arange(n, dtype=np.uint32))
    first = np.full(n, n, np.uint32)
    starts = np.flatnonzero(np.r_[True, keys[1:] != keys[:-1]])
    first[keys[starts]] = starts
    np.testing.assert_array_equal(offsets, first)
    x, v, rho, pid = sim.get_state(ids=True)
    assert np.array_equal(np.sort(pid), np.arange(n, dtype=np.uint32))
    assert np.isfinite(x).all() and np.isfinite(v).all() and np.isfinite(rho).all()


def test_obstacle_1m_substeps_bit_exact():
    w = CF.obstacle_1m()
    print(f"case obstacle_1m: {w.n} particles, icosphere mesh obstacle", flush=True)
    grids = V.boundary_grids(w.shapes, w.h)
    p = _params(w)
    sim = V.Sim(p, grids)
    sim.set_state(w.x, w.v)
    orc = H.OracleSim(p, grids)
    orc.set_state(w.x, w.v)
    sim.substeps(2)
    orc.substeps(2)
    _same(sim, orc)
    # boundary terms (the block starts 0.032 from the sphere, farther than h:
    # zero until the fluid reaches it, but sampled for every particle in the grid box)
    for a, b in zip(sim.get_boundary(), orc.get_boundary()):
        np.testing.assert_array_equal(np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel())
    # the substeps after the first re-sort against the previous sorted keys
    sim.substeps(8)
    orc.substeps(8)
    _same(sim, orc)
    del orc
    sim.substeps(10)
    _properties(sim, w.n)


def test_obstacle_1m_contact_phase_bit_exact():
    """The benchmark's contact-phase state: 1,500 substeps on the GPU bring the
    fluid onto the floor (pressure active on most particles, thousands of keys
    changing bucket per substep); from that state 2 + 8 substeps are compared
    with the oracle bit for bit."""
    w = CF.obstacle_1m_contact()
    print(f"case obstacle_1m_contact: {w.n} particles, {w.settle} settling substeps", flush=True)
    grids = V.boundary_grids(w.shapes, w.h)
    p = _params(w)
    sim = V.Sim(p, grids)
    sim.set_state(w.x, w.v)
    sim.substeps(w.settle)
    x, v, rho, _ = sim.get_state()
    assert np.mean(rho > 1000.0) > 0.5  # pressure acts on most particles
    sim.set_state(x, v)
    orc = H.OracleSim(p, grids)
    orc.set_state(x, v)
    sim.substeps(2)
    orc.substeps(2)
    _same(sim, orc)
    sim.substeps(8)
    orc.substeps(8)
    _same(sim, orc)
    assert sim.changed_keys() > 1000
    for a, b in zip(sim.get_boundary(), orc.get_boundary()):
        np.testing.assert_array_equal(np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel())


def test_uniform_8m_substep_bit_exact():
    w = CF.uniform_stress()
    print(f"case uniform_8m: {w.n} particles", flush=True)
    grids = V.boundary_grids(w.shapes, w.h)
    p = _params(w)
    sim = V.Sim(p, grids)
    sim.set_state(w.x, w.v)
    orc = H.OracleSim(p, grids)
    orc.set_state(w.x, w.v)
    sim.substeps(1)
    orc.substeps(1)
    _same(sim, orc)
    # fast particles: thousands of keys change per substep (the multi-tile re-sort)
    sim.substeps(2)
    orc.substeps(2)
    _same(sim, orc)
    del orc
    sim.substeps(4)
    _properties(sim, w.n)


def test_batched_256_scenes_match_single_scenes():
    ws = CF.batched_dam_breaks()
    print(f"case 256 scenes: {sum(w.n for w in ws)} particles", flush=True)
    grids = [V.boundary_grids(w.shapes, w.h) for w in ws]
    params = [_params(w) for w in ws]
    batch = V.Batch(params, grids)
    for k, w in enumerate(ws):
        batch.set_state(k, w.x, w.v)
    batch.substeps(10)
    # the smallest, the largest and a few in between, each against its own vfs_sim
    order = np.argsort([w.n for w in ws])
    for k in [int(order[0]), int(order[-1]), 17, 100, 201]:
        sim = V.Sim(params[k], grids[k])
        sim.set_state(ws[k].x, ws[k].v)
        sim.substeps(10)
        for a, b in zip(batch.get_state(k, ids=True), sim.get_state(ids=True)):
            np.testing.assert_array_equal(a, b)
        for a, b in zip(batch.get_hash(k), sim.get_hash()):
            np.testing.assert_array_equal(a, b)
        sim.close()
    for k in range(0, 256, 15):
        x, v, rho, _ = batch.get_state(k)
        assert np.isfinite(x).all() and np.isfinite(rho).all()
    batch.close()


def test_adjoint_65k_matches_fp64_adjoint_of_the_same_trajectory():
    w = CF.dam_break(side=(64, 32, 32), seed=CF.SEED_BASE + 4, steps=200)
    print(f"case adjoint: {w.n} particles, 200 substeps", flush=True)
    grids = V.boundary_grids(w.shapes, w.h)
    sim = V.Sim(V.make_params(w.n, w.h, dt=w.dt, box_size=w.box_size, box_pos=w.box_pos), grids)
    dims = (32, 32, 32)
    spacing = tuple(b / (d - 1) for b, d in zip(w.box_size, dims))
    adj = V.Adjoint(sim, 200, 20, (0.0, 0.0, 0.0), spacing, dims)
    v_t = (0.01 * CF.normal(CF.SEED_BASE + 4, 3 * w.n)).astype(np.float32).reshape(w.n, 3)
    _, target = adj.forward(w.x, v_t)
    loss, _ = adj.forward(w.x, w.v, target)
    gx, gv, gnu = adj.backward()
    assert loss > 0 and np.isfinite(gv).all() and np.isfinite(gnu)
    # FP32 central differences are noise at this size (the loss moves by a few
    # ulps of its FP32 trajectory); the reference is the FP64 adjoint of the
    # FP32 trajectory the GPU reproduces bit for bit
    L, mx, mv, mnu = H.oracle_adjoint_f32_trajectory(sim.params, grids, w.x, w.v, 200, (0.0, 0.0, 0.0), spacing,
                                                     dims, target.astype(np.float64))
    ex = np.linalg.norm(gx - mx) / np.linalg.norm(mx)
    ev = np.linalg.norm(gv - mv) / np.linalg.norm(mv)
    print(f"65,536 particles, 200 substeps: x0 {ex:.2e}, v0 {ev:.2e}, viscosity {gnu:.8e} vs {mnu:.8e}")
    # both evaluate the loss and the reverse substeps in double on the same
    # float tape (measured: 2.5e-8, the rounding of the float gradients)
    assert loss == pytest.approx(L, rel=1e-9)
    assert ex < 1e-6 and ev < 1e-6
    assert gnu == pytest.approx(mnu, rel=1e-6)
    adj.close()
    sim.close()
