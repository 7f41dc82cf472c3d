"""GPU tests of batched scenes (vfs_batch_*): every scene of a batch evolves
bit for bit as it does alone in a vfs_sim (and so as the CPU oracle)."""
import ctypes

import numpy as np
import pytest

import harness as H
import vkfluidsim_amd as V
from vkfluidsim_amd import configs as CF

pytestmark = pytest.mark.gpu


def _scene_params(w, flags=V.VFS_FLAG_TRACK_IDS):
    return V.make_params(w.n, w.h, dt=w.dt, box_size=w.box_size, box_pos=w.box_pos, flags=flags)


def _grids(w, res=None):
    return [V.volume_map(kind, sp, res or r, w.h) for kind, sp, r in w.shapes]


def _mixed_scenes():
    """Uneven sizes (not multiples of 64, one of 7 particles), different boxes,
    h and dt, with and without boundary objects."""
    ws = CF.batched_dam_breaks(n_scenes=5, map_res=16, n_min=300, n_max=5000)
    ws[2].shapes = []  # penalty walls only
    x = np.array([[0.10, 0.10, 0.10], [0.12, 0.11, 0.10], [0.15, 0.10, 0.12], [0.30, 0.30, 0.30],
                  [0.31, 0.33, 0.29], [0.11, 0.13, 0.14], [0.50, 0.10, 0.10]], np.float32)
    tiny = CF.Workload("tiny", 7, 0.05, 1e-4, 10, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), x, np.zeros_like(x), [])
    ws.insert(1, tiny)
    return ws


def _assert_scene_equal(batch, k, sim):
    for a, b in zip(batch.get_state(k, ids=True), sim.get_state(ids=True)):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(batch.get_hash(k), sim.get_hash()):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(batch.get_boundary(k), sim.get_boundary()):
        np.testing.assert_array_equal(a, b)


def test_batch_scenes_match_single_scenes_bit_exact():
    ws = _mixed_scenes()
    grids = [_grids(w) for w in ws]
    batch = V.Batch([_scene_params(w) for w in ws], grids)
    ns, npart, entries = batch.info()
    assert ns == len(ws) and npart == sum(w.n for w in ws) and entries % 64 == 0
    for k, w in enumerate(ws):
        batch.set_state(k, w.x, w.v)
    batch.substeps(5)
    for k, w in enumerate(ws):
        sim = V.Sim(_scene_params(w), grids[k])
        sim.set_state(w.x, w.v)
        sim.substeps(5)
        _assert_scene_equal(batch, k, sim)
        sim.close()
    # and one scene against the oracle directly
    w = ws[3]
    orc = H.OracleSim(_scene_params(w), grids[3])
    orc.set_state(w.x, w.v)
    orc.substeps(5)
    np.testing.assert_array_equal(batch.get_state(3)[0].astype(np.float64), orc.get_state()[0])
    np.testing.assert_array_equal(batch.get_hash(3)[1], orc.get_hash()[1])


def test_batch_tunings_and_eager_launches_are_bit_identical():
    ws = _mixed_scenes()
    grids = [_grids(w) for w in ws]
    ref = None
    for tune, flags in (((8, 16, 128, 1), V.VFS_FLAG_TRACK_IDS), ((10, 12, 256, 0), V.VFS_FLAG_TRACK_IDS),
                        ((9, 8, 128, 0), V.VFS_FLAG_TRACK_IDS | V.VFS_FLAG_NO_GRAPH)):
        batch = V.Batch([_scene_params(w, flags) for w in ws], grids)
        batch.set_tuning(*tune)
        for k, w in enumerate(ws):
            batch.set_state(k, w.x, w.v)
        batch.substeps(4)
        got = [batch.get_state(k, ids=True) for k in range(len(ws))]
        if ref is None:
            ref = got
        else:
            for a, b in zip(ref, got):
                for u, t in zip(a, b):
                    np.testing.assert_array_equal(u, t)
        batch.close()


def test_batch_scene_params_and_state_are_per_scene():
    ws = _mixed_scenes()[:3]
    grids = [_grids(w) for w in ws]
    batch = V.Batch([_scene_params(w) for w in ws], grids)
    for k, w in enumerate(ws):
        batch.set_state(k, w.x, w.v)
    p = _scene_params(ws[2])
    p.viscosity_strenght = 0.5
    p.gravity[1] = -3.0
    V._lib.call("vfs_batch_set_params", batch.handle, 2, ctypes.byref(p))
    batch.substeps(3)
    batch.synchronize()
    for k, w in enumerate(ws):
        sim = V.Sim(p if k == 2 else _scene_params(w), grids[k])
        sim.set_state(w.x, w.v)
        sim.substeps(3)
        _assert_scene_equal(batch, k, sim)


def test_batch_autotune_restores_state_and_may_change_cell_size():
    ws = _mixed_scenes()
    grids = [_grids(w) for w in ws]
    batch = V.Batch([_scene_params(w) for w in ws], grids)
    for k, w in enumerate(ws):
        batch.set_state(k, w.x, w.v)
    batch.substeps(2)
    before = [batch.get_state(k, ids=True) for k in range(len(ws))]
    best, ms = batch.autotune(2, allow_cell_size=True)
    assert ms > 0 and best.cell_scale in (1.0, 1.25, 1.5)
    after = [batch.get_state(k, ids=True) for k in range(len(ws))]
    for a, b in zip(before, after):
        for u, t in zip(a, b):
            np.testing.assert_array_equal(u, t)
    batch.substeps(2)
    for k in range(len(ws)):
        x, v, rho, _ = batch.get_state(k)
        assert np.isfinite(x).all() and np.isfinite(rho).all()


def test_batch_rejects_bad_input():
    w = CF.dam_break(side=(4, 4, 4), map_res=8)
    p = _scene_params(w)
    q = _scene_params(w, flags=0)
    with pytest.raises(V.VfsError):
        V.Batch([p, q])  # flags differ
    batch = V.Batch([p, p])
    with pytest.raises(V.VfsError):
        batch.set_state(2, w.x, w.v)
    with pytest.raises(V.VfsError):
        batch.get_hash(0)  # no substep yet
