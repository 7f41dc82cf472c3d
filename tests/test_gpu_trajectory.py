"""Trajectory export (vfs_trajectory_*): frames read back from the file equal
the simulation state at the recorded substeps, bit for bit, and recording does
not change the run."""
import numpy as np
import pytest

import vkfluidsim_amd as V
from vkfluidsim_amd import configs as CF

pytestmark = pytest.mark.gpu


def _sim(w, flags):
    grids = V.boundary_grids(w.shapes, w.h)
    p = V.make_params(w.n, w.h, dt=w.dt, box_size=w.box_size, box_pos=w.box_pos, flags=flags)
    sim = V.Sim(p, grids)
    sim.set_state(w.x, w.v)
    return sim


def test_recorded_frames_match_state(tmp_path):
    w = CF.dam_break(side=(10, 8, 12), map_res=24)
    print(f"case dam break {w.n} particles", flush=True)
    sim = _sim(w, V.VFS_FLAG_TRACK_IDS)
    path = tmp_path / "run.vfstraj"
    everything = V.VFS_TRAJ_POSITIONS | V.VFS_TRAJ_VELOCITIES | V.VFS_TRAJ_DENSITIES | V.VFS_TRAJ_IDS
    want = []
    with V.Trajectory(sim, path, everything) as t:
        done = 0
        for k in (0, 5, 7, 1):
            sim.substeps(k) if k else None
            done += k
            t.record(done, done * w.dt)
            want.append((done, sim.get_state(ids=True)))
    tr = V.read_trajectory(str(path))
    assert tr["n_particles"] == w.n and tr["fields"] == everything
    assert tr["objects"].shape == (1, 16)
    np.testing.assert_array_equal(tr["substep"], [d for d, _ in want])
    np.testing.assert_allclose(tr["time"], [d * w.dt for d, _ in want])
    for f, (_, (x, v, rho, pid)) in enumerate(want):
        np.testing.assert_array_equal(tr["positions"][f], x)
        np.testing.assert_array_equal(tr["velocities"][f], v)
        np.testing.assert_array_equal(tr["densities"][f], rho)
        np.testing.assert_array_equal(tr["particle_ids"][f], pid)
    # recording does not disturb the simulation
    ref = _sim(w, V.VFS_FLAG_TRACK_IDS)
    ref.substeps(13)
    np.testing.assert_array_equal(ref.get_state()[0], want[-1][1][0])


def test_positions_only_and_errors(tmp_path):
    w = CF.dam_break(side=(6, 6, 6), map_res=20)
    sim = _sim(w, 0)
    with pytest.raises(V.VfsError, match="VFS_FLAG_TRACK_IDS"):
        V.Trajectory(sim, tmp_path / "ids.vfstraj", V.VFS_TRAJ_IDS)
    with pytest.raises(V.VfsError):
        V.Trajectory(sim, tmp_path / "none.vfstraj", 0)
    p = tmp_path / "x.vfstraj"
    t = V.Trajectory(sim, p, V.VFS_TRAJ_POSITIONS)
    t.record(0, 0.0)
    sim.substeps(3)
    t.record(3, 3 * w.dt)
    t.close()
    tr = V.read_trajectory(str(p))
    assert set(k for k in tr if k in ("positions", "velocities", "densities", "particle_ids")) == {"positions"}
    assert tr["positions"].shape == (2, w.n, 3)
    np.testing.assert_array_equal(tr["positions"][1], sim.get_state()[0])
    assert p.stat().st_size == 64 + 64 + 2 * (16 + 12 * w.n)
