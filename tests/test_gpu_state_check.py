"""Failure detection on the product path: NaN injected into the positions
(hash and state against the oracle, whose float-to-cell conversion maps NaN to
cell 0), the opt-in state check (vfs_sim_check_state, VFS_FLAG_CHECK_STATE ->
VFS_ERR_DIVERGED) on NaN, escaped and diverging states, and a stable run that
passes it."""
import numpy as np
import pytest

import harness as H
import vkfluidsim_amd as V
from vkfluidsim_amd import configs as CF

pytestmark = pytest.mark.gpu


def _dam_break(flags=V.VFS_FLAG_TRACK_IDS):
    w = CF.dam_break(side=(12, 12, 12), map_res=24)
    grid = V.volume_map(*w.shapes[0][:2], w.shapes[0][2], w.h)
    p = V.make_params(w.n, w.h, dt=w.dt, box_size=w.box_size, flags=flags)
    return w, grid, p


def test_nan_position_matches_oracle_and_is_reported():
    w, grid, p = _dam_break()
    x = w.x.copy()
    x[5, 1] = np.nan      # a particle inside the block
    x[700, :] = np.nan    # and one with no finite coordinate
    sim = V.Sim(p, [grid])
    sim.set_state(x, w.v)
    orc = H.OracleSim(p, [grid])
    orc.set_state(x, w.v)
    for k in range(3):
        sim.substeps(1)
        orc.substeps(1)
        print(f"substep {k}", flush=True)
        for a, b in zip(sim.get_hash(), orc.get_hash()):
            np.testing.assert_array_equal(a, b)
        gx, gv, grho, gpid = sim.get_state(ids=True)
        ox, ov, orho, opid, _ = orc.get_state()
        np.testing.assert_array_equal(gpid, opid)
        # NaN compares equal to NaN here: the same entries are non-finite
        np.testing.assert_array_equal(gx.astype(np.float64), ox)
        np.testing.assert_array_equal(gv.astype(np.float64), ov)
        np.testing.assert_array_equal(grho.astype(np.float64), orho)
    bad_rows = int(np.sum(~(np.isfinite(ox).all(axis=1) & np.isfinite(ov).all(axis=1))))
    assert bad_rows >= 2
    bad, out = sim.check_state(0.0)
    assert bad == bad_rows
    sim.close()


def test_check_state_flag_reports_nan_and_escaped_particles():
    w, grid, p = _dam_break(flags=V.VFS_FLAG_CHECK_STATE)
    sim = V.Sim(p, [grid])
    sim.set_state(w.x, w.v)
    sim.substeps(20)  # a stable run passes the check
    assert sim.check_state(0.5 * w.h) == (0, 0)
    x, v, _, _ = sim.get_state()
    x[3] = (50.0, 0.4, 0.4)  # far outside the 1.6 x 0.8 x 0.8 box
    x[9, 2] = np.inf
    sim.set_state(x, v)
    with pytest.raises(V.VfsError, match="1 particles with a non-finite position or velocity, 1 outside"):
        sim.substeps(1)
    # the explicit count with a margin that also takes the pushed particle
    bad, out = sim.check_state(0.0)
    assert bad >= 1 and out >= 1
    sim.close()


def test_check_state_flag_catches_the_divergence_of_the_configured_1m_scene():
    """The 1M obstacle scene as configured (viscosity 0.01) falls freely and
    diverges when it reaches the floor; with the flag the step call that runs
    into it fails instead of returning a garbage state."""
    w = CF.obstacle_1m(mesh=False)
    grids = V.boundary_grids(w.shapes, w.h)
    sim = V.Sim(V.make_params(w.n, w.h, dt=w.dt, box_size=w.box_size, box_pos=w.box_pos,
                              flags=V.VFS_FLAG_CHECK_STATE), grids)
    sim.set_state(w.x, w.v)
    done = 0
    with pytest.raises(V.VfsError, match="VFS_ERR_DIVERGED"):
        while done < 1000:
            sim.substeps(25)
            done += 25
    print(f"divergence reported after {done + 25} substeps", flush=True)
    assert 250 <= done < 600
    sim.close()
