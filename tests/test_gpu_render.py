"""Render-buffer handoff (vfs_render_*, SPHModel::CreateDataBuffers /
CopyDataBuffers): the buffers hold exactly the simulation's state in the
reference's layout, in this process and, mapped through IPC handles, in a
renderer process of its own."""
import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import vkfluidsim_amd as V
from vkfluidsim_amd import configs as CF

pytestmark = pytest.mark.gpu

HERE = Path(__file__).resolve().parent


def _sim(flags=V.VFS_FLAG_STORE_ACCEL):
    w = CF.dam_break(side=(12, 12, 12), map_res=24)
    grid = V.volume_map(*w.shapes[0][:2], w.shapes[0][2], w.h)
    sim = V.Sim(V.make_params(w.n, w.h, dt=w.dt, box_size=w.box_size, flags=flags), [grid])
    sim.set_state(w.x, w.v)
    return sim


def test_copy_data_buffers_hold_the_state():
    sim = _sim()
    print(f"case render copy: {sim.n} particles", flush=True)
    sim.substeps(7)
    with V.RenderBuffers(sim.n, with_accelerations=True) as rb:
        rb.copy_from(sim)
        x, v, rho, a = rb.read()
        sx, sv, srho, _ = sim.get_state()
        np.testing.assert_array_equal(x, sx)
        np.testing.assert_array_equal(v, sv)
        np.testing.assert_array_equal(rho, srho)
        assert np.isfinite(a).all() and np.abs(a).max() > 0
        # the next copy follows the next substeps
        sim.substeps(3)
        rb.copy_from(sim)
        x2, _, _, _ = rb.read()
        np.testing.assert_array_equal(x2, sim.get_state()[0])
        assert not np.array_equal(x2, x)
    # accelerations are only handed over when the simulation stores them
    plain = _sim(flags=0)
    plain.substeps(1)
    with V.RenderBuffers(plain.n, with_accelerations=True) as rb:
        with pytest.raises(V.VfsError, match="STORE_ACCEL"):
            rb.copy_from(plain)
    with V.RenderBuffers(plain.n + 1) as rb:
        with pytest.raises(V.VfsError, match="particle count"):
            rb.copy_from(plain)
    plain.close()
    sim.close()


def test_renderer_process_maps_the_buffers(tmp_path):
    sim = _sim()
    print(f"case render ipc: {sim.n} particles, consumer process", flush=True)
    sim.substeps(5)
    rb = V.RenderBuffers(sim.n, with_accelerations=True)
    rb.copy_from(sim)
    rb.synchronize()
    handles = tmp_path / "handles.bin"
    handles.write_bytes(rb.export())
    out = tmp_path / "seen.npz"
    env = dict(os.environ)
    proc = subprocess.run([sys.executable, str(HERE / "render_consumer.py"), str(handles), str(out)], env=env,
                          capture_output=True, text=True, timeout=180)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    seen = np.load(out)
    x, v, rho, a = rb.read()
    sx, sv, srho, _ = sim.get_state()
    np.testing.assert_array_equal(seen["x"], sx)
    np.testing.assert_array_equal(seen["v"], sv)
    np.testing.assert_array_equal(seen["rho"], srho)
    np.testing.assert_array_equal(seen["a"], a)
    assert bool(seen["export_refused"])
    rb.close()
    sim.close()
