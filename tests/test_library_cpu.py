"""CPU-tier checks of the C ABI library: it loads, exports every declared
symbol, and rejects bad arguments before touching a device."""
import ctypes as C
from pathlib import Path

import pytest

import vkfluidsim_amd as V
from vkfluidsim_amd import _lib

ROOT = Path(__file__).resolve().parents[1]
SCENES = ROOT / "tests" / "golden" / "scenes"


def test_every_declared_symbol_is_exported():
    names = _lib.symbols_declared(ROOT / "include" / "vfs_c.h")
    assert len(names) >= 40
    missing = [n for n in names if not _lib.exported(n)]
    assert missing == []
    # and the binding covers all of them
    assert sorted(_lib.SIGNATURES) == names


def test_abi_version_and_device_count():
    assert V.lib.vfs_abi_version() == 2
    assert V.device_count() >= 0


def _err(fn, *args):
    rc = fn(*args)
    return rc, V.lib.vfs_last_error().decode()


def test_sim_create_rejects_bad_parameters():
    out = C.c_void_p()
    rc, msg = _err(V.lib.vfs_sim_create, None, None, 0, C.byref(out))
    assert rc == -1 and "null" in msg
    p = V.make_params(10, 0.04, dt=1e-4)
    p.iterations = 0
    rc, msg = _err(V.lib.vfs_sim_create, C.byref(p), None, 0, C.byref(out))
    assert rc == -1 and "iterations" in msg
    p = V.make_params(-5, 0.04, dt=1e-4)
    rc, msg = _err(V.lib.vfs_sim_create, C.byref(p), None, 0, C.byref(out))
    assert rc == -1 and "n_particles" in msg
    p = V.make_params(10, 0.04, dt=1e-4, cell_size=0.02)
    rc, msg = _err(V.lib.vfs_sim_create, C.byref(p), None, 0, C.byref(out))
    assert rc == -1 and "cell_size" in msg
    # the smoothing radius keeps the force terms' divisors where the force
    # pass's reciprocal is exact
    for h in (0.0, 2.0 ** -41, 2.0 ** 41, float("inf"), float("nan")):
        p = V.make_params(10, h, dt=1e-4, cell_size=1.0)
        rc, msg = _err(V.lib.vfs_sim_create, C.byref(p), None, 0, C.byref(out))
        assert rc == -1 and "smooth_radius" in msg
    p = V.make_params(10, 0.04, dt=1e-4)
    rc, msg = _err(V.lib.vfs_sim_create, C.byref(p), None, 1, C.byref(out))
    assert rc == -1 and "objects" in msg
    o = V.BoundaryObject()
    rc, msg = _err(V.lib.vfs_sim_create, C.byref(p), C.byref(o), 1, C.byref(out))
    assert rc == -1 and "grid" in msg


def test_sizes_and_workspaces():
    ws = C.c_size_t(0)
    assert V.lib.vfs_sort_workspace_size(1 << 30, 5, C.byref(ws)) == -1
    assert V.lib.vfs_sort_workspace_size(1000, 999, C.byref(ws)) == 0 and ws.value > 8000
    assert V.lib.vfs_scan_workspace_size(1000, C.byref(ws)) == 0 and ws.value > 0
    h = C.c_void_p()
    assert V.lib.vfs_spatial_hash_create(10, 0.0, C.byref(h)) == -1
    assert V.lib.vfs_sim_substeps(None, 1, None) == -1


@pytest.mark.parametrize("name,code,needle", [("malformed.json", -4, "parse error"),
                                              ("missing_key.json", -4, "outerWallStiffness"),
                                              ("does_not_exist.json", -4, "cannot open"),
                                              ("mesh_objects.json", -4, "cannot open OBJ file")])
def test_scene_loader_errors(name, code, needle):
    out = C.c_void_p()
    rc, msg = _err(V.lib.vfs_scene_load, str(SCENES / name).encode(), b"", C.byref(out))
    assert rc == code and needle in msg
    assert out.value is None


def test_trajectory_argument_checks(tmp_path):
    out = C.c_void_p()
    rc, msg = _err(V.lib.vfs_trajectory_open, None, str(tmp_path / "t").encode(), 1, C.byref(out))
    assert rc == -1 and "null" in msg
    assert V.lib.vfs_trajectory_record(None, 0, 0.0) == -1
    assert V.lib.vfs_trajectory_close(None) == 0


def test_cpp_facade_example_links():
    """The C++ facade examples link against libvfs.so and start (usage path,
    no device work)."""
    import subprocess
    from pathlib import Path
    for name in ("facade_scene", "facade_lague"):
        exe = Path(__file__).resolve().parent / "cpp" / name
        if not exe.exists():
            subprocess.run(["make", "-C", str(exe.parent)], check=True, capture_output=True)
        r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
        assert r.returncode == 2 and "usage" in r.stderr


def test_render_buffer_argument_checks():
    assert C.sizeof(V._lib.RenderBuffers) == 48 and C.sizeof(V._lib.RenderIpc) == 264
    rc, msg = _err(V.lib.vfs_render_buffers_create, 10, 0, None)
    assert rc == -1 and "null" in msg
    rc, msg = _err(V.lib.vfs_render_buffers_create, 1 << 28, 0, C.byref(V._lib.RenderBuffers()))
    assert rc == -1 and "2^28" in msg
    rc, msg = _err(V.lib.vfs_sim_copy_data_buffers, None, None, None)
    assert rc == -1 and "null" in msg
    rc, msg = _err(V.lib.vfs_render_buffers_export, None, None)
    assert rc == -1 and "null" in msg
    rc, msg = _err(V.lib.vfs_render_buffers_open, None, None)
    assert rc == -1 and "null" in msg
    b = V._lib.RenderBuffers()
    rc, msg = _err(V.lib.vfs_render_buffers_export, C.byref(b), C.byref(V._lib.RenderIpc()))
    assert rc == -1 and "incomplete" in msg
    b.positions = b.velocities = b.densities = 1
    b.n_particles = 200_000_000
    rc, msg = _err(V.lib.vfs_render_buffers_export, C.byref(b), C.byref(V._lib.RenderIpc()))
    assert rc == -1 and "2 GiB" in msg
    assert V.lib.vfs_render_buffers_destroy(None) == 0


def test_state_check_flag_is_for_single_scenes():
    p = V.make_params(10, 0.04, dt=1e-4, flags=V.VFS_FLAG_CHECK_STATE)
    arr = (V.Params * 2)(p, p)
    counts = (C.c_uint32 * 2)(0, 0)
    out = C.c_void_p()
    rc, msg = _err(V.lib.vfs_batch_create, arr, 2, None, counts, C.byref(out))
    assert rc == -1 and "VFS_FLAG_CHECK_STATE" in msg
    rc, msg = _err(V.lib.vfs_sim_check_state, None, 0.0, None, None)
    assert rc == -1 and "null" in msg
