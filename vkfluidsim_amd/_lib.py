"""ctypes binding of libvfs.so (include/vfs_c.h).

The library is the product: every call below goes to the gfx950 kernels.  If
libvfs.so is missing the import of this module raises; there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libvfs.so"

VFS_FLAG_TRACK_IDS = 0x1
VFS_FLAG_NO_GRAPH = 0x2
VFS_FLAG_STORE_ACCEL = 0x4
VFS_FLAG_CHECK_STATE = 0x8
VFS_NEIGHBOUR_CAPACITY = 64

VFS_SHAPE_BOX = 0
VFS_SHAPE_SPHERE = 1
VFS_SHAPE_BOX_WALLS = 2

STATUS = {0: "VFS_OK", -1: "VFS_ERR_INVALID", -2: "VFS_ERR_HIP", -3: "VFS_ERR_STATE",
          -4: "VFS_ERR_IO", -5: "VFS_ERR_NOMEM", -6: "VFS_ERR_DIVERGED"}


class VfsError(RuntimeError):
    def __init__(self, code: int, what: str, msg: str):
        super().__init__(f"{what}: {STATUS.get(code, code)}: {msg}")
        self.code = code


class LagueParams(C.Structure):
    """vfs_lague_params (LagueModel::Parameters, lague_model.h:19-24)."""
    _fields_ = [("wall_damping_factor", C.c_float), ("pressure_multiplier", C.c_float),
                ("near_pressure_multiplier", C.c_float), ("viscosity_strenght", C.c_float)]


class Params(C.Structure):
    """vfs_params (SimulationParameters + SPHModel::Parameters + WCSPH parameters)."""
    _fields_ = [
        ("gravity", C.c_float * 3),
        ("smooth_radius", C.c_float),
        ("time_scale", C.c_float),
        ("iterations", C.c_int32),
        ("n_particles", C.c_int32),
        ("fixed_dt", C.c_float),
        ("target_density", C.c_float),
        ("bbox_size", C.c_float * 3),
        ("bbox_pos", C.c_float * 3),
        ("wall_stiffness", C.c_float),
        ("stiffness", C.c_float),
        ("expoent", C.c_float),
        ("viscosity_strenght", C.c_float),
        ("cell_size", C.c_float),
        ("flags", C.c_uint32),
    ]


class BoundaryObject(C.Structure):
    """vfs_boundary_object (BoundaryObjectInfo with host grids)."""
    _fields_ = [
        ("transform", C.c_float * 16),
        ("rotation", C.c_float * 16),
        ("box_size", C.c_float * 3),
        ("box_pos", C.c_float * 3),
        ("resolution", C.c_uint32 * 3),
        ("reserved", C.c_uint32),
        ("sdf_grid", C.POINTER(C.c_float)),
        ("volume_map_grid", C.POINTER(C.c_float)),
    ]


class Tuning(C.Structure):
    """vfs_tuning: launch configuration (bit-identical results for every setting)."""
    _fields_ = [("sort_bits", C.c_int32), ("sort_items", C.c_int32), ("block_size", C.c_int32),
                ("stage_rows", C.c_int32), ("cell_scale", C.c_float)]

    def as_tuple(self):
        return (self.sort_bits, self.sort_items, self.block_size, self.stage_rows, self.cell_scale)


class DensityGrid(C.Structure):
    """vfs_density_grid: sample points origin + (i, j, k) * spacing."""
    _fields_ = [("origin", C.c_float * 3), ("spacing", C.c_float * 3), ("dims", C.c_uint32 * 3)]


class DeviceBuffers(C.Structure):
    _fields_ = [
        ("positions", C.c_void_p),
        ("velocities", C.c_void_p),
        ("densities", C.c_void_p),
        ("accelerations", C.c_void_p),
        ("particle_ids", C.c_void_p),
        ("sorted_keys", C.c_void_p),
        ("sorted_indices", C.c_void_p),
        ("cell_start", C.c_void_p),
        ("n_particles", C.c_uint32),
    ]


class RenderBuffers(C.Structure):
    _fields_ = [
        ("positions", C.c_void_p),
        ("velocities", C.c_void_p),
        ("densities", C.c_void_p),
        ("accelerations", C.c_void_p),
        ("ready", C.c_void_p),
        ("n_particles", C.c_uint32),
        ("opened", C.c_uint32),
    ]


class RenderIpc(C.Structure):
    _fields_ = [
        ("buffers", (C.c_ubyte * 64) * 4),
        ("n_particles", C.c_uint32),
        ("has_accelerations", C.c_uint32),
    ]


assert C.sizeof(Params) == 84
assert C.sizeof(BoundaryObject) == 184

_vp = C.c_void_p
_u32 = C.c_uint32
_i32 = C.c_int
_sz = C.c_size_t
_fp = C.POINTER(C.c_float)
_up = C.POINTER(C.c_uint32)

# name -> (restype, argtypes)
SIGNATURES = {
    "vfs_last_error": (C.c_char_p, []),
    "vfs_abi_version": (_i32, []),
    "vfs_device_count": (_i32, [C.POINTER(C.c_int)]),
    "vfs_dev_malloc": (_i32, [C.POINTER(_vp), _sz]),
    "vfs_dev_free": (_i32, [_vp]),
    "vfs_memcpy_h2d": (_i32, [_vp, _vp, _sz]),
    "vfs_memcpy_d2h": (_i32, [_vp, _vp, _sz]),
    "vfs_memcpy_d2d": (_i32, [_vp, _vp, _sz]),
    "vfs_memset_d32": (_i32, [_vp, _u32, _sz]),
    "vfs_stream_create": (_i32, [C.POINTER(_vp)]),
    "vfs_stream_destroy": (_i32, [_vp]),
    "vfs_stream_synchronize": (_i32, [_vp]),
    "vfs_sim_create": (_i32, [C.POINTER(Params), C.POINTER(BoundaryObject), _u32, C.POINTER(_vp)]),
    "vfs_sim_destroy": (_i32, [_vp]),
    "vfs_sim_set_state": (_i32, [_vp, _fp, _fp, _u32, C.c_int64]),
    "vfs_sim_set_params": (_i32, [_vp, C.POINTER(Params)]),
    "vfs_sim_get_params": (_i32, [_vp, C.POINTER(Params)]),
    "vfs_sim_step": (_i32, [_vp, _vp]),
    "vfs_sim_substeps": (_i32, [_vp, _u32, _vp]),
    "vfs_sim_synchronize": (_i32, [_vp]),
    "vfs_sim_get_state": (_i32, [_vp, _fp, _fp, _fp, _up]),
    "vfs_sim_get_hash": (_i32, [_vp, _up, _up, _up]),
    "vfs_sim_get_boundary": (_i32, [_vp, _fp, _fp]),
    "vfs_sim_device_buffers": (_i32, [_vp, C.POINTER(DeviceBuffers)]),
    "vfs_sim_get_tuning": (_i32, [_vp, C.POINTER(Tuning)]),
    "vfs_sim_set_tuning": (_i32, [_vp, C.POINTER(Tuning)]),
    "vfs_sim_autotune": (_i32, [_vp, _u32, C.POINTER(Tuning), C.POINTER(C.c_float)]),
    "vfs_sim_profile_stages": (_i32, [_vp, _u32, _fp, _up, C.POINTER(C.c_char_p), _u32, _up]),
    "vfs_sim_neighbour_stats": (_i32, [_vp, C.POINTER(C.c_double), _up, _up]),
    "vfs_sim_sort_stats": (_i32, [_vp, _up]),
    "vfs_sim_check_state": (_i32, [_vp, C.c_float, _up, _up]),
    "vfs_batch_create": (_i32, [_vp, _u32, _vp, _up, C.POINTER(_vp)]),
    "vfs_batch_destroy": (_i32, [_vp]),
    "vfs_batch_info": (_i32, [_vp, _up, C.POINTER(C.c_uint64), _up]),
    "vfs_batch_set_state": (_i32, [_vp, _u32, _fp, _fp]),
    "vfs_batch_set_params": (_i32, [_vp, _u32, _vp]),
    "vfs_batch_substeps": (_i32, [_vp, _u32, _vp]),
    "vfs_batch_synchronize": (_i32, [_vp]),
    "vfs_batch_get_state": (_i32, [_vp, _u32, _fp, _fp, _fp, _up]),
    "vfs_batch_get_hash": (_i32, [_vp, _u32, _up, _up, _up]),
    "vfs_batch_get_boundary": (_i32, [_vp, _u32, _fp, _fp]),
    "vfs_batch_get_tuning": (_i32, [_vp, C.POINTER(Tuning)]),
    "vfs_batch_set_tuning": (_i32, [_vp, C.POINTER(Tuning)]),
    "vfs_batch_autotune": (_i32, [_vp, _u32, _i32, C.POINTER(Tuning), C.POINTER(C.c_float)]),
    "vfs_batch_profile_stages": (_i32, [_vp, _u32, _fp, _up, C.POINTER(C.c_char_p), _u32, _up]),
    "vfs_batch_neighbour_stats": (_i32, [_vp, C.POINTER(C.c_double), _up, _up]),
    "vfs_adjoint_create": (_i32, [_vp, _u32, _u32, _vp, C.POINTER(_vp)]),
    "vfs_adjoint_destroy": (_i32, [_vp]),
    "vfs_adjoint_forward": (_i32, [_vp, _vp, _vp, _vp, _vp, C.POINTER(C.c_double)]),
    "vfs_adjoint_backward": (_i32, [_vp, _vp, _vp, C.POINTER(C.c_double)]),
    "vfs_adjoint_profile_backward": (_i32, [_vp, _fp, _up, C.POINTER(C.c_char_p), _u32, _up]),
    "vfs_spatial_hash_create": (_i32, [_u32, C.c_float, C.POINTER(_vp)]),
    "vfs_spatial_hash_set_cell_size": (_i32, [_vp, C.c_float]),
    "vfs_spatial_hash_run": (_i32, [_vp, _vp, _u32, _vp]),
    "vfs_spatial_hash_buffers": (_i32, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    "vfs_spatial_hash_destroy": (_i32, [_vp]),
    "vfs_sort_workspace_size": (_i32, [_u32, _u32, C.POINTER(_sz)]),
    "vfs_sort_pairs": (_i32, [_vp, _vp, _u32, _u32, _vp, _sz, _vp]),
    "vfs_sort_pairs_noalloc": (_i32, [_vp, _vp, _u32, _u32, _vp]),
    "vfs_scan_workspace_size": (_i32, [_u32, C.POINTER(_sz)]),
    "vfs_scan_exclusive": (_i32, [_vp, _u32, _vp, _sz, _vp]),
    "vfs_spatial_offsets": (_i32, [_vp, _vp, _u32, _i32, _vp]),
    "vfs_reorder_f32x3": (_i32, [_vp, _vp, _vp, _u32, _vp]),
    "vfs_volume_map_generate": (_i32, [_i32, C.POINTER(C.c_double), _up, C.c_double, C.c_double,
                                       C.c_double, _fp, _fp, _fp, _fp, _i32]),
    "vfs_mesh_load_obj": (_i32, [C.c_char_p, C.POINTER(_vp)]),
    "vfs_mesh_create": (_i32, [_fp, _u32, _up, _u32, C.POINTER(_vp)]),
    "vfs_mesh_info": (_i32, [_vp, _up, _up, _fp, _fp]),
    "vfs_mesh_get": (_i32, [_vp, _fp, _up]),
    "vfs_mesh_properties": (_i32, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "vfs_mesh_signed_distance": (_i32, [_vp, _fp, _u32, C.POINTER(C.c_double)]),
    "vfs_mesh_volume_map_generate": (_i32, [_vp, _up, C.c_double, C.c_double, C.c_double, _fp, _fp, _fp, _fp]),
    "vfs_mesh_destroy": (_i32, [_vp]),
    "vfs_trajectory_open": (_i32, [_vp, C.c_char_p, _u32, C.POINTER(_vp)]),
    "vfs_trajectory_record": (_i32, [_vp, C.c_uint64, C.c_double]),
    "vfs_trajectory_close": (_i32, [_vp]),
    "vfs_render_buffers_create": (_i32, [_u32, C.c_int, C.POINTER(RenderBuffers)]),
    "vfs_render_buffers_destroy": (_i32, [C.POINTER(RenderBuffers)]),
    "vfs_sim_copy_data_buffers": (_i32, [_vp, C.POINTER(RenderBuffers), _vp]),
    "vfs_render_buffers_export": (_i32, [C.POINTER(RenderBuffers), C.POINTER(RenderIpc)]),
    "vfs_render_buffers_open": (_i32, [C.POINTER(RenderIpc), C.POINTER(RenderBuffers)]),
    "vfs_render_buffers_wait": (_i32, [C.POINTER(RenderBuffers), _vp]),
    "vfs_render_buffers_synchronize": (_i32, [C.POINTER(RenderBuffers)]),
    "vfs_render_buffers_read": (_i32, [C.POINTER(RenderBuffers), _vp, _vp, _vp, _vp]),
    "vfs_lague_create": (_i32, [C.POINTER(Params), C.POINTER(LagueParams), C.POINTER(_vp)]),
    "vfs_lague_set_params": (_i32, [_vp, C.POINTER(LagueParams)]),
    "vfs_lague_get_params": (_i32, [_vp, C.POINTER(LagueParams)]),
    "vfs_scene_load": (_i32, [C.c_char_p, C.c_char_p, C.POINTER(_vp)]),
    "vfs_scene_sim": (_i32, [_vp, C.POINTER(_vp)]),
    "vfs_scene_reset": (_i32, [_vp]),
    "vfs_scene_step": (_i32, [_vp, _vp]),
    "vfs_scene_boundary_objects": (_i32, [_vp, C.POINTER(BoundaryObject), _u32, _up]),
    "vfs_scene_destroy": (_i32, [_vp]),
}


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built: run __graft_entry__.build() or make -C vkfluidsim_amd")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def check(code: int, what: str) -> None:
    if code != 0:
        msg = lib.vfs_last_error()
        raise VfsError(code, what, msg.decode() if msg else "")


def call(name: str, *args) -> None:
    check(getattr(lib, name)(*args), name)


def device_count() -> int:
    c = C.c_int(0)
    call("vfs_device_count", C.byref(c))
    return c.value


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_fp)


def uptr(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags.c_contiguous
    return a.ctypes.data_as(_up)


class DeviceArray:
    """Device allocation owned by Python (plumbing for tests and the bench)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = _vp()
        call("vfs_dev_malloc", C.byref(p), max(self.nbytes, 1))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        if a.nbytes:
            call("vfs_memcpy_h2d", d.ptr, a.ctypes.data, a.nbytes)
        return d

    def to_host(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            call("vfs_memcpy_d2h", out.ctypes.data, self.ptr, out.nbytes)
        return out

    def free(self) -> None:
        if self.ptr:
            lib.vfs_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def symbols_declared(header: Path) -> list[str]:
    """Names of the functions include/vfs_c.h declares."""
    import re
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vfs_[a-z0-9_]+)\s*\(", text)))


def exported(name: str) -> bool:
    try:
        getattr(lib, name)
        return True
    except AttributeError:
        return False


if os.environ.get("VFS_DEBUG_LOAD"):
    print(f"libvfs loaded from {LIB_PATH}")
