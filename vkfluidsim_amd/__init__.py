"""vkfluidsim_amd: MI355X (gfx950) WCSPH substep engine behind a C ABI.

Python mirror of the reference model interface (luihabl/VkFluidSim
src/models/model.h:12-114, src/models/wcsph_with_boundary_model.h:14-49) over
libvfs.so.  Every method calls the HIP library; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import (BoundaryObject, DeviceArray, LagueParams, Params, Tuning, VfsError, VFS_FLAG_CHECK_STATE,
                   VFS_FLAG_NO_GRAPH, VFS_FLAG_STORE_ACCEL,
                   VFS_FLAG_TRACK_IDS, VFS_NEIGHBOUR_CAPACITY, VFS_SHAPE_BOX, VFS_SHAPE_BOX_WALLS, VFS_SHAPE_SPHERE, call, fptr, lib, uptr)

__all__ = ["Sim", "Batch", "Adjoint", "Params", "Tuning", "BoundaryGrid", "SpatialHash", "VfsError", "make_params", "volume_map",
           "sort_pairs", "scan_exclusive", "spatial_offsets", "reorder_f32x3", "Scene", "lib",
           "VFS_FLAG_TRACK_IDS", "VFS_FLAG_NO_GRAPH", "VFS_FLAG_STORE_ACCEL", "VFS_FLAG_CHECK_STATE",
           "VFS_NEIGHBOUR_CAPACITY",
           "VFS_SHAPE_BOX", "VFS_SHAPE_SPHERE", "VFS_SHAPE_BOX_WALLS", "device_count"]

device_count = _lib.device_count

IDENTITY16 = np.eye(4, dtype=np.float32).reshape(-1)


def make_params(n: int, h: float, *, dt: float, box_size=(1.0, 1.0, 1.0), box_pos=(0.0, 0.0, 0.0),
                rho0: float = 1000.0, stiffness: float = 1000.0, gamma: float = 7.0, viscosity: float = 0.01,
                gravity=(0.0, -9.81, 0.0), wall_stiffness: float = 1e4, iterations: int = 1,
                time_scale: float = 1.0, cell_size: float = 0.0, flags: int = 0) -> Params:
    p = Params()
    p.gravity[:] = [float(g) for g in gravity]
    p.smooth_radius = h
    p.time_scale = time_scale
    p.iterations = iterations
    p.n_particles = n
    p.fixed_dt = dt
    p.target_density = rho0
    p.bbox_size[:] = [float(s) for s in box_size]
    p.bbox_pos[:] = [float(s) for s in box_pos]
    p.wall_stiffness = wall_stiffness
    p.stiffness = stiffness
    p.expoent = gamma
    p.viscosity_strenght = viscosity
    p.cell_size = cell_size
    p.flags = flags
    return p


@dataclass
class BoundaryGrid:
    """One boundary object: SDF + volume-map grids and its transforms."""
    sdf: np.ndarray
    vm: np.ndarray
    resolution: tuple
    box_size: tuple
    box_pos: tuple
    transform: np.ndarray = field(default_factory=lambda: IDENTITY16.copy())
    rotation: np.ndarray = field(default_factory=lambda: IDENTITY16.copy())

    def as_struct(self) -> BoundaryObject:
        o = BoundaryObject()
        o.transform[:] = [float(v) for v in np.asarray(self.transform, np.float32)]
        o.rotation[:] = [float(v) for v in np.asarray(self.rotation, np.float32)]
        o.box_size[:] = [float(v) for v in self.box_size]
        o.box_pos[:] = [float(v) for v in self.box_pos]
        o.resolution[:] = [int(v) for v in self.resolution]
        o.sdf_grid = fptr(self.sdf)
        o.volume_map_grid = fptr(self.vm)
        return o


def volume_map(kind: int, shape_params, resolution, support_radius: float, *, tolerance: float = 0.0,
               margin: float | None = None) -> BoundaryGrid:
    """SDF + volume map of an analytic shape on the GPU (vfs_volume_map_generate)."""
    if margin is None:
        margin = 8.0 * support_radius
    res = np.ascontiguousarray(np.asarray(resolution, dtype=np.uint32))
    sp = np.zeros(6, dtype=np.float64)
    sp[: len(shape_params)] = shape_params
    cnt = int(np.prod(res.astype(np.int64)))
    sdf = np.empty(cnt, np.float32)
    vm = np.empty(cnt, np.float32)
    bs = np.empty(3, np.float32)
    bp = np.empty(3, np.float32)
    call("vfs_volume_map_generate", kind, sp.ctypes.data_as(C.POINTER(C.c_double)), uptr(res),
         support_radius, tolerance, margin, fptr(sdf), fptr(vm), fptr(bs), fptr(bp), 0)
    return BoundaryGrid(sdf, vm, tuple(int(r) for r in res), tuple(float(v) for v in bs),
                        tuple(float(v) for v in bp))


def boundary_grids(shapes, support_radius: float) -> list:
    """Boundary grids of a workload's shapes: (kind, params, resolution) for an
    analytic shape, ("mesh", (vertices, triangles), resolution) for a mesh."""
    out = []
    for kind, sp, res in shapes:
        if kind == "mesh":
            m = Mesh.from_arrays(*sp)
            out.append(m.volume_map(res, support_radius))
            m.close()
        else:
            out.append(volume_map(kind, sp, res, support_radius))
    return out


def make_lague_params(wall_damping_factor: float = 0.95, pressure_multiplier: float = 288.0,
                      near_pressure_multiplier: float = 2.16, viscosity: float = 0.03) -> LagueParams:
    """vfs_lague_params; the defaults are LagueModel's (lague_model.cpp:26-34)."""
    return LagueParams(wall_damping_factor, pressure_multiplier, near_pressure_multiplier, viscosity)


VFS_TRAJ_POSITIONS, VFS_TRAJ_VELOCITIES, VFS_TRAJ_DENSITIES, VFS_TRAJ_IDS = 1, 2, 4, 8
_TRAJ_HEADER = np.dtype([("magic", "S8"), ("version", "<u4"), ("n_particles", "<u4"), ("fields", "<u4"),
                         ("n_objects", "<u4"), ("n_frames", "<u8"), ("box_pos", "<f4", 3),
                         ("box_size", "<f4", 3), ("reserved", "<u4", 2)])


class Trajectory:
    """Recorder of particle-state frames into a trajectory file
    (vfs_trajectory_*); read files back with read_trajectory."""

    def __init__(self, sim, path: str, fields: int = VFS_TRAJ_POSITIONS | VFS_TRAJ_VELOCITIES):
        h = C.c_void_p()
        call("vfs_trajectory_open", sim.handle if hasattr(sim, "handle") else sim, str(path).encode(), int(fields),
             C.byref(h))
        self.handle = h.value

    def record(self, substep: int, time: float) -> None:
        call("vfs_trajectory_record", self.handle, int(substep), float(time))

    def close(self) -> None:
        if self.handle:
            call("vfs_trajectory_close", self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class RenderBuffers:
    """Render buffers in the reference's data-buffer layout (vfs_render_*):
    packed float3 positions, velocities (and accelerations), float densities.
    Fill them with copy_from(sim); hand them to another process with
    export() -> bytes and RenderBuffers.open(bytes) there."""

    def __init__(self, n_particles: int = 0, with_accelerations: bool = False, _opened=None):
        self.b = _lib.RenderBuffers()
        if _opened is not None:
            call("vfs_render_buffers_open", C.byref(_opened), C.byref(self.b))
        else:
            call("vfs_render_buffers_create", int(n_particles), int(bool(with_accelerations)), C.byref(self.b))
        self.n = int(self.b.n_particles)

    @classmethod
    def open(cls, handles: bytes) -> "RenderBuffers":
        ipc = _lib.RenderIpc.from_buffer_copy(handles)
        return cls(_opened=ipc)

    def copy_from(self, sim, stream=None) -> None:
        call("vfs_sim_copy_data_buffers", sim.handle, C.byref(self.b), stream)

    def export(self) -> bytes:
        ipc = _lib.RenderIpc()
        call("vfs_render_buffers_export", C.byref(self.b), C.byref(ipc))
        return bytes(ipc)

    def wait(self, stream=None) -> None:
        call("vfs_render_buffers_wait", C.byref(self.b), stream)

    def synchronize(self) -> None:
        call("vfs_render_buffers_synchronize", C.byref(self.b))

    def read(self):
        """(positions (n, 3), velocities (n, 3), densities (n,), accelerations (n, 3) or None)."""
        x = np.empty((self.n, 3), np.float32)
        v = np.empty((self.n, 3), np.float32)
        rho = np.empty(self.n, np.float32)
        a = np.empty((self.n, 3), np.float32) if self.b.accelerations else None
        call("vfs_render_buffers_read", C.byref(self.b), x.ctypes.data, v.ctypes.data, rho.ctypes.data,
             a.ctypes.data if a is not None else None)
        return x, v, rho, a

    def close(self) -> None:
        if self.b.positions:
            call("vfs_render_buffers_destroy", C.byref(self.b))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def read_trajectory(path: str) -> dict:
    """A trajectory file as numpy arrays: header fields, objects (k, 16),
    substep (f,), time (f,) and per field (f, n, 3) or (f, n) arrays."""
    raw = np.fromfile(path, dtype=np.uint8)
    hdr = raw[:64].view(_TRAJ_HEADER)[0]
    if hdr["magic"] != b"VFSTRAJ1" or hdr["version"] != 1:
        raise ValueError(f"{path}: not a version-1 trajectory file")
    n, fields, k, f = int(hdr["n_particles"]), int(hdr["fields"]), int(hdr["n_objects"]), int(hdr["n_frames"])
    off = 64
    out = {"n_particles": n, "fields": fields, "box_pos": tuple(hdr["box_pos"]), "box_size": tuple(hdr["box_size"]),
           "objects": raw[off:off + 64 * k].view("<f4").reshape(k, 16)}
    off += 64 * k
    parts = [("substep", "<u8", 1), ("time", "<f8", 1)]
    for bit, name, dt, w in [(1, "positions", "<f4", 3), (2, "velocities", "<f4", 3), (4, "densities", "<f4", 1),
                             (8, "particle_ids", "<u4", 1)]:
        if fields & bit:
            parts.append((name, dt, w * n))
    frame = np.dtype([(name, dt, (cnt,)) for name, dt, cnt in parts])
    frames = raw[off:off + frame.itemsize * f].view(frame)
    out["substep"] = frames["substep"][:, 0].copy()
    out["time"] = frames["time"][:, 0].copy()
    for bit, name, w in [(1, "positions", 3), (2, "velocities", 3), (4, "densities", 1), (8, "particle_ids", 1)]:
        if fields & bit:
            a = frames[name].copy()
            out[name] = a.reshape(f, n, 3) if w == 3 else a.reshape(f, n)
    return out


class Mesh:
    """A triangle mesh boundary object (vfs_mesh_*): OBJ loading on the host,
    signed distances and boundary grids on the GPU."""

    def __init__(self, handle: int):
        self.handle = handle

    @classmethod
    def from_obj(cls, path: str) -> "Mesh":
        h = C.c_void_p()
        call("vfs_mesh_load_obj", str(path).encode(), C.byref(h))
        return cls(h.value)

    @classmethod
    def from_arrays(cls, vertices, triangles) -> "Mesh":
        v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1)
        t = np.ascontiguousarray(triangles, dtype=np.uint32).reshape(-1)
        h = C.c_void_p()
        call("vfs_mesh_create", fptr(v), v.size // 3, uptr(t), t.size // 3, C.byref(h))
        return cls(h.value)

    def info(self):
        nv, nt = C.c_uint32(), C.c_uint32()
        lo, hi = np.empty(3, np.float32), np.empty(3, np.float32)
        call("vfs_mesh_info", self.handle, C.byref(nv), C.byref(nt), fptr(lo), fptr(hi))
        return nv.value, nt.value, lo, hi

    def properties(self):
        """(closed, acute): see vfs_mesh_properties."""
        c, a = C.c_int(), C.c_int()
        call("vfs_mesh_properties", self.handle, C.byref(c), C.byref(a))
        return bool(c.value), bool(a.value)

    def arrays(self):
        nv, nt, _, _ = self.info()
        v = np.empty(3 * nv, np.float32)
        t = np.empty(3 * nt, np.uint32)
        call("vfs_mesh_get", self.handle, fptr(v), uptr(t))
        return v.reshape(nv, 3), t.reshape(nt, 3)

    def signed_distance(self, points) -> np.ndarray:
        p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1)
        out = np.empty(p.size // 3, np.float64)
        call("vfs_mesh_signed_distance", self.handle, fptr(p), p.size // 3,
             out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def volume_map(self, resolution, support_radius: float, *, tolerance: float = 0.0,
                   margin: float | None = None) -> BoundaryGrid:
        if margin is None:
            margin = 8.0 * support_radius
        res = np.ascontiguousarray(np.asarray(resolution, dtype=np.uint32))
        cnt = int(np.prod(res.astype(np.int64)))
        sdf, vm = np.empty(cnt, np.float32), np.empty(cnt, np.float32)
        bs, bp = np.empty(3, np.float32), np.empty(3, np.float32)
        call("vfs_mesh_volume_map_generate", self.handle, uptr(res), support_radius, tolerance, margin,
             fptr(sdf), fptr(vm), fptr(bs), fptr(bp))
        return BoundaryGrid(sdf, vm, tuple(int(r) for r in res), tuple(float(v) for v in bs),
                            tuple(float(v) for v in bp))

    def close(self) -> None:
        if self.handle:
            call("vfs_mesh_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Sim:
    """WCSPHWithBoundaryModel through the C ABI (one scene, one GPU)."""

    def __init__(self, params: Params, objects: list[BoundaryGrid] | None = None):
        self.params = params
        self.n = int(params.n_particles)
        self._objects = list(objects or [])
        arr = (BoundaryObject * max(1, len(self._objects)))()
        for k, o in enumerate(self._objects):
            arr[k] = o.as_struct()
        h = C.c_void_p()
        call("vfs_sim_create", C.byref(params), arr if self._objects else None, len(self._objects), C.byref(h))
        self.handle = h.value

    # SetParticleState (model.cpp:139-147)
    def set_state(self, x: np.ndarray, v: np.ndarray | None = None, offset: int = 0, count: int = -1) -> None:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1)
        v = np.zeros_like(x) if v is None else np.ascontiguousarray(v, dtype=np.float32).reshape(-1)
        if count < 0:
            count = x.size // 3
        call("vfs_sim_set_state", self.handle, fptr(x), fptr(v), offset, count)

    def set_params(self, params: Params) -> None:
        call("vfs_sim_set_params", self.handle, C.byref(params))
        self.params = params

    def step(self, stream=None) -> None:
        call("vfs_sim_step", self.handle, stream)

    def substeps(self, k: int, stream=None) -> None:
        call("vfs_sim_substeps", self.handle, int(k), stream)

    def synchronize(self) -> None:
        call("vfs_sim_synchronize", self.handle)

    def get_state(self, ids: bool = False):
        n = self.n
        x = np.empty(3 * n, np.float32)
        v = np.empty(3 * n, np.float32)
        rho = np.empty(n, np.float32)
        pid = np.empty(n, np.uint32) if ids else None
        call("vfs_sim_get_state", self.handle, fptr(x), fptr(v), fptr(rho), uptr(pid) if ids else None)
        return x.reshape(n, 3), v.reshape(n, 3), rho, pid

    def get_hash(self):
        n = self.n
        k = np.empty(n, np.uint32)
        p = np.empty(n, np.uint32)
        o = np.empty(n, np.uint32)
        call("vfs_sim_get_hash", self.handle, uptr(k), uptr(p), uptr(o))
        return k, p, o

    def get_boundary(self):
        n = self.n
        psi = np.empty(n, np.float32)
        g = np.empty(3 * n, np.float32)
        call("vfs_sim_get_boundary", self.handle, fptr(psi), fptr(g))
        return psi, g.reshape(n, 3)

    def get_tuning(self) -> Tuning:
        t = Tuning()
        call("vfs_sim_get_tuning", self.handle, C.byref(t))
        return t

    def set_tuning(self, sort_bits=8, sort_items=16, block_size=128, stage_rows=1, cell_scale=1.0) -> None:
        t = Tuning(sort_bits, sort_items, block_size, stage_rows, cell_scale)
        call("vfs_sim_set_tuning", self.handle, C.byref(t))

    def autotune(self, substeps_per_trial: int = 5):
        """Fastest launch configuration on the current scene: (Tuning, ms per substep)."""
        t = Tuning()
        ms = C.c_float(0)
        call("vfs_sim_autotune", self.handle, int(substeps_per_trial), C.byref(t), C.byref(ms))
        return t, ms.value

    def device_buffers(self) -> _lib.DeviceBuffers:
        b = _lib.DeviceBuffers()
        call("vfs_sim_device_buffers", self.handle, C.byref(b))
        return b

    def get_accelerations(self) -> np.ndarray:
        """Accelerations of the last substep (n x 3; needs VFS_FLAG_STORE_ACCEL)."""
        b = self.device_buffers()
        if not b.accelerations:
            raise VfsError(-3, "get_accelerations", "the simulation was created without VFS_FLAG_STORE_ACCEL")
        call("vfs_sim_synchronize", self.handle)
        a = np.empty(4 * self.n, np.float32)
        if self.n:
            call("vfs_memcpy_d2h", a.ctypes.data, b.accelerations, a.nbytes)
        return a.reshape(self.n, 4)[:, :3].copy()

    def profile(self, k: int):
        """[(stage, ms per substep, launches per substep)] over k eager substeps."""
        m = 16
        ms = (C.c_float * m)()
        la = (C.c_uint32 * m)()
        names = (C.c_char_p * m)()
        cnt = C.c_uint32(0)
        call("vfs_sim_profile_stages", self.handle, int(k), ms, la, names, m, C.byref(cnt))
        return [(names[i].decode(), float(ms[i]), int(la[i])) for i in range(cnt.value)]

    def neighbour_stats(self):
        """(mean, max, overflowed) list entries per particle of the last substep."""
        mean, mx, over = C.c_double(0), C.c_uint32(0), C.c_uint32(0)
        call("vfs_sim_neighbour_stats", self.handle, C.byref(mean), C.byref(mx), C.byref(over))
        return mean.value, mx.value, over.value

    def check_state(self, margin: float) -> tuple:
        """(particles with a non-finite x or v, particles farther than margin outside the box)."""
        bad, out = C.c_uint32(0), C.c_uint32(0)
        call("vfs_sim_check_state", self.handle, float(margin), C.byref(bad), C.byref(out))
        return int(bad.value), int(out.value)

    def changed_keys(self) -> int:
        """Keys that changed bucket in the last substep (n after a full sort)."""
        c = C.c_uint32(0)
        call("vfs_sim_sort_stats", self.handle, C.byref(c))
        return int(c.value)

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib.vfs_sim_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class SpatialHash:
    """SpatialHash (src/compute/spatial_hash.h:9-34) on device positions."""

    def __init__(self, n: int, cell_size: float):
        h = C.c_void_p()
        call("vfs_spatial_hash_create", int(n), float(cell_size), C.byref(h))
        self.handle = h.value
        self.n = int(n)

    def run(self, positions_dev_ptr: int, stride_floats: int = 3, stream=None) -> None:
        call("vfs_spatial_hash_run", self.handle, positions_dev_ptr, stride_floats, stream)

    def results(self):
        k, i, o = C.c_void_p(), C.c_void_p(), C.c_void_p()
        call("vfs_spatial_hash_buffers", self.handle, C.byref(k), C.byref(i), C.byref(o))
        call("vfs_stream_synchronize", None)
        out = []
        for ptr in (k.value, i.value, o.value):
            a = np.empty(self.n, np.uint32)
            if self.n:
                call("vfs_memcpy_d2h", a.ctypes.data, ptr, a.nbytes)
            out.append(a)
        return tuple(out)

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib.vfs_spatial_hash_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sort_pairs(keys: np.ndarray, max_value: int | None = None):
    """GPUCountSort::Run on host data: returns (sorted_keys, permutation)."""
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    n = keys.size
    if max_value is None:
        max_value = int(keys.max()) if n else 0
    ws = C.c_size_t(0)
    call("vfs_sort_workspace_size", n, max_value, C.byref(ws))
    dk = DeviceArray.from_host(keys)
    di = DeviceArray(4 * n)
    dw = DeviceArray(ws.value)
    call("vfs_sort_pairs", di.ptr, dk.ptr, n, max_value, dw.ptr, ws.value, None)
    call("vfs_stream_synchronize", None)
    return dk.to_host(np.uint32, n), di.to_host(np.uint32, n)


def scan_exclusive(values: np.ndarray) -> np.ndarray:
    values = np.ascontiguousarray(values, dtype=np.uint32)
    n = values.size
    ws = C.c_size_t(0)
    call("vfs_scan_workspace_size", n, C.byref(ws))
    d = DeviceArray.from_host(values)
    w = DeviceArray(ws.value)
    call("vfs_scan_exclusive", d.ptr, n, w.ptr, ws.value, None)
    call("vfs_stream_synchronize", None)
    return d.to_host(np.uint32, n)


def spatial_offsets(sorted_keys: np.ndarray, init: bool = True, prefill: np.ndarray | None = None) -> np.ndarray:
    sorted_keys = np.ascontiguousarray(sorted_keys, dtype=np.uint32)
    n = sorted_keys.size
    dk = DeviceArray.from_host(sorted_keys)
    do = DeviceArray.from_host(prefill.astype(np.uint32)) if prefill is not None else DeviceArray(4 * n)
    call("vfs_spatial_offsets", dk.ptr, do.ptr, n, 1 if init else 0, None)
    call("vfs_stream_synchronize", None)
    return do.to_host(np.uint32, n)


def reorder_f32x3(perm: np.ndarray, src: np.ndarray) -> np.ndarray:
    perm = np.ascontiguousarray(perm, dtype=np.uint32)
    src = np.ascontiguousarray(src, dtype=np.float32).reshape(-1)
    n = perm.size
    dp = DeviceArray.from_host(perm)
    ds = DeviceArray.from_host(src)
    dd = DeviceArray(src.nbytes)
    call("vfs_reorder_f32x3", dp.ptr, ds.ptr, dd.ptr, n, None)
    call("vfs_stream_synchronize", None)
    return dd.to_host(np.float32, 3 * n).reshape(n, 3)


class Lague(Sim):
    """LagueModel (Clavet double-density) through the C ABI (vfs_lague_*); the
    vfs_sim methods of Sim apply."""

    def __init__(self, params: Params, lague: LagueParams | None = None):
        self.params = params
        self.n = int(params.n_particles)
        self._objects = []
        h = C.c_void_p()
        call("vfs_lague_create", C.byref(params), C.byref(lague) if lague is not None else None, C.byref(h))
        self.handle = h.value

    def set_lague_params(self, lague: LagueParams) -> None:
        call("vfs_lague_set_params", self.handle, C.byref(lague))

    def get_lague_params(self) -> LagueParams:
        lp = LagueParams()
        call("vfs_lague_get_params", self.handle, C.byref(lp))
        return lp


class Batch:
    """Independent scenes stepped together (vfs_batch_*): scene k has params[k]
    and the boundary grids objects[k]."""

    def __init__(self, params: list, objects: list | None = None):
        self.n_scenes = len(params)
        objects = objects or [[] for _ in params]
        assert len(objects) == self.n_scenes
        self.params = list(params)
        self.sizes = [int(p.n_particles) for p in params]
        self._objects = [list(o) for o in objects]
        flat = [g for objs in self._objects for g in objs]
        arr = (BoundaryObject * max(1, len(flat)))()
        for k, g in enumerate(flat):
            arr[k] = g.as_struct()
        pa = (Params * self.n_scenes)(*params)
        counts = np.array([len(o) for o in self._objects], np.uint32)
        h = C.c_void_p()
        call("vfs_batch_create", pa, self.n_scenes, arr if flat else None, uptr(counts), C.byref(h))
        self.handle = h.value

    def info(self):
        ns, npart, entries = C.c_uint32(0), C.c_uint64(0), C.c_uint32(0)
        call("vfs_batch_info", self.handle, C.byref(ns), C.byref(npart), C.byref(entries))
        return ns.value, npart.value, entries.value

    def set_state(self, scene: int, x: np.ndarray, v: np.ndarray | None = None) -> None:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1)
        v = np.zeros_like(x) if v is None else np.ascontiguousarray(v, dtype=np.float32).reshape(-1)
        call("vfs_batch_set_state", self.handle, int(scene), fptr(x), fptr(v))

    def substeps(self, k: int, stream=None) -> None:
        call("vfs_batch_substeps", self.handle, int(k), stream)

    def synchronize(self) -> None:
        call("vfs_batch_synchronize", self.handle)

    def get_state(self, scene: int, ids: bool = False):
        n = self.sizes[scene]
        x = np.empty(3 * n, np.float32)
        v = np.empty(3 * n, np.float32)
        rho = np.empty(n, np.float32)
        pid = np.empty(n, np.uint32) if ids else None
        call("vfs_batch_get_state", self.handle, int(scene), fptr(x), fptr(v), fptr(rho),
             uptr(pid) if ids else None)
        return x.reshape(n, 3), v.reshape(n, 3), rho, pid

    def get_hash(self, scene: int):
        n = self.sizes[scene]
        k, p, o = (np.empty(n, np.uint32) for _ in range(3))
        call("vfs_batch_get_hash", self.handle, int(scene), uptr(k), uptr(p), uptr(o))
        return k, p, o

    def get_boundary(self, scene: int):
        n = self.sizes[scene]
        psi = np.empty(n, np.float32)
        g = np.empty(3 * n, np.float32)
        call("vfs_batch_get_boundary", self.handle, int(scene), fptr(psi), fptr(g))
        return psi, g.reshape(n, 3)

    def get_tuning(self) -> Tuning:
        t = Tuning()
        call("vfs_batch_get_tuning", self.handle, C.byref(t))
        return t

    def set_tuning(self, sort_bits=8, sort_items=16, block_size=128, stage_rows=1, cell_scale=1.0) -> None:
        t = Tuning(sort_bits, sort_items, block_size, stage_rows, cell_scale)
        call("vfs_batch_set_tuning", self.handle, C.byref(t))

    def autotune(self, substeps_per_trial: int = 5, allow_cell_size: bool = False):
        t = Tuning()
        ms = C.c_float(0)
        call("vfs_batch_autotune", self.handle, int(substeps_per_trial), 1 if allow_cell_size else 0, C.byref(t),
             C.byref(ms))
        return t, ms.value

    def profile(self, k: int):
        m = 16
        ms = (C.c_float * m)()
        la = (C.c_uint32 * m)()
        names = (C.c_char_p * m)()
        cnt = C.c_uint32(0)
        call("vfs_batch_profile_stages", self.handle, int(k), ms, la, names, m, C.byref(cnt))
        return [(names[i].decode(), float(ms[i]), int(la[i])) for i in range(cnt.value)]

    def neighbour_stats(self):
        mean, mx, over = C.c_double(0), C.c_uint32(0), C.c_uint32(0)
        call("vfs_batch_neighbour_stats", self.handle, C.byref(mean), C.byref(mx), C.byref(over))
        return mean.value, mx.value, over.value

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib.vfs_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Adjoint:
    """Differentiable substeps of a Sim (vfs_adjoint_*): loss = mean squared
    difference between the density field sampled on a grid and a target."""

    def __init__(self, sim: "Sim", n_substeps: int, checkpoint_every: int, origin, spacing, dims):
        self.sim = sim
        self.n = sim.n
        g = _lib.DensityGrid()
        g.origin[:] = [float(v) for v in origin]
        g.spacing[:] = [float(v) for v in spacing]
        g.dims[:] = [int(v) for v in dims]
        self.grid = g
        self.G = int(np.prod([int(v) for v in dims]))
        h = C.c_void_p()
        call("vfs_adjoint_create", sim.handle, int(n_substeps), int(checkpoint_every), C.byref(g), C.byref(h))
        self.handle = h.value
        self._x0 = DeviceArray(12 * self.n)
        self._v0 = DeviceArray(12 * self.n)
        self._target = DeviceArray(4 * self.G)
        self._density = DeviceArray(4 * self.G)
        self._gx = DeviceArray(12 * self.n)
        self._gv = DeviceArray(12 * self.n)

    def forward(self, x0: np.ndarray, v0: np.ndarray, target: np.ndarray | None = None):
        """(loss, sampled density field) for the initial state (x0, v0)."""
        for dev, a in ((self._x0, x0), (self._v0, v0)):
            a = np.ascontiguousarray(a, np.float32).reshape(-1)
            call("vfs_memcpy_h2d", dev.ptr, a.ctypes.data, a.nbytes)
        tptr = None
        if target is not None:
            t = np.ascontiguousarray(target, np.float32).reshape(-1)
            call("vfs_memcpy_h2d", self._target.ptr, t.ctypes.data, t.nbytes)
            tptr = self._target.ptr
        loss = C.c_double(0)
        call("vfs_adjoint_forward", self.handle, self._x0.ptr, self._v0.ptr, tptr, self._density.ptr,
             C.byref(loss))
        return loss.value, self._density.to_host(np.float32, self.G)

    def backward(self):
        """(d loss / d x0, d loss / d v0, d loss / d viscosity)."""
        gnu = C.c_double(0)
        call("vfs_adjoint_backward", self.handle, self._gx.ptr, self._gv.ptr, C.byref(gnu))
        gx = self._gx.to_host(np.float32, 3 * self.n).reshape(self.n, 3)
        gv = self._gv.to_host(np.float32, 3 * self.n).reshape(self.n, 3)
        return gx, gv, gnu.value

    def profile_backward(self):
        """[(stage, total ms over the reverse pass, timed groups)] of one more
        reverse pass of the last forward (vfs_adjoint_profile_backward)."""
        m = 24
        ms = (C.c_float * m)()
        la = (C.c_uint32 * m)()
        names = (C.c_char_p * m)()
        cnt = C.c_uint32(0)
        call("vfs_adjoint_profile_backward", self.handle, ms, la, names, m, C.byref(cnt))
        return [(names[i].decode(), float(ms[i]), int(la[i])) for i in range(cnt.value)]

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib.vfs_adjoint_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """A JSON scene (LoadScene + GenericScene, src/util/scene_loader.cpp:113-140)."""

    def __init__(self, path: str, resource_root: str = ""):
        h = C.c_void_p()
        call("vfs_scene_load", path.encode(), resource_root.encode(), C.byref(h))
        self.handle = h.value
        s = C.c_void_p()
        call("vfs_scene_sim", self.handle, C.byref(s))
        self.sim_handle = s.value
        p = Params()
        call("vfs_sim_get_params", self.sim_handle, C.byref(p))
        self.params = p
        self.n = int(p.n_particles)

    def step(self, stream=None) -> None:
        call("vfs_scene_step", self.handle, stream)

    def reset(self) -> None:
        call("vfs_scene_reset", self.handle)

    def get_state(self, ids: bool = False):
        n = self.n
        x = np.empty(3 * n, np.float32)
        v = np.empty(3 * n, np.float32)
        rho = np.empty(n, np.float32)
        pid = np.empty(n, np.uint32) if ids else None
        call("vfs_sim_get_state", self.sim_handle, fptr(x), fptr(v), fptr(rho), uptr(pid) if ids else None)
        return x.reshape(n, 3), v.reshape(n, 3), rho, pid

    def boundary_objects(self) -> list:
        """The scene's boundary objects as BoundaryGrid copies (grids, grid box,
        world-to-object matrix and rotation, column-major float32)."""
        cnt = C.c_uint32(0)
        call("vfs_scene_boundary_objects", self.handle, None, 0, C.byref(cnt))
        if cnt.value == 0:
            return []
        arr = (BoundaryObject * cnt.value)()
        call("vfs_scene_boundary_objects", self.handle, arr, cnt.value, C.byref(cnt))
        out = []
        for o in arr:
            res = tuple(int(r) for r in o.resolution)
            m = res[0] * res[1] * res[2]
            sdf = np.ctypeslib.as_array(o.sdf_grid, (m,)).copy()
            vm = np.ctypeslib.as_array(o.volume_map_grid, (m,)).copy()
            out.append(BoundaryGrid(sdf, vm, res, tuple(o.box_size), tuple(o.box_pos),
                                    transform=np.array(o.transform, np.float32),
                                    rotation=np.array(o.rotation, np.float32)))
        return out

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib.vfs_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
