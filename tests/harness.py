"""Test harness: ctypes binding of the CPU oracle (oracle/liboracle.so) and
helpers shared by the CPU and GPU tests.  The oracle is the checker only."""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_SO = ROOT / "oracle" / "liboracle.so"


def _load_oracle() -> C.CDLL:
    if not ORACLE_SO.exists():
        subprocess.run(["make", "-C", str(ROOT / "oracle")], check=True, capture_output=True)
    L = C.CDLL(str(ORACLE_SO))
    vp, u32, fp, up, dp = C.c_void_p, C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_uint32), C.POINTER(C.c_double)
    sig = {
        "orc_hash_cell": (u32, [C.c_int32] * 3),
        "orc_hash_keys": (None, [fp, u32, C.c_float, u32, up]),
        "orc_stable_sort": (None, [up, u32, u32, up, up]),
        "orc_offsets": (None, [up, u32, up]),
        "orc_exclusive_scan": (None, [up, u32, up]),
        "orc_W_f32": (C.c_float, [vp, C.c_float]),
        "orc_dW_f32": (C.c_float, [vp, C.c_float]),
        "orc_W_f64": (C.c_double, [vp, C.c_double]),
        "orc_dW_f64": (C.c_double, [vp, C.c_double]),
        "orc_eos_f32": (C.c_float, [vp, C.c_float]),
        "orc_uniforms_f32": (None, [vp, fp]),
        "orc_sample_grid_f32": (C.c_float, [fp, up, fp]),
        "orc_sample_grid_f64": (C.c_double, [fp, up, dp]),
        "orc_sim_create": (vp, [vp, vp, u32, C.c_int]),
        "orc_sim_destroy": (None, [vp]),
        "orc_sim_gradient_vector": (C.c_int, [vp, u32, dp, dp]),
        "orc_sim_set_threads": (None, [vp, C.c_int]),
        "orc_sim_set_state": (None, [vp, fp, fp, u32, u32]),
        "orc_sim_set_state_f64": (None, [vp, dp, dp, u32, u32]),
        "orc_adjoint_run": (C.c_int, [vp, u32, vp, dp, dp, dp, dp, dp, dp]),
        "orc_adjoint_run_f32_trajectory": (C.c_int, [vp, vp, u32, vp, dp, dp, dp, dp, dp]),
        "orc_sim_substeps": (None, [vp, u32]),
        "orc_sim_get_state": (None, [vp, dp, dp, dp, up, dp]),
        "orc_sim_get_state_f32": (None, [vp, fp, fp, fp]),
        "orc_sim_get_hash": (None, [vp, up, up, up]),
        "orc_sim_get_boundary": (None, [vp, dp, dp]),
        "orc_gauss_legendre": (None, [C.c_int, dp, dp]),
        "orc_volume_map": (C.c_int, [C.c_int, dp, up, C.c_double, C.c_double, C.c_double, fp, fp, fp, fp, C.c_int]),
        "orc_max_threads": (C.c_int, []),
        "orc_lague_create": (vp, [vp, vp, C.c_int, C.c_int]),
        "orc_lague_destroy": (None, [vp]),
        "orc_lague_set_state": (None, [vp, fp, fp]),
        "orc_lague_substeps": (None, [vp, u32]),
        "orc_lague_get_state": (None, [vp, dp, dp, dp, dp, up]),
        "orc_lague_kernels": (None, [vp, C.c_double, dp]),
        "orc_mesh_pseudonormals": (C.c_int, [fp, u32, up, u32, fp, fp, fp]),
        "orc_point_triangle": (C.c_int, [dp, dp, u32, dp, dp, C.POINTER(C.c_int)]),
        "orc_mesh_signed_distance": (C.c_int, [fp, u32, up, u32, fp, u32, dp, C.c_int]),
        "orc_mesh_volume_map": (C.c_int, [fp, u32, up, u32, up, C.c_double, C.c_double, C.c_double, fp, fp, fp, fp,
                                          C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


ORC = _load_oracle()


def fp(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


def up(a):
    assert a.dtype == np.uint32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def threads() -> int:
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except Exception:
        return 4


class OracleSim:
    """CPU restatement of one scene (FP32 parity oracle, or the FP64 twin)."""

    def __init__(self, params, grids=(), fp64: bool = False, nthreads: int | None = None):
        from vkfluidsim_amd._lib import BoundaryObject
        self.n = int(params.n_particles)
        self._grids = list(grids)
        arr = (BoundaryObject * max(1, len(self._grids)))()
        for k, g in enumerate(self._grids):
            arr[k] = g.as_struct()
        self.h = ORC.orc_sim_create(C.addressof(params), C.addressof(arr), len(self._grids), 1 if fp64 else 0)
        assert self.h
        ORC.orc_sim_set_threads(self.h, nthreads or threads())

    def set_state(self, x, v=None):
        x = np.ascontiguousarray(x, np.float32).reshape(-1)
        v = np.zeros_like(x) if v is None else np.ascontiguousarray(v, np.float32).reshape(-1)
        ORC.orc_sim_set_state(self.h, fp(x), fp(v), 0, x.size // 3)

    def set_state_f64(self, x, v=None):
        x = np.ascontiguousarray(x, np.float64).reshape(-1)
        v = np.zeros_like(x) if v is None else np.ascontiguousarray(v, np.float64).reshape(-1)
        ORC.orc_sim_set_state_f64(self.h, dp(x), dp(v), 0, x.size // 3)

    def substeps(self, k: int):
        ORC.orc_sim_substeps(self.h, int(k))

    def get_state(self):
        n = self.n
        x = np.empty(3 * n)
        v = np.empty(3 * n)
        rho = np.empty(n)
        pid = np.empty(n, np.uint32)
        a = np.empty(3 * n)
        ORC.orc_sim_get_state(self.h, dp(x), dp(v), dp(rho), up(pid), dp(a))
        return x.reshape(n, 3), v.reshape(n, 3), rho, pid, a.reshape(n, 3)

    def get_hash(self):
        n = self.n
        k = np.empty(n, np.uint32)
        p = np.empty(n, np.uint32)
        o = np.empty(n, np.uint32)
        ORC.orc_sim_get_hash(self.h, up(k), up(p), up(o))
        return k, p, o

    def gradient_vector(self, k, x_local):
        x = np.ascontiguousarray(x_local, np.float64)
        out = np.empty(3)
        assert ORC.orc_sim_gradient_vector(self.h, int(k), dp(x), dp(out)) == 0
        return out

    def get_boundary(self):
        n = self.n
        psi = np.empty(n)
        g = np.empty(3 * n)
        ORC.orc_sim_get_boundary(self.h, dp(psi), dp(g))
        return psi, g.reshape(n, 3)

    def __del__(self):
        if getattr(self, "h", None):
            ORC.orc_sim_destroy(self.h)
            self.h = None


def oracle_volume_map(kind, shape_params, resolution, support_radius, tolerance=0.0, margin=None, nthreads=None):
    from vkfluidsim_amd import BoundaryGrid
    if margin is None:
        margin = 8.0 * support_radius
    res = np.ascontiguousarray(np.asarray(resolution, np.uint32))
    sp = np.zeros(6)
    sp[: len(shape_params)] = shape_params
    cnt = int(np.prod(res.astype(np.int64)))
    sdf = np.empty(cnt, np.float32)
    vm = np.empty(cnt, np.float32)
    bs = np.empty(3, np.float32)
    bpos = np.empty(3, np.float32)
    rc = ORC.orc_volume_map(kind, dp(sp), up(res), support_radius, tolerance, margin, fp(sdf), fp(vm), fp(bs),
                            fp(bpos), nthreads or threads())
    assert rc == 0
    return BoundaryGrid(sdf, vm, tuple(int(r) for r in res), tuple(float(v) for v in bs),
                        tuple(float(v) for v in bpos))


class OracleLague:
    """The oracle's LagueModel restatement (FP32, or the FP64 twin)."""

    def __init__(self, params, lague, fp64=False, nthreads=None):
        self.n = int(params.n_particles)
        self._p, self._l = params, lague
        self.h = ORC.orc_lague_create(C.byref(params), C.byref(lague), int(fp64), nthreads or threads())
        assert self.h

    def set_state(self, x, v=None):
        x = np.ascontiguousarray(x, np.float32).reshape(-1)
        v = np.zeros_like(x) if v is None else np.ascontiguousarray(v, np.float32).reshape(-1)
        ORC.orc_lague_set_state(self.h, fp(x), fp(v))

    def substeps(self, k):
        ORC.orc_lague_substeps(self.h, k)

    def get_state(self):
        n = self.n
        x, v = np.empty(3 * n), np.empty(3 * n)
        rho, near = np.empty(n), np.empty(n)
        pid = np.empty(n, np.uint32)
        ORC.orc_lague_get_state(self.h, dp(x), dp(v), dp(rho), dp(near), up(pid))
        return x.reshape(n, 3), v.reshape(n, 3), rho, near, pid

    def kernels(self, d):
        out = np.empty(5)
        ORC.orc_lague_kernels(self.h, float(d), dp(out))
        return out

    def __del__(self):
        if getattr(self, "h", None):
            ORC.orc_lague_destroy(self.h)
            self.h = None


def _mesh_arrays(vertices, triangles):
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1)
    t = np.ascontiguousarray(triangles, np.uint32).reshape(-1)
    return v, t


def oracle_mesh_signed_distance(vertices, triangles, points, nthreads=None):
    """Exhaustive-search signed distance of the oracle (float query points)."""
    v, t = _mesh_arrays(vertices, triangles)
    p = np.ascontiguousarray(points, np.float32).reshape(-1)
    out = np.empty(p.size // 3)
    assert ORC.orc_mesh_signed_distance(fp(v), v.size // 3, up(t), t.size // 3, fp(p), p.size // 3, dp(out),
                                        nthreads or threads()) == 0
    return out


def oracle_mesh_volume_map(vertices, triangles, resolution, support_radius, tolerance=0.0, margin=None,
                           nthreads=None):
    from vkfluidsim_amd import BoundaryGrid
    if margin is None:
        margin = 8.0 * support_radius
    v, t = _mesh_arrays(vertices, triangles)
    res = np.ascontiguousarray(np.asarray(resolution, np.uint32))
    cnt = int(np.prod(res.astype(np.int64)))
    sdf, vm = np.empty(cnt, np.float32), np.empty(cnt, np.float32)
    bs, bpos = np.empty(3, np.float32), np.empty(3, np.float32)
    assert ORC.orc_mesh_volume_map(fp(v), v.size // 3, up(t), t.size // 3, up(res), support_radius, tolerance, margin,
                                   fp(sdf), fp(vm), fp(bs), fp(bpos), nthreads or threads()) == 0
    return BoundaryGrid(sdf, vm, tuple(int(r) for r in res), tuple(float(x) for x in bs),
                        tuple(float(x) for x in bpos))


def oracle_pseudonormals(vertices, triangles):
    v, t = _mesh_arrays(vertices, triangles)
    nt, nv = t.size // 3, v.size // 3
    face, edge, vert = np.empty(3 * nt, np.float32), np.empty(9 * nt, np.float32), np.empty(3 * nv, np.float32)
    assert ORC.orc_mesh_pseudonormals(fp(v), nv, up(t), nt, fp(face), fp(edge), fp(vert)) == 0
    return face.reshape(nt, 3), edge.reshape(nt, 3, 3), vert.reshape(nv, 3)


def oracle_point_triangle(points, tris9):
    p = np.ascontiguousarray(points, np.float64).reshape(-1)
    t = np.ascontiguousarray(tris9, np.float64).reshape(-1)
    n = p.size // 3
    d2, pt, feat = np.empty(n), np.empty(3 * n), np.empty(n, np.int32)
    assert ORC.orc_point_triangle(dp(p), dp(t), n, dp(d2), dp(pt), feat.ctypes.data_as(C.POINTER(C.c_int))) == 0
    return d2, pt.reshape(n, 3), feat


def write_obj(path, vertices, faces, extra=""):
    """OBJ text: `v` lines, then `f` lines of 1-based indices (faces may be
    polygons), with `extra` appended verbatim."""
    lines = [f"v {x:.9g} {y:.9g} {z:.9g}" for x, y, z in np.asarray(vertices, np.float64)]
    lines += ["f " + " ".join(str(int(i) + 1) for i in f) for f in faces]
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n" + extra)


def hash_cells(c):
    """HashCell3D of integer cells (m, 3) in numpy (mod 2^32), for brute-force
    neighbour counts; the oracle's own hash is pinned by the known answers."""
    u = (np.asarray(c, np.int64) + 25) & 0xFFFFFFFF
    lo, hi = u % 50, u // 50
    h = lo[:, 0] + 50 * (lo[:, 1] + 50 * lo[:, 2]) + hi[:, 0] * 15823 + hi[:, 1] * 9737333 + hi[:, 2] * 440817757
    return h & 0xFFFFFFFF


class GridC(C.Structure):
    _fields_ = [("origin", C.c_float * 3), ("spacing", C.c_float * 3), ("dims", C.c_uint32 * 3)]


def oracle_adjoint(params, grids, x0, v0, steps, origin, spacing, dims, target=None, nthreads=None):
    """FP64 host adjoint: (loss, density field, d loss/d x0, d loss/d v0, d loss/d viscosity)."""
    o = OracleSim(params, grids, fp64=True, nthreads=nthreads)
    o.set_state_f64(x0, v0)
    g = GridC()
    g.origin[:] = [float(v) for v in origin]
    g.spacing[:] = [float(v) for v in spacing]
    g.dims[:] = [int(v) for v in dims]
    n = int(params.n_particles)
    G = int(np.prod([int(v) for v in dims]))
    t = None if target is None else np.ascontiguousarray(target, np.float64)
    loss = np.zeros(1)
    rho = np.empty(G)
    gx = np.empty(3 * n)
    gv = np.empty(3 * n)
    gnu = np.zeros(1)
    rc = ORC.orc_adjoint_run(o.h, int(steps), C.addressof(g), dp(t) if t is not None else None, dp(loss), dp(gx),
                             dp(gv), dp(gnu), dp(rho))
    assert rc == 0
    return float(loss[0]), rho, gx.reshape(n, 3), gv.reshape(n, 3), float(gnu[0])


def oracle_adjoint_f32_trajectory(params, grids, x0, v0, steps, origin, spacing, dims, target, nthreads=None):
    """FP64 adjoint of the FP32 oracle trajectory (the GPU's, bit for bit):
    (loss, d loss/d x0, d loss/d v0, d loss/d viscosity)."""
    o32 = OracleSim(params, grids, nthreads=nthreads)
    o32.set_state(x0, v0)
    o64 = OracleSim(params, grids, fp64=True, nthreads=nthreads)
    g = GridC()
    g.origin[:] = [float(v) for v in origin]
    g.spacing[:] = [float(v) for v in spacing]
    g.dims[:] = [int(v) for v in dims]
    n = int(params.n_particles)
    t = np.ascontiguousarray(target, np.float64)
    loss, gnu = np.zeros(1), np.zeros(1)
    gx, gv = np.empty(3 * n), np.empty(3 * n)
    rc = ORC.orc_adjoint_run_f32_trajectory(o32.h, o64.h, int(steps), C.addressof(g), dp(t), dp(loss), dp(gx), dp(gv),
                                            dp(gnu))
    assert rc == 0
    return float(loss[0]), gx.reshape(n, 3), gv.reshape(n, 3), float(gnu[0])


def oracle_keys(x, cell, table):
    x = np.ascontiguousarray(x, np.float32).reshape(-1)
    n = x.size // 3
    keys = np.empty(n, np.uint32)
    ORC.orc_hash_keys(fp(x), n, cell, table, up(keys))
    return keys


def oracle_sort(keys, max_value):
    keys = np.ascontiguousarray(keys, np.uint32)
    n = keys.size
    sk = np.empty(n, np.uint32)
    perm = np.empty(n, np.uint32)
    ORC.orc_stable_sort(up(keys), n, max_value, up(sk), up(perm))
    return sk, perm


def oracle_offsets(sorted_keys):
    sorted_keys = np.ascontiguousarray(sorted_keys, np.uint32)
    n = sorted_keys.size
    o = np.empty(n, np.uint32)
    ORC.orc_offsets(up(sorted_keys), n, up(o))
    return o


def rel_err(a, b) -> float:
    """max |a - b| over the field, relative to max |b| (the field scale)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(float(np.max(np.abs(b))) if b.size else 0.0, 1e-30)
    return float(np.max(np.abs(a - b))) / scale if a.size else 0.0


def by_id(arr, pid):
    out = np.empty_like(arr)
    out[pid] = arr
    return out
