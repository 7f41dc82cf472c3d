"""Synthetic inputs of the benchmark workloads (host-side numpy).

Seeded with splitmix64 (seed 0x5EED0000 + workload index): the k-th uniform
of a stream is (mix(seed + (k+1) * golden) >> 40) * 2^-24; Gaussians use
Box-Muller on consecutive pairs.  Lattices are x-major with z innermost, as
the reference seeds them (src/models/model.cpp:52-63).
"""
from __future__ import annotations

import ctypes
import ctypes.util
from dataclasses import dataclass, field

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
SEED_BASE = 0x5EED0000


def splitmix64(seed: int, count: int, start: int = 0) -> np.ndarray:
    with np.errstate(over="ignore"):
        k = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, count: int, start: int = 0) -> np.ndarray:
    """Uniform float32 in [0, 1) with 24 random bits (exactly representable)."""
    return ((splitmix64(seed, count, start) >> np.uint64(40)).astype(np.float64) * 2.0 ** -24).astype(np.float32)


def normal(seed: int, count: int, start: int = 0) -> np.ndarray:
    m = (count + 1) // 2
    u = uniform(seed, 2 * m, start).astype(np.float64)
    u1 = np.maximum(u[0::2], 2.0 ** -25)
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
    out = np.empty(2 * m)
    out[0::2] = z[:m]
    out[1::2] = z[m:]
    return out[:count]


def lattice(counts, spacing: float, origin, jitter: float = 0.0, seed: int = 0) -> np.ndarray:
    """(N, 3) float32 block lattice, x outer, z inner; the arithmetic matches
    GenericScene::Reset with the spacing/jitter extension keys."""
    nx, ny, nz = (int(c) for c in counts)
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    idx = np.stack([i.reshape(-1), j.reshape(-1), k.reshape(-1)], axis=1).astype(np.float32)
    x = np.float32(spacing) * idx + np.asarray(origin, dtype=np.float32)[None, :]
    if jitter:
        u = uniform(seed, x.size).reshape(-1, 3)
        x = x + (np.float32(2.0) * u - np.float32(1.0)) * np.float32(jitter)
    return np.ascontiguousarray(x, dtype=np.float32)


@dataclass
class Workload:
    name: str
    n: int
    h: float
    dt: float
    steps: int
    box_size: tuple
    box_pos: tuple
    x: np.ndarray
    v: np.ndarray
    # analytic boundary shapes: (kind, params, resolution)
    shapes: list = field(default_factory=list)
    params_kw: dict = field(default_factory=dict)
    # substeps run from (x, v) on the GPU before the workload's measured state
    settle: int = 0


def icosphere(subdivisions: int, radius: float = 1.0, center=(0.0, 0.0, 0.0)):
    """Icosphere mesh: the icosahedron with every triangle split into four
    `subdivisions` times, midpoints pushed to the sphere (level 4: 2,562
    vertices, 5,120 triangles).  Triangles are oriented outward.  Returns
    (float32 vertices (V, 3), uint32 triangles (F, 3))."""
    g = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, g, 0), (1, g, 0), (-1, -g, 0), (1, -g, 0), (0, -1, g), (0, 1, g), (0, -1, -g), (0, 1, -g),
         (g, 0, -1), (g, 0, 1), (-g, 0, -1), (-g, 0, 1)]
    verts = [np.asarray(p, np.float64) / np.linalg.norm(p) for p in v]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdivisions):
        mid = {}

        def midpoint(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                p = verts[a] + verts[b]
                verts.append(p / np.linalg.norm(p))
                mid[key] = len(verts) - 1
            return mid[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.asarray(verts) * radius + np.asarray(center, np.float64)
    F = np.asarray(faces, np.int64)
    # outward orientation
    n = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    flip = np.einsum("ij,ij->i", n, V[F].mean(axis=1) - np.asarray(center)) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    return V.astype(np.float32), F.astype(np.uint32)


def box_mesh(center, half_extents):
    """Closed box of 12 outward triangles (two per face)."""
    c, e = np.asarray(center, np.float64), np.asarray(half_extents, np.float64)
    corners = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], np.float64)
    V = c + corners * e
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    F = []
    for a, b, cc, d in quads:
        F += [(a, b, cc), (a, cc, d)]
    F = np.asarray(F, np.int64)
    n = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    flip = np.einsum("ij,ij->i", n, V[F].mean(axis=1) - c) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    return V.astype(np.float32), F.astype(np.uint32)


def torus_mesh(R: float = 1.0, r: float = 0.4, nu: int = 24, nv: int = 12, center=(0.0, 0.0, 0.0)):
    """Closed torus around the z axis: (nu x nv) quads split into outward
    triangles (a non-convex test shape)."""
    u = np.arange(nu) * 2 * np.pi / nu
    w = np.arange(nv) * 2 * np.pi / nv
    uu, ww = np.meshgrid(u, w, indexing="ij")
    pts = np.stack([(R + r * np.cos(ww)) * np.cos(uu), (R + r * np.cos(ww)) * np.sin(uu), r * np.sin(ww)], -1)
    V = pts.reshape(-1, 3)
    F = []
    for i in range(nu):
        for j in range(nv):
            a, b = i * nv + j, ((i + 1) % nu) * nv + j
            c, d = ((i + 1) % nu) * nv + (j + 1) % nv, i * nv + (j + 1) % nv
            F += [(a, b, c), (a, c, d)]
    F = np.asarray(F, np.int64)
    cen = V[F].mean(axis=1)
    ring = cen.copy()
    ring[:, 2] = 0
    ring[:, :2] *= (R / np.linalg.norm(ring[:, :2], axis=1))[:, None]
    n = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    flip = np.einsum("ij,ij->i", n, cen - ring) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    return (V + np.asarray(center, np.float64)).astype(np.float32), F.astype(np.uint32)


def dam_break(side=(32, 32, 32), spacing: float = 0.02, jitter: float = 0.001, seed: int = SEED_BASE,
              steps: int = 100, map_res: int = 64, name: str = "wcsph_dam_break") -> Workload:
    """Plain WCSPH dam break: lattice block with spacing 0.02 in a 1.6 x 0.8 x 0.8
    box, h = 2 spacing, dt = 2e-4, box walls as an inverted-box volume map."""
    h = 2.0 * spacing
    box = (1.6, 0.8, 0.8)
    x = lattice(side, spacing, (spacing, spacing, 0.08), jitter, seed)
    v = np.zeros_like(x)
    walls = (2, (0.8, 0.4, 0.4, 0.8, 0.4, 0.4), (map_res,) * 3)
    return Workload(name, x.shape[0], h, 2e-4, steps, box, (0.0, 0.0, 0.0), x, v, [walls])


def obstacle_1m(spacing: float = 0.006, map_res: int = 128, side=(128, 64, 128), steps: int = 500,
                walls: bool = False, mesh: bool = True) -> Workload:
    """1,048,576-particle lattice (128 x 64 x 128) in a 2 x 1 x 1 box with a centred
    obstacle of radius 0.2 (128^3 SDF/volume map), dt = 1e-4, box walls as the
    penalty force.  The obstacle is an icosphere mesh of subdivision 4 (2,562
    vertices, 5,120 triangles), or with mesh=False the analytic sphere.  The
    spacing is 0.006 (h = 0.012) so that the block fits the box next to the
    sphere.

    The lattice starts at 0.8 rho0, so it falls freely (zero pressure) for about
    320 substeps and then hits the penalty floor; with the model's 1/h-scaled
    kernel gradient the impact blows the scene up.  A resolved box-wall volume
    map (walls=True, node spacing <= h/2) does not help: particles seeded h/2
    from two walls get psi ~ 1 and are kicked out at once.  Timings are taken in
    the first ~250 substeps."""
    h = 2.0 * spacing
    x = lattice(side, spacing, (spacing, spacing, spacing), 0.1 * spacing, SEED_BASE + 1)
    v = np.zeros_like(x)
    if mesh:
        shapes = [("mesh", icosphere(4, 0.2, (1.0, 0.5, 0.5)), (map_res,) * 3)]
    else:
        shapes = [(1, (1.0, 0.5, 0.5, 0.2), (map_res,) * 3)]
    if walls:
        res = tuple(int(np.ceil((e + 16.0 * h) / (0.5 * h))) + 1 for e in (2.0, 1.0, 1.0))
        shapes.append((2, (1.0, 0.5, 0.5, 1.0, 0.5, 0.5), res))
    return Workload("wcsph_obstacle_1m", x.shape[0], h, 1e-4, steps, (2.0, 1.0, 1.0), (0.0, 0.0, 0.0), x, v,
                    shapes)


def obstacle_1m_contact(settle: int = 1500) -> Workload:
    """The 1,048,576-particle obstacle workload in its contact phase: the same
    lattice, box, obstacle, h = 0.012 and dt = 1e-4, with viscosity 1e-4, run
    `settle` substeps from rest so that the fluid has hit the floor (at 1,500
    substeps two thirds of the particles are above rest density, ~6,000 keys
    change bucket per substep, 36 listed neighbours per particle).

    Viscosity 0.01 is unstable here at any practical dt: with the reference's
    kernel gradient (an extra 1/h) the viscous term's rate grows like
    nu / h^3, about 1.6e6 /s at h = 0.012, so the explicit update diverges at
    the first contact for dt = 1e-4 and dt = 1e-5 alike (and for nu = 1e-3 at
    dt = 1e-4); nu = 1e-4 is stable through 3,000 substeps.  The viscous sum
    is evaluated in every substep whatever nu is."""
    w = obstacle_1m()
    return Workload("wcsph_obstacle_1m_contact", w.n, w.h, w.dt, w.steps, w.box_size, w.box_pos, w.x, w.v,
                    w.shapes, {"viscosity": 1e-4}, settle)


def _cbrtf(x: np.float32) -> np.float32:
    """std::cbrt of a float as the reference's host code computes it (the C
    library's cbrtf; numpy's float32 cbrt can differ in the last bit)."""
    try:
        libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
        libm.cbrtf.restype = ctypes.c_float
        libm.cbrtf.argtypes = [ctypes.c_float]
        return np.float32(libm.cbrtf(float(x)))
    except (OSError, AttributeError):
        return np.float32(np.cbrt(np.float64(x)))


def compact_block(counts, density: float, pos=(0.0, 0.0, 0.0)) -> np.ndarray:
    """SpawnCompactParticlesInBox (model.cpp:36-63) for the box a scene sizes
    from `counts` (dam_break_scene.cpp:42-50): spacing cbrt(1 / density) in
    float32, x-major with z innermost, at the box position."""
    step = _cbrtf(np.float32(1.0) / np.float32(density))
    size = step * (np.asarray(counts, np.float32) + np.float32(1.0))
    n_axis = [int(np.round(size[a] / step)) - 1 for a in range(3)]
    i, j, k = np.meshgrid(*[np.arange(m, dtype=np.float32) for m in n_axis], indexing="ij")
    pts = np.stack([step * i.ravel(), step * j.ravel(), step * k.ravel()], 1) + np.asarray(pos, np.float32)
    return pts[: int(np.prod(counts))].astype(np.float32)


def lague_dam_break(counts=(50, 60, 50), steps: int = 100) -> Workload:
    """The reference's Lague dam break (dam_break_scene.cpp:8-50): a 50 x 60 x 50
    block (150,000 particles) at target density 630 in a 23 x 10 x 10 box,
    dt = 1/120 over 3 iterations, smooth radius 0.2, the block centred in z."""
    h, rho0 = 0.2, 630.0
    box = (23.0, 10.0, 10.0)
    step = _cbrtf(np.float32(1.0) / np.float32(rho0))
    size_z = step * np.float32(counts[2] + 1)
    pos = (0.0, 0.0, float(np.float32(0.0) + (np.float32(box[2]) - size_z) / np.float32(2.0)))
    x = compact_block(counts, rho0, pos)
    v = np.zeros_like(x)
    n = x.shape[0]
    return Workload("lague_dam_break", n, h, 1.0 / 120.0, steps, box, (0.0, 0.0, 0.0), x, v, [],
                    {"rho0": rho0, "iterations": 3})


def uniform_stress(n: int = 8_388_608, h: float = 0.02, map_res: int = 256, steps: int = 200) -> Workload:
    """n particles uniform in a 4 x 2 x 2 box with N(0, 0.5) velocities, box
    walls, dt = 1e-4, viscosity 1e-4 (0.01 diverges, as in the 1M obstacle
    scene).  Random placement puts pairs far closer than the lattice spacing;
    their pressure, with the reference's 1/h-scaled gradient, throws them apart
    at up to ~6,600 m/s in the first substep, and ~0.2% of the particles end
    outside the box, but the state stays finite through the 200 substeps, each
    re-sorting 150,000-290,000 changed keys."""
    seed = SEED_BASE + 2
    u = uniform(seed, 3 * n).reshape(n, 3)
    x = (u * np.array([4.0, 2.0, 2.0], dtype=np.float32)).astype(np.float32)
    v = (0.5 * normal(seed ^ 0xABCDEF, 3 * n)).astype(np.float32).reshape(n, 3)
    walls = (2, (2.0, 1.0, 1.0, 2.0, 1.0, 1.0), (map_res,) * 3)
    return Workload("wcsph_uniform_stress", n, h, 1e-4, steps, (4.0, 2.0, 2.0), (0.0, 0.0, 0.0), x, v, [walls],
                    {"viscosity": 1e-4})


def batched_dam_breaks(n_scenes: int = 256, seed: int = SEED_BASE + 3, map_res: int = 32, steps: int = 200,
                       n_min: int = 2048, n_max: int = 65536) -> list:
    """Independent dam breaks of uneven sizes: n_s = round(n_min (n_max/n_min)^u)
    (u uniform), box dimensions uniform in [0.5, 2] per axis, spacing
    d_s = cbrt(V_box / (3 n_s)) shrunk by 1% steps until the left third of the
    box holds n_s lattice points, h_s = 2 d_s, jitter 0.05 d_s, dt = 2e-4, box
    walls as a map_res^3 inverted-box volume map.  Returns a list of Workload."""
    u = uniform(seed, 4 * n_scenes).astype(np.float64).reshape(n_scenes, 4)
    out = []
    for s in range(n_scenes):
        n = int(round(n_min * (n_max / n_min) ** u[s, 0]))
        box = tuple(float(np.float32(0.5 + 1.5 * u[s, a])) for a in (1, 2, 3))
        d = float(np.cbrt(box[0] * box[1] * box[2] / (3.0 * n)))
        while True:
            cx = int((box[0] / 3.0) / d)
            cy = int((box[1] - 2.0 * d) / d)
            cz = int((box[2] - 2.0 * d) / d)
            if cx * cy * cz >= n:
                break
            d *= 0.99
        x = lattice((cx, cy, cz), d, (d, d, d), 0.05 * d, seed + 1 + s)[:n]
        h = 2.0 * d
        walls = (2, (box[0] / 2, box[1] / 2, box[2] / 2, box[0] / 2, box[1] / 2, box[2] / 2), (map_res,) * 3)
        out.append(Workload(f"dam_break_batch_{s}", n, h, 2e-4, steps, box, (0.0, 0.0, 0.0),
                            np.ascontiguousarray(x), np.zeros_like(x), [walls]))
    return out


WORKLOADS = {
    "dam_break_32k": lambda: dam_break(),
    "obstacle_1m": lambda: obstacle_1m(),
    "obstacle_1m_contact": lambda: obstacle_1m_contact(),
    "lague_dam_break": lambda: lague_dam_break(),
    "obstacle_1m_wall_map": lambda: obstacle_1m(walls=True),
    "uniform_8m": lambda: uniform_stress(),
    "dam_break_65k": lambda: dam_break(side=(64, 32, 32), seed=SEED_BASE + 4, steps=200,
                                       name="wcsph_dam_break_65k"),
}
