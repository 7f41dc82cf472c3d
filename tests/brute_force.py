"""Independent float64 numpy evaluations of the WCSPH terms, written from the
shader formulas (not from the oracle or the kernels): O(N^2) pair sums for
density, pressure and viscosity, and the closed form of the boundary volume
of a flat wall.  Used to pin the CPU oracle and, directly, the GPU kernels.

Reference formulas (luihabl/VkFluidSim):
  CubicSpline / GradCubicSpline      assets/shaders/simulation/kernels/kernels_3d.slang:62-91
  CalculateDensities                 assets/shaders/simulation/wcsph_with_boundary_model.slang:153-197
  PressureFromDensityEOS             wcsph_with_boundary_model.slang:199-203
  PressureAccel                      wcsph_with_boundary_model.slang:205-262
  ViscousAccel                       wcsph_with_boundary_model.slang:273-324
  CalculateBoundaryVolume            wcsph_with_boundary_model.slang:102-151
  ParticleVolume (mass / rho0)       wcsph_with_boundary_model.slang:39-41
The neighbour loops visit the bucket of each of the 27 stencil cells' keys
(hash % n), so a pair whose bucket key appears k times among i's 27 stencil
keys is summed k times; the sums below weight each pair by that count.
"""
from __future__ import annotations

import math

import numpy as np

import harness as H


def w_cubic(r, h):
    """Cubic spline W(r), sigma = 8 / (pi h^3) (textbook form)."""
    r = np.asarray(r, np.float64)
    q = r / h
    s = 8.0 / (math.pi * h ** 3)
    inner = s * (6 * q ** 3 - 6 * q ** 2 + 1)
    outer = 2 * s * (1 - q) ** 3
    return np.where(q <= 0.5, inner, np.where(q <= 1.0, outer, 0.0))


def dw_shader(r, h):
    """The shader's GradCubicSpline: (1/h) dW/dr for r > 1e-9 and q <= 1."""
    r = np.asarray(r, np.float64)
    q = r / h
    s = 8.0 / (math.pi * h ** 3)
    dwdr = np.where(q <= 0.5, 6 * s / h * q * (3 * q - 2), -6 * s / h * (1 - q) ** 2)
    return np.where((r > 1e-9) & (q <= 1.0), dwdr / h, 0.0)


def eos(rho, rho0, B, gamma):
    return B * ((np.maximum(rho, rho0) / rho0) ** gamma - 1.0)


# The shaders' literals are float: 0.8 and 0.01 enter as their float32 values.
F32_0_8 = float(np.float32(0.8))
F32_0_01 = float(np.float32(0.01))


def particle_mass(h, rho0):
    return rho0 * F32_0_8 * h ** 3 / 8.0


def cells_of(x32, cell):
    """GetCell3D in float32 (true division, then floor), as the shader does."""
    return np.floor(np.asarray(x32, np.float32) / np.float32(cell)).astype(np.int64)


def pairs_within(x, h, include_self):
    """(i, j, xij = x_i - x_j, r) for every ordered pair with r <= h."""
    d = x[:, None, :] - x[None, :, :]
    r2 = np.einsum("ijk,ijk->ij", d, d)
    mask = r2 <= h * h
    if not include_self:
        np.fill_diagonal(mask, False)
    i, j = np.nonzero(mask)
    return i, j, d[i, j], np.sqrt(r2[i, j])


def multiplicity(x32, cell, n, i, j):
    """How often the bucket of particle j is visited from particle i: the
    number of i's 27 stencil cells whose key (hash % n) equals j's key."""
    c = cells_of(x32, cell)
    off = np.array([(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)], np.int64)
    stencil = (c[:, None, :] + off[None, :, :]).reshape(-1, 3)
    sk = (H.hash_cells(stencil) % n).reshape(-1, 27)
    key = H.hash_cells(c) % n
    return np.sum(sk[i] == key[j][:, None], axis=1).astype(np.float64)


def density(x32, h, rho0, psi=None):
    """rho_i = sum_j k_ij m W(r_ij) (self included) + rho0 psi_i."""
    x = np.asarray(x32, np.float64)
    n = x.shape[0]
    i, j, _, r = pairs_within(x, h, include_self=True)
    k = multiplicity(x32, h, n, i, j)
    rho = np.zeros(n)
    np.add.at(rho, i, k * particle_mass(h, rho0) * w_cubic(r, h))
    if psi is not None:
        rho = rho + rho0 * np.asarray(psi, np.float64)
    return rho


def pressure_accel(x32, rho, h, rho0, B, gamma, grad_psi=None):
    """-[sum_{j != i} k_ij m (p_i/rho_i^2 + p_j/rho_j^2) W'(r) x_ij/r
         + (p_i/rho_i^2 + p_i/rho0^2) rho0 grad psi_i]."""
    x = np.asarray(x32, np.float64)
    rho = np.asarray(rho, np.float64)
    n = x.shape[0]
    p = eos(rho, rho0, B, gamma)
    pd = np.where(rho > 0, p / (rho * rho), 0.0)
    i, j, xij, r = pairs_within(x, h, include_self=False)
    k = multiplicity(x32, h, n, i, j)
    xn = np.where(r[:, None] > 0, xij / np.where(r > 0, r, 1.0)[:, None], np.array([0.0, 1.0, 0.0]))
    s = k * particle_mass(h, rho0) * (pd[i] + pd[j]) * dw_shader(r, h)
    f = np.zeros((n, 3))
    np.add.at(f, i, s[:, None] * xn)
    if grad_psi is not None:
        pb = np.where(rho > 0, p / (rho0 * rho0), 0.0)
        f += ((pd + pb) * rho0)[:, None] * np.asarray(grad_psi, np.float64)
    return -f


def viscous_accel(x32, v32, rho, h, rho0, nu):
    """2 (d + 2) nu sum_{j != i} k_ij (m / max(rho_j, 1e-6)) (v_ij . x_ij)
       W'(r) (x_ij / r) / (r^2 + 0.01 h^2), d = 3."""
    x = np.asarray(x32, np.float64)
    v = np.asarray(v32, np.float64)
    rho = np.asarray(rho, np.float64)
    n = x.shape[0]
    i, j, xij, r = pairs_within(x, h, include_self=False)
    k = multiplicity(x32, h, n, i, j)
    xn = np.where(r[:, None] > 0, xij / np.where(r > 0, r, 1.0)[:, None], np.array([0.0, 1.0, 0.0]))
    vij = v[i] - v[j]
    s = k * particle_mass(h, rho0) / np.maximum(rho[j], 1e-6) * np.einsum("ij,ij->i", vij, xij)
    s = s * dw_shader(r, h) / (r * r + F32_0_01 * h * h)
    f = np.zeros((n, 3))
    np.add.at(f, i, s[:, None] * xn)
    return 2.0 * (3 + 2.0) * nu * f


def flat_wall_boundary(x_world, T, R, slope, y_wall, volume, h):
    """Boundary volume of one object whose object-space SDF is the linear field
    slope * (y - y_wall) and whose volume map is the constant `volume`
    (CalculateBoundaryVolume).  T: world -> object 4x4, R: the object's
    rotation 4x4 (row-major numpy arrays).  Returns (psi, grad psi)."""
    x = np.asarray(x_world, np.float64)
    xl = x @ T[:3, :3].T + T[:3, 3]
    dist = slope * (xl[:, 1] - y_wall)  # the SDF sample
    active = (dist >= 0) & (dist < h) & (volume > 0)
    # normal: object-space gradient (0, slope, 0), rotated, over its length
    nrm = R[:3, :3] @ np.array([0.0, slope, 0.0]) / abs(slope)
    d = np.maximum(dist + 0.125 * h, 0.5 * h)
    # x_ij = x - (x - d n): d n, whose length is d |n| (|n| = 1 up to the
    # float rotation's rounding)
    xij = d[:, None] * nrm[None, :]
    r = np.linalg.norm(xij, axis=1)
    psi = np.where(active, volume * w_cubic(r, h), 0.0)
    grad = np.where(active[:, None], (volume * dw_shader(r, h) / r)[:, None] * xij, 0.0)
    return psi, grad


# ------------------------------------------------------------------ scenes
H_CLOUD = float(np.float32(0.04))  # the smoothing radius as the float parameter holds it


def sheared_cloud(n=2450, seed=11, extent=0.235):
    """n particles uniform in a cube of side `extent` (mean density above
    rho0 at h = 0.04) with a velocity shear along x plus seeded noise.  With
    n = 2450 the keys (hash % n) of cells one z step apart equal those one y
    step apart (2500 = 50 mod 2450), so buckets are visited twice."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, extent, (n, 3)).astype(np.float32) + np.float32(0.3)
    v = np.zeros((n, 3), np.float32)
    v[:, 0] = 4.0 * (x[:, 1] - 0.3 - extent / 2)
    v += rng.normal(0.0, 0.05, (n, 3)).astype(np.float32)
    return x, v


def rotation_matrix(axis, angle_deg):
    """Rodrigues rotation (row-major 4x4)."""
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    t = math.radians(angle_deg)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    M = np.eye(4)
    M[:3, :3] = np.eye(3) + math.sin(t) * K + (1 - math.cos(t)) * (K @ K)
    return M


def translation_matrix(t):
    M = np.eye(4)
    M[:3, 3] = t
    return M


def wall_grid(slope, y_wall, volume, object_matrix, res=(6, 5, 5), lo=(-0.625, -0.5, -0.5), size=(1.25, 1.0, 1.0)):
    """A boundary object whose object-space SDF is slope * (y - y_wall) on a
    grid over [lo, lo + size] and whose volume map is constant (the default
    nodes and the callers' slopes and walls are binary fractions, so the float
    grid holds the linear field exactly).  Returns the
    BoundaryGrid (float32 matrices, column-major as the C ABI wants) and the
    float64 world->object and rotation matrices rebuilt from those floats."""
    from vkfluidsim_amd import BoundaryGrid
    nx, ny, nz = res
    ys = lo[1] + np.arange(ny) * (size[1] / (ny - 1))
    # index i + nx (j + ny k): k slowest, i fastest
    sdf = np.broadcast_to((slope * (ys - y_wall))[None, :, None], (nz, ny, nx)).astype(np.float32).reshape(-1)
    vm = np.full(nx * ny * nz, volume, np.float32)
    T = np.linalg.inv(object_matrix).astype(np.float32)
    R = np.eye(4, dtype=np.float32)
    R[:3, :3] = object_matrix[:3, :3]
    g = BoundaryGrid(np.ascontiguousarray(sdf), vm, tuple(res), tuple(float(s) for s in size),
                     tuple(float(p) for p in lo), transform=T.T.reshape(-1).copy(),
                     rotation=R.T.reshape(-1).copy())
    return g, T.astype(np.float64), R.astype(np.float64)


def wall_probe_points(object_matrix, y_wall, h, n=400, seed=5, spread=0.3):
    """World positions at heights -0.3h ... 1.3h above the wall, far enough
    apart laterally that their bounds stay inside the grid."""
    rng = np.random.default_rng(seed)
    local = np.empty((n, 4))
    local[:, 0] = rng.uniform(-spread, spread, n)
    local[:, 2] = rng.uniform(-spread, spread, n)
    local[:, 1] = y_wall + np.linspace(-0.3 * h, 1.3 * h, n)
    local[:, 3] = 1.0
    return (local @ object_matrix.T)[:, :3].astype(np.float32)
