"""Pins of the oracle's force and boundary terms against independent float64
numpy evaluations written from the shader formulas (tests/brute_force.py):
O(N^2) density, pressure and viscosity sums on a dense sheared cloud (hash
collisions included through the stencil-key multiplicities), the closed form
of a flat wall's boundary volume (object rotated, translated, SDF of slope
!= 1), the boundary pressure on lone particles, and GradientVector on linear
and kinked fields.  Each check is shown to fail for the plausible mistakes
listed in the planted-mistake tests (a sign flip, a dropped mass or 1/h factor, p/rho in place
of p/rho^2, 2 (d + 2) with d = 2, a reflected boundary point ...)."""
import numpy as np
import pytest

import brute_force as BF
import harness as H
from vkfluidsim_amd import make_params

RHO0, B, GAMMA = 1000.0, 1000.0, 7.0
NU = float(np.float32(0.05))  # the viscosity as the float parameter holds it
TOL = {True: 1e-10, False: 2e-5}  # relative to the field's largest magnitude: FP64 twin, FP32 oracle


def _cloud_run(fp64, *, nu, stiffness, objects=(), x=None, v=None, h=BF.H_CLOUD):
    if x is None:
        x, v = BF.sheared_cloud()
    p = make_params(x.shape[0], h, dt=1e-5, gravity=(0.0, 0.0, 0.0), viscosity=nu, stiffness=stiffness,
                    box_size=(10.0, 10.0, 10.0), box_pos=(-5.0, -5.0, -5.0), flags=1)
    o = H.OracleSim(p, list(objects), fp64=fp64, nthreads=4)
    o.set_state(x, v)
    o.substeps(1)
    _, _, rho, pid, a = o.get_state()
    psi, grad = o.get_boundary()
    # the substep's forces act on the sorted input state: x0[pid], v0[pid]
    return x[pid], v[pid], rho, a, psi, grad


@pytest.fixture(scope="module")
def cloud():
    out = {}
    for fp64 in (True, False):
        xs, vs, rho_p, a_p, _, _ = _cloud_run(fp64, nu=0.0, stiffness=B)
        xs2, vs2, rho_v, a_v, _, _ = _cloud_run(fp64, nu=NU, stiffness=0.0)
        np.testing.assert_array_equal(xs, xs2)
        out[fp64] = dict(x=xs, v=vs, rho=rho_p, a_p=a_p, a_v=a_v)
    return out


def test_cloud_exercises_pressure_and_key_collisions(cloud):
    c = cloud[True]
    n = c["x"].shape[0]
    assert np.mean(c["rho"] > RHO0) > 0.5  # pressure acts on most particles
    i, j, _, _ = BF.pairs_within(c["x"].astype(np.float64), BF.H_CLOUD, include_self=True)
    k = BF.multiplicity(c["x"], BF.H_CLOUD, n, i, j)
    assert k.min() >= 1 and np.sum(k > 1) > 100  # buckets visited twice occur


@pytest.mark.parametrize("fp64", [True, False])
def test_density_matches_brute_force(cloud, fp64):
    c = cloud[fp64]
    want = BF.density(c["x"], BF.H_CLOUD, RHO0)
    assert H.rel_err(c["rho"], want) < TOL[fp64]


@pytest.mark.parametrize("fp64", [True, False])
def test_pressure_matches_brute_force(cloud, fp64):
    c = cloud[fp64]
    rho = BF.density(c["x"], BF.H_CLOUD, RHO0)
    want = BF.pressure_accel(c["x"], rho, BF.H_CLOUD, RHO0, B, GAMMA)
    assert np.abs(want).max() > 10.0
    assert H.rel_err(c["a_p"], want) < TOL[fp64]


@pytest.mark.parametrize("fp64", [True, False])
def test_viscosity_matches_brute_force(cloud, fp64):
    c = cloud[fp64]
    rho = BF.density(c["x"], BF.H_CLOUD, RHO0)
    want = BF.viscous_accel(c["x"], c["v"], rho, BF.H_CLOUD, RHO0, NU)
    assert np.abs(want).max() > 1.0
    assert H.rel_err(c["a_v"], want) < TOL[fp64]


def _mutant_pressure(x, rho, h, kind):
    m = BF.particle_mass(h, RHO0)
    p = BF.eos(rho, RHO0, B, GAMMA)
    n = x.shape[0]
    i, j, xij, r = BF.pairs_within(x.astype(np.float64), h, include_self=False)
    k = BF.multiplicity(x, h, n, i, j)
    xn = xij / r[:, None]
    pd = p / rho ** 2 if kind != "p_over_rho" else p / rho
    dw = BF.dw_shader(r, h) * (h if kind == "no_extra_inv_h" else 1.0)
    mass = 1.0 if kind == "no_mass" else m
    kk = np.ones_like(k) if kind == "no_multiplicity" else k
    s = kk * mass * (pd[i] + pd[j]) * dw
    f = np.zeros((n, 3))
    np.add.at(f, i, s[:, None] * xn)
    return f if kind == "sign" else -f


@pytest.mark.parametrize("kind", ["sign", "no_mass", "no_extra_inv_h", "p_over_rho", "no_multiplicity"])
def test_pressure_check_fails_for_planted_mistakes(cloud, kind):
    c = cloud[False]
    rho = BF.density(c["x"], BF.H_CLOUD, RHO0)
    assert H.rel_err(c["a_p"], _mutant_pressure(c["x"], rho, BF.H_CLOUD, kind)) > 100 * TOL[False]


@pytest.mark.parametrize("kind", ["d_is_2", "sign", "rho_i", "no_denominator_eps"])
def test_viscosity_check_fails_for_planted_mistakes(cloud, kind):
    c = cloud[False]
    x, v, h = c["x"], c["v"], BF.H_CLOUD
    rho = BF.density(x, h, RHO0)
    if kind == "d_is_2":
        want = BF.viscous_accel(x, v, rho, h, RHO0, NU) * (2 * (2 + 2)) / (2 * (3 + 2))
    elif kind == "sign":
        want = -BF.viscous_accel(x, v, rho, h, RHO0, NU)
    else:
        n = x.shape[0]
        X, V = x.astype(np.float64), v.astype(np.float64)
        i, j, xij, r = BF.pairs_within(X, h, include_self=False)
        k = BF.multiplicity(x, h, n, i, j)
        dens = rho[i] if kind == "rho_i" else rho[j]
        eps = 0.0 if kind == "no_denominator_eps" else BF.F32_0_01 * h * h
        s = k * BF.particle_mass(h, RHO0) / dens * np.einsum("ij,ij->i", V[i] - V[j], xij)
        s = s * BF.dw_shader(r, h) / (r * r + eps)
        want = np.zeros((n, 3))
        np.add.at(want, i, s[:, None] * xij / r[:, None])
        want *= 10.0 * NU
    assert H.rel_err(c["a_v"], want) > 100 * TOL[False]


# ------------------------------------------------------------ boundary volume
H_WALL = float(np.float32(0.05))
WALL_CASES = {
    # slope, wall height, volume-map value, object matrix
    "identity": (1.0, 0.0, 0.75, np.eye(4)),
    "rotated": (1.0, 0.0625, 0.75, BF.translation_matrix([0.1, 0.2, -0.05]) @ BF.rotation_matrix([1, 2, 0.5], 37.0)),
    "rotated_slope": (1.5, -0.125, 0.375,
                      BF.translation_matrix([-0.2, 0.1, 0.3]) @ BF.rotation_matrix([0, 0, 1], 120.0)),
}


def _wall_run(case, fp64):
    slope, y_wall, vol, M = WALL_CASES[case]
    g, T, R = BF.wall_grid(slope, y_wall, vol, M)
    x = BF.wall_probe_points(M, y_wall, H_WALL / slope)
    p = make_params(x.shape[0], H_WALL, dt=1e-5, gravity=(0.0, 0.0, 0.0), viscosity=0.0,
                    box_size=(10.0, 10.0, 10.0), box_pos=(-5.0, -5.0, -5.0), flags=1)
    o = H.OracleSim(p, [g], fp64=fp64, nthreads=4)
    o.set_state(x)
    o.substeps(1)
    _, _, _, pid, _ = o.get_state()
    psi, grad = o.get_boundary()
    return x[pid], psi, grad, T, R


@pytest.mark.parametrize("fp64", [True, False])
@pytest.mark.parametrize("case", sorted(WALL_CASES))
def test_boundary_volume_flat_wall_closed_form(case, fp64):
    slope, y_wall, vol, M = WALL_CASES[case]
    xs, psi, grad, T, R = _wall_run(case, fp64)
    want_psi, want_grad = BF.flat_wall_boundary(xs, T, R, slope, y_wall, vol, H_WALL)
    active = want_psi > 0
    assert 50 < active.sum() < len(psi) - 50  # points below, inside and beyond the support
    assert np.all(psi[~active] == 0.0) and np.all(grad[~active] == 0.0)
    assert H.rel_err(psi, want_psi) < TOL[fp64]
    assert H.rel_err(grad, want_grad) < TOL[fp64]
    # grad psi points into the wall: against the rotated normal
    nrm = R[:3, :3] @ np.array([0.0, 1.0, 0.0])
    assert np.all(grad[active] @ nrm < 0)


@pytest.mark.parametrize("kind", ["reflected", "no_floor", "no_offset", "no_rotation", "unnormalised", "W_for_grad"])
def test_boundary_check_fails_for_planted_mistakes(kind):
    case = "rotated_slope"
    slope, y_wall, vol, M = WALL_CASES[case]
    xs, psi, grad, T, R = _wall_run(case, False)
    x = xs.astype(np.float64)
    xl = x @ T[:3, :3].T + T[:3, 3]
    dist = slope * (xl[:, 1] - y_wall)
    active = (dist >= 0) & (dist < H_WALL)
    nrm = R[:3, :3] @ np.array([0.0, 1.0, 0.0])
    d = np.maximum(dist + 0.125 * H_WALL, 0.5 * H_WALL)
    if kind == "no_floor":
        d = dist + 0.125 * H_WALL
    elif kind == "no_offset":
        d = np.maximum(dist, 0.5 * H_WALL)
    elif kind == "no_rotation":
        nrm = np.array([0.0, 1.0, 0.0])
    elif kind == "unnormalised":
        d = d * slope
    sgn = -1.0 if kind == "reflected" else 1.0
    gw = BF.w_cubic(d, H_WALL) if kind == "W_for_grad" else BF.dw_shader(d, H_WALL)
    want_psi = np.where(active, vol * BF.w_cubic(d, H_WALL), 0.0)
    want_grad = np.where(active[:, None], sgn * (vol * gw)[:, None] * nrm[None, :], 0.0)
    err = max(H.rel_err(psi, want_psi), H.rel_err(grad, want_grad))
    assert err > 100 * TOL[False]


@pytest.mark.parametrize("fp64", [True, False])
def test_boundary_pressure_on_lone_particles(fp64):
    """Particles 3h apart above a wall: no pair terms, so rho = k m W(0) + rho0
    psi (k: visits of the own bucket) and the pressure acceleration is -(p/rho^2 + p/rho0^2) rho0 grad psi,
    pushing away from the wall wherever rho > rho0."""
    M = BF.translation_matrix([0.0, 0.0, 0.0]) @ BF.rotation_matrix([1, 0, 1], 25.0)
    vol = float(np.float32(20 * H_WALL ** 3))  # psi ~ 1.3 near the wall: rho > rho0 there
    g, T, R = BF.wall_grid(1.0, 0.0, vol, M)
    k = np.arange(-2, 3) * 3 * H_WALL
    gx, gz = np.meshgrid(k, k, indexing="ij")
    heights = np.linspace(0.0, 0.95 * H_WALL, gx.size)
    local = np.stack([gx.ravel(), heights, gz.ravel(), np.ones(gx.size)], -1)
    x = (local @ M.T)[:, :3].astype(np.float32)
    p = make_params(x.shape[0], H_WALL, dt=1e-5, gravity=(0.0, 0.0, 0.0), viscosity=0.0,
                    box_size=(10.0, 10.0, 10.0), box_pos=(-5.0, -5.0, -5.0), flags=1)
    o = H.OracleSim(p, [g], fp64=fp64, nthreads=2)
    o.set_state(x)
    o.substeps(1)
    _, _, rho, pid, a = o.get_state()
    psi, grad = o.get_boundary()
    want_psi, want_grad = BF.flat_wall_boundary(x[pid], T, R, 1.0, 0.0, vol, H_WALL)
    # self terms only (times the visits of the own bucket: 25 keys repeat among 27 stencil cells)
    want_rho = BF.density(x[pid], H_WALL, RHO0, psi=want_psi)
    assert H.rel_err(rho, want_rho) < TOL[fp64]
    pr = BF.eos(want_rho, RHO0, B, GAMMA)
    want_a = -((pr / want_rho ** 2 + pr / RHO0 ** 2) * RHO0)[:, None] * want_grad
    assert H.rel_err(a, want_a) < TOL[fp64]
    nrm = R[:3, :3] @ np.array([0.0, 1.0, 0.0])
    pushed = (want_rho > RHO0 * 1.001) & (want_psi > 0)
    assert pushed.sum() >= 5 and np.all(a[pushed] @ nrm > 0)


# ------------------------------------------------------------ GradientVector
@pytest.mark.parametrize("fp64", [True, False])
def test_gradient_vector_exact_on_linear_sdf(fp64):
    from vkfluidsim_amd import BoundaryGrid
    h = 0.04
    # binary-fraction nodes and coefficients: the float grid holds the field exactly
    res = (5, 7, 9)
    lo, size = np.array([-0.25, -0.5, -0.375]), np.array([0.5, 0.75, 1.0])
    c0, c = 0.0625, np.array([0.75, -1.25, 2.5])
    ax = [lo[a] + np.arange(res[a]) * (size[a] / (res[a] - 1)) for a in range(3)]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    f = c0 + c[0] * X + c[1] * Y + c[2] * Z
    sdf = np.ascontiguousarray(np.transpose(f, (2, 1, 0)).reshape(-1), np.float32)
    g = BoundaryGrid(sdf, np.zeros_like(sdf), res, tuple(size), tuple(lo))
    o = H.OracleSim(make_params(1, h, dt=1e-5), [g], fp64=fp64, nthreads=1)
    rng = np.random.default_rng(1)
    for q in rng.uniform(lo, lo + size, (100, 3)):
        np.testing.assert_allclose(o.gradient_vector(0, q), c, rtol=0, atol=1e-9 if fp64 else 2e-3)


def test_gradient_vector_step_is_a_tenth_of_h():
    """A kinked field |x - k| (slopes -1 / +1 across a node at k): at a point
    a distance delta < eps left of the kink the central difference is
    (s1 (eps + delta) + s2 (eps - delta)) / (2 eps) = -delta / eps, which
    fixes eps = 0.1 h (and the 1 / (2 eps) scale)."""
    from vkfluidsim_amd import BoundaryGrid
    h = 0.25
    res = (9, 3, 3)
    lo, size = np.array([0.0, 0.0, 0.0]), np.array([1.0, 0.25, 0.25])
    xs = lo[0] + np.arange(res[0]) * 0.125
    kink = xs[4]
    f = np.broadcast_to(np.abs(xs - kink)[None, None, :], (res[2], res[1], res[0]))
    g = BoundaryGrid(np.ascontiguousarray(f, np.float32).reshape(-1), np.zeros(int(np.prod(res)), np.float32), res,
                     tuple(size), tuple(lo))
    o = H.OracleSim(make_params(1, h, dt=1e-5), [g], fp64=True, nthreads=1)
    eps = float(np.float32(0.1)) * h  # the shader's 0.1 is a float literal
    for delta in (0.002, 0.008, 0.015, 0.024):
        got = o.gradient_vector(0, np.array([kink - delta, 0.125, 0.125]))
        assert got[0] == pytest.approx(-delta / eps, abs=1e-9)
        assert got[1] == pytest.approx(0.0, abs=1e-9) and got[2] == pytest.approx(0.0, abs=1e-9)
    # beyond eps the slope is the plain one
    assert o.gradient_vector(0, np.array([kink - 0.03, 0.125, 0.125]))[0] == pytest.approx(-1.0, abs=1e-9)
