"""The GPU kernels against the independent float64 numpy evaluations of
tests/brute_force.py (not only against the oracle): density, pressure and
viscosity accelerations of a dense sheared cloud with repeated bucket keys,
the boundary volume of a rotated flat wall, and the boundary pressure on lone
particles.  Tolerances are relative to each field's largest magnitude."""
import numpy as np
import pytest

import brute_force as BF
import harness as H
import vkfluidsim_amd as V

pytestmark = pytest.mark.gpu

RHO0, B, GAMMA = 1000.0, 1000.0, 7.0
NU = float(np.float32(0.05))
TOL = 2e-5
H_WALL = float(np.float32(0.05))


def _gpu_run(x, v, h, *, nu, stiffness, objects=()):
    if v is None:
        v = np.zeros_like(x)
    p = V.make_params(x.shape[0], h, dt=1e-5, gravity=(0.0, 0.0, 0.0), viscosity=nu, stiffness=stiffness,
                      box_size=(10.0, 10.0, 10.0), box_pos=(-5.0, -5.0, -5.0),
                      flags=V.VFS_FLAG_TRACK_IDS | V.VFS_FLAG_STORE_ACCEL)
    with V.Sim(p, list(objects)) as sim:
        sim.set_state(x, v)
        sim.substeps(1)
        _, _, rho, pid = sim.get_state(ids=True)
        a = sim.get_accelerations()
        psi, grad = sim.get_boundary()
    return x[pid], v[pid], rho, a, psi, grad


def test_gpu_density_pressure_viscosity_match_brute_force():
    x, v = BF.sheared_cloud()
    h = BF.H_CLOUD
    xs, vs, rho, a_p, _, _ = _gpu_run(x, v, h, nu=0.0, stiffness=B)
    xs2, _, _, a_v, _, _ = _gpu_run(x, v, h, nu=NU, stiffness=0.0)
    np.testing.assert_array_equal(xs, xs2)
    want_rho = BF.density(xs, h, RHO0)
    assert np.mean(want_rho > RHO0) > 0.5
    assert H.rel_err(rho, want_rho) < TOL
    assert H.rel_err(a_p, BF.pressure_accel(xs, want_rho, h, RHO0, B, GAMMA)) < TOL
    assert H.rel_err(a_v, BF.viscous_accel(xs, vs, want_rho, h, RHO0, NU)) < TOL


@pytest.mark.parametrize("slope,y_wall,vol,axis,angle", [(1.0, 0.0625, 0.75, (1, 2, 0.5), 37.0),
                                                         (1.5, -0.125, 0.375, (0, 0, 1), 120.0)])
def test_gpu_boundary_volume_flat_wall_closed_form(slope, y_wall, vol, axis, angle):
    M = BF.translation_matrix([0.1, 0.2, -0.05]) @ BF.rotation_matrix(axis, angle)
    g, T, R = BF.wall_grid(slope, y_wall, vol, M)
    x = BF.wall_probe_points(M, y_wall, H_WALL / slope)
    xs, _, _, _, psi, grad = _gpu_run(x, None, H_WALL, nu=0.0, stiffness=B, objects=[g])
    want_psi, want_grad = BF.flat_wall_boundary(xs, T, R, slope, y_wall, vol, H_WALL)
    active = want_psi > 0
    assert 50 < active.sum() < len(psi) - 50
    assert np.all(psi[~active] == 0.0)
    assert H.rel_err(psi, want_psi) < TOL
    assert H.rel_err(grad, want_grad) < TOL
    nrm = R[:3, :3] @ np.array([0.0, 1.0, 0.0])
    assert np.all(grad[active] @ nrm < 0)


def test_gpu_boundary_pressure_on_lone_particles():
    M = BF.rotation_matrix([1, 0, 1], 25.0)
    vol = float(np.float32(20 * H_WALL ** 3))
    g, T, R = BF.wall_grid(1.0, 0.0, vol, M)
    k = np.arange(-2, 3) * 3 * H_WALL
    gx, gz = np.meshgrid(k, k, indexing="ij")
    local = np.stack([gx.ravel(), np.linspace(0.0, 0.95 * H_WALL, gx.size), gz.ravel(), np.ones(gx.size)], -1)
    x = (local @ M.T)[:, :3].astype(np.float32)
    xs, _, rho, a, psi, grad = _gpu_run(x, None, H_WALL, nu=0.0, stiffness=B, objects=[g])
    want_psi, want_grad = BF.flat_wall_boundary(xs, T, R, 1.0, 0.0, vol, H_WALL)
    want_rho = BF.density(xs, H_WALL, RHO0, psi=want_psi)
    assert H.rel_err(rho, want_rho) < TOL
    pr = BF.eos(want_rho, RHO0, B, GAMMA)
    want_a = -((pr / want_rho ** 2 + pr / RHO0 ** 2) * RHO0)[:, None] * want_grad
    assert H.rel_err(a, want_a) < TOL
    nrm = R[:3, :3] @ np.array([0.0, 1.0, 0.0])
    pushed = (want_rho > RHO0 * 1.001) & (want_psi > 0)
    assert pushed.sum() >= 5 and np.all(a[pushed] @ nrm > 0)
