"""Seeded sweep of random scenes against the FP32 CPU oracle.

Each case draws a particle count from tens to tens of thousands (ragged: not a
multiple of the block, wave or sort tile), a smoothing radius, a hash cell of h
or up to 1.6 h, a box offset into negative coordinates, EOS stiffness and
exponent, viscosity, a gravity direction and zero to three boundary objects
(box walls, a translated sphere, a rotated and scaled box).  The particles are a
jittered block plus a uniform cloud that reaches past the box, with some exact
duplicates (pairs at r = 0) and random velocities.  Every case runs 24 substeps
and must match the oracle bit for bit after each group of 8: sorted keys,
permutation, offsets, positions, velocities, densities and boundary terms (the
kernels round every operation like the oracle, so any difference is a bug)."""
import numpy as np
import pytest

import harness as H
import vkfluidsim_amd as V
from vkfluidsim_amd import configs as CF

pytestmark = pytest.mark.gpu

CASES = 40


def _inverse_translation(pos):
    t = np.eye(4, dtype=np.float32)
    t[3, :3] = -np.asarray(pos, np.float32)  # column-major: column 3 holds the translation
    return t.reshape(-1)


def _rotated_box(r, h, center):
    half = r.uniform(0.04, 0.15, 3)
    grid = V.volume_map(V.VFS_SHAPE_BOX, (0.0, 0.0, 0.0, *half), (int(r.integers(12, 24)),) * 3, h)
    ang = r.uniform(-np.pi, np.pi)
    c, s = np.cos(ang), np.sin(ang)
    axis = int(r.integers(0, 3))
    rot = np.eye(3)
    i, j = [(1, 2), (0, 2), (0, 1)][axis]
    rot[i, i], rot[i, j], rot[j, i], rot[j, j] = c, -s, s, c
    scale = r.uniform(0.7, 1.5, 3)
    inv3 = np.diag(1.0 / scale) @ rot.T
    inv = np.eye(4)
    inv[:3, :3] = inv3
    inv[:3, 3] = -inv3 @ center
    grid.transform = inv.T.astype(np.float32).reshape(-1)  # column-major storage
    r4 = np.eye(4)
    r4[:3, :3] = rot
    grid.rotation = r4.T.astype(np.float32).reshape(-1)
    return grid


def _case(seed):
    r = np.random.default_rng(7000 + seed)
    n = int(np.exp(r.uniform(np.log(40.0), np.log(60000.0))))
    h = float(r.uniform(0.02, 0.06))
    box = r.uniform(0.4, 1.4, 3)
    box_pos = r.uniform(-0.7, 0.2, 3)
    cell = 0.0 if r.random() < 0.5 else float(h * r.uniform(1.0, 1.6))

    spacing = h / r.uniform(1.6, 2.4)
    n_block = int(n * r.uniform(0.5, 0.9))
    side = int(np.ceil(n_block ** (1.0 / 3.0)))
    block = CF.lattice((side, side, side), spacing, box_pos + spacing, 0.1 * spacing, 100 + seed)[:n_block]
    cloud = (box_pos + r.uniform(-0.1, 1.1, (n - n_block, 3)) * box).astype(np.float32)
    x = np.concatenate([block, cloud]).astype(np.float32)
    dup = r.choice(n, size=min(n // 10, 20), replace=False)
    x[dup] = x[(dup + 1) % n]  # exact duplicates: pairs at r = 0
    v = (0.3 * r.standard_normal((n, 3))).astype(np.float32)

    g = r.standard_normal(3)
    g = 9.81 * g / np.linalg.norm(g)
    params = V.make_params(n, h, dt=float(r.uniform(0.5, 2.0) * 1e-4 * h / 0.04), box_size=box, box_pos=box_pos,
                           stiffness=float(r.uniform(200.0, 3000.0)), gamma=float(r.choice([1.0, 7.0])),
                           viscosity=float(10.0 ** r.uniform(-4.0, -1.0)), gravity=g, cell_size=cell,
                           flags=V.VFS_FLAG_TRACK_IDS)

    grids = []
    if r.random() < 0.5:
        c = box_pos + 0.5 * box
        grids.append(V.volume_map(V.VFS_SHAPE_BOX_WALLS, (*c, *(0.5 * box)), (int(r.integers(16, 33)),) * 3, h))
    if r.random() < 0.5:
        sphere = V.volume_map(V.VFS_SHAPE_SPHERE, (0.0, 0.0, 0.0, r.uniform(0.05, 0.2)),
                              (int(r.integers(12, 28)),) * 3, h)
        sphere.transform = _inverse_translation(box_pos + r.uniform(0.2, 0.8, 3) * box)
        grids.append(sphere)
    if r.random() < 0.5:
        grids.append(_rotated_box(r, h, box_pos + r.uniform(0.2, 0.8, 3) * box))
    return params, grids, x, v


def _assert_same(sim, orc, what):
    for a, b, name in zip(sim.get_hash(), orc.get_hash(), ("keys", "permutation", "offsets")):
        np.testing.assert_array_equal(a, b, err_msg=f"{what}: {name}")
    x, v, rho, pid = sim.get_state(ids=True)
    rx, rv, rrho, rpid, _ = orc.get_state()
    np.testing.assert_array_equal(pid, rpid, err_msg=f"{what}: ids")
    np.testing.assert_array_equal(x.astype(np.float64), rx, err_msg=f"{what}: x")
    np.testing.assert_array_equal(v.astype(np.float64), rv, err_msg=f"{what}: v")
    np.testing.assert_array_equal(rho.astype(np.float64), rrho, err_msg=f"{what}: rho")
    psi, gr = sim.get_boundary()
    rpsi, rgr = orc.get_boundary()
    np.testing.assert_array_equal(psi.astype(np.float64), rpsi, err_msg=f"{what}: psi")
    np.testing.assert_array_equal(gr.astype(np.float64), rgr, err_msg=f"{what}: grad psi")


@pytest.mark.parametrize("seed", range(CASES))
def test_random_scene_bit_exact(seed):
    params, grids, x, v = _case(seed)
    print(f"case {seed}: n={params.n_particles} h={params.smooth_radius:.4f} cell={params.cell_size:.4f} "
          f"objects={len(grids)}", flush=True)
    sim = V.Sim(params, grids)
    sim.set_state(x, v)
    orc = H.OracleSim(params, grids)
    orc.set_state(x, v)
    for k in range(3):
        sim.substeps(8)
        orc.substeps(8)
        _assert_same(sim, orc, f"case {seed} after {8 * (k + 1)} substeps")


def test_random_scenes_as_one_batch_match_single_scenes_and_the_oracle():
    """The first 16 random scenes in one batch (each with its own smoothing
    radius, hash cell, box, gravity, EOS and boundary objects), 16 substeps:
    every scene equals its own vfs_sim bit for bit (state, hash tables,
    boundary terms) and the oracle's run of it."""
    cases = [_case(seed) for seed in range(16)]
    batch = V.Batch([p for p, _, _, _ in cases], [g for _, g, _, _ in cases])
    for k, (_, _, x, v) in enumerate(cases):
        batch.set_state(k, x, v)
    batch.substeps(16)
    for k, (params, grids, x, v) in enumerate(cases):
        print(f"scene {k}: n={params.n_particles}", flush=True)
        sim = V.Sim(params, grids)
        sim.set_state(x, v)
        sim.substeps(16)
        for a, b in zip(batch.get_state(k, ids=True), sim.get_state(ids=True)):
            np.testing.assert_array_equal(a, b, err_msg=f"scene {k}: state")
        for a, b in zip(batch.get_hash(k), sim.get_hash()):
            np.testing.assert_array_equal(a, b, err_msg=f"scene {k}: hash")
        for a, b in zip(batch.get_boundary(k), sim.get_boundary()):
            np.testing.assert_array_equal(a, b, err_msg=f"scene {k}: boundary")
        sim.close()
        orc = H.OracleSim(params, grids)
        orc.set_state(x, v)
        orc.substeps(16)
        bx, bv, brho, bpid = batch.get_state(k, ids=True)
        rx, rv, rrho, rpid, _ = orc.get_state()
        np.testing.assert_array_equal(bpid, rpid, err_msg=f"scene {k}: ids vs oracle")
        np.testing.assert_array_equal(bx.astype(np.float64), rx, err_msg=f"scene {k}: x vs oracle")
        np.testing.assert_array_equal(bv.astype(np.float64), rv, err_msg=f"scene {k}: v vs oracle")
        np.testing.assert_array_equal(brho.astype(np.float64), rrho, err_msg=f"scene {k}: rho vs oracle")
    batch.close()
