"""The reference's own scenes, committed as data under tests/golden/reference_scene/:
the dam break (assets/scenes/dam_break.json: parameters, the reference's lattice
seeding, frames against the oracle bit for bit) and the obstacle scene
(assets/scenes/obstacle.json with models/box.obj and models/suzanne.obj) loaded
through the C ABI: the JSON transforms
against hand-computed known answers and an independent T R S composition
(scene_loader.cpp:85-105, transform.cpp:34-46, generic_scene.cpp:134-136),
the OBJ triangles against a plain parse of the file, the SDF and volume-map
grids against the oracle's exhaustive search (suzanne is open, so the
generator searches every point) and the box against its closed-form distance,
and whole frames against the oracle bit for bit."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

import harness as H
import vkfluidsim_amd as V
from vkfluidsim_amd import configs as CF

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent / "golden" / "reference_scene"
SCENE = ROOT / "obstacle.json"


@pytest.fixture(scope="module")
def scene():
    s = V.Scene(str(SCENE), str(ROOT))
    yield s
    s.close()


def _object_matrix(obj):
    """T(position) R(angle about axis) S(scale), row-major float64."""
    p = np.array(obj.get("position", [0, 0, 0]), np.float64)
    s = np.array(obj.get("scale", [1, 1, 1]), np.float64)
    R = np.eye(4)
    if "rotation" in obj:
        a = np.array(obj["rotation"]["axis"], np.float64)
        a /= np.linalg.norm(a)
        t = math.radians(obj["rotation"]["angle"])
        K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
        R[:3, :3] = np.eye(3) + math.sin(t) * K + (1 - math.cos(t)) * K @ K
    T = np.eye(4)
    T[:3, 3] = p
    return T @ R @ np.diag([*s, 1.0]), R


def _parse_obj(path):
    """Positions and polygons of an OBJ file (v and f lines only)."""
    verts, faces = [], []
    for line in Path(path).read_text().splitlines():
        tok = line.split()
        if not tok:
            continue
        if tok[0] == "v":
            verts.append([float(t) for t in tok[1:4]])
        elif tok[0] == "f":
            faces.append([int(t.split("/")[0]) - 1 for t in tok[1:]])
    return np.array(verts, np.float32), faces


def test_scene_parameters(scene):
    cfg = json.loads(SCENE.read_text())
    p = scene.params
    assert scene.n == 50 * 20 * 40
    assert p.smooth_radius == pytest.approx(cfg["simulationParameters"]["smoothRadius"])
    assert p.iterations == 3 and p.fixed_dt == pytest.approx(0.008333)
    assert p.stiffness == 1000.0 and p.expoent == 7.0 and p.viscosity_strenght == pytest.approx(0.01)
    assert p.wall_stiffness == 10000.0
    assert list(p.bbox_size) == [10.0, 10.0, 10.0]


def test_transforms_known_answers_and_composition(scene):
    cfg = json.loads(SCENE.read_text())
    objs = scene.boundary_objects()
    assert len(objs) == 2
    for g, d in zip(objs, cfg["boundaryObjects"]):
        Tw2o = np.asarray(g.transform, np.float64).reshape(4, 4).T  # column-major in the ABI
        Rot = np.asarray(g.rotation, np.float64).reshape(4, 4).T
        M, R = _object_matrix(d)
        np.testing.assert_allclose(Tw2o, np.linalg.inv(M), atol=2e-7 * np.abs(np.linalg.inv(M)).max())
        np.testing.assert_allclose(Rot, R, atol=1e-7)
        np.testing.assert_allclose(Tw2o @ M, np.eye(4), atol=1e-6)
    # hand-computed: -90 degrees about y maps (x, y, z) to (-z, y, x); the box's
    # corner (1, 1, 1) scaled by (4, 3, 0.5) and rotated is (-0.5, 3, 4), so it
    # sits at (5.5, 2.9, 6.1) + (-0.5, 3, 4) = (5.0, 5.9, 10.1) in the world
    box = np.asarray(objs[0].transform, np.float64).reshape(4, 4).T
    np.testing.assert_allclose(box @ [5.0, 5.9, 10.1, 1.0], [1, 1, 1, 1], atol=1e-6)
    np.testing.assert_allclose(box @ [6.0, -0.1, 2.1, 1.0], [-1, -1, -1, 1], atol=1e-6)
    # suzanne (no scale): local (0, 0, 1) is world (7.5, 1, 1) + (-1, 0, 0)
    suz = np.asarray(objs[1].transform, np.float64).reshape(4, 4).T
    np.testing.assert_allclose(suz @ [6.5, 1.0, 1.0, 1.0], [0, 0, 1, 1], atol=1e-6)


def test_obj_triangles_match_the_file():
    # suzanne: 968 triangles, the same corner positions in the same order
    v, f = V.Mesh.from_obj(str(ROOT / "models" / "suzanne.obj")).arrays()
    pv, pf = _parse_obj(ROOT / "models" / "suzanne.obj")
    assert all(len(p) == 3 for p in pf) and len(f) == len(pf) == 968
    np.testing.assert_array_equal(v[f.astype(np.int64)], pv[np.array(pf)])
    # box: 6 quads -> 12 triangles covering each face once (area 4 each)
    v, f = V.Mesh.from_obj(str(ROOT / "models" / "box.obj")).arrays()
    assert len(f) == 12
    T = v[f.astype(np.int64)].astype(np.float64)
    area = 0.5 * np.linalg.norm(np.cross(T[:, 1] - T[:, 0], T[:, 2] - T[:, 0]), axis=1)
    np.testing.assert_allclose(area, 2.0)
    assert abs(np.sum(area) - 24.0) < 1e-9


@pytest.mark.parametrize("k,name", [(0, "box.obj"), (1, "suzanne.obj")])
def test_scene_grids_match_oracle(scene, k, name):
    g = scene.boundary_objects()[k]
    h = float(scene.params.smooth_radius)
    v, f = V.Mesh.from_obj(str(ROOT / "models" / name)).arrays()
    print(f"object {name}: {len(f)} triangles, grid {g.resolution}", flush=True)
    o = H.oracle_mesh_volume_map(v, f, g.resolution, h)
    assert tuple(g.box_size) == tuple(o.box_size) and tuple(g.box_pos) == tuple(o.box_pos)
    np.testing.assert_array_equal(g.sdf, o.sdf)
    np.testing.assert_allclose(g.vm, o.vm, rtol=2e-7, atol=1e-12)
    assert np.count_nonzero(g.vm) > 100
    if name == "box.obj":
        # closed-form distance to the cube [-1, 1]^3 at the grid nodes (inside negative)
        res = g.resolution
        ax = [g.box_pos[a] + np.arange(res[a]) * (np.float64(g.box_size[a]) / (res[a] - 1)) for a in range(3)]
        Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
        q = np.abs(np.stack([X, Y, Z], -1)) - 1.0
        d = np.linalg.norm(np.maximum(q, 0), axis=-1) + np.minimum(np.max(q, axis=-1), 0)
        np.testing.assert_allclose(g.sdf, d.reshape(-1), atol=2e-6)


def test_frames_match_oracle_bit_for_bit(scene):
    scene.reset()
    x0, v0, _, _ = scene.get_state()
    grids = scene.boundary_objects()
    orc = H.OracleSim(scene.params, grids)
    orc.set_state(x0, v0)
    touching = 0
    for frame in range(30):
        scene.step()  # one frame: `iterations` = 3 substeps
        orc.substeps(int(scene.params.iterations))
        x, v, rho, _ = scene.get_state()
        rx, rv, rrho, _, _ = orc.get_state()
        psi, _ = V.Sim.get_boundary(_SimView(scene))
        rpsi, _ = orc.get_boundary()
        touching = max(touching, int(np.count_nonzero(psi)))
        print(f"frame {frame}: {np.count_nonzero(psi)} particles near a boundary object", flush=True)
        np.testing.assert_array_equal(x.astype(np.float64), rx)
        np.testing.assert_array_equal(v.astype(np.float64), rv)
        np.testing.assert_array_equal(rho.astype(np.float64), rrho)
        np.testing.assert_array_equal(psi.astype(np.float64), rpsi)
    assert touching > 100  # the fluid reaches the box


class _SimView:
    """The scene's simulation handle with the attributes V.Sim's readers use."""

    def __init__(self, scene):
        self.handle = scene.sim_handle
        self.n = scene.n


def test_reference_dam_break_scene_loads_seeds_and_steps_like_the_oracle():
    cfg = json.loads((ROOT / "dam_break.json").read_text())
    sc = V.Scene(str(ROOT / "dam_break.json"), str(ROOT))
    p = sc.params
    assert sc.n == 30 * 50 * 50
    assert p.smooth_radius == pytest.approx(0.2) and p.iterations == 3 and p.fixed_dt == pytest.approx(0.008333)
    assert p.stiffness == 1000.0 and p.expoent == 7.0 and p.viscosity_strenght == pytest.approx(0.01)
    assert list(p.bbox_size) == [10.0, 10.0, 8.0] and list(p.gravity) == pytest.approx([0.0, -9.81, 0.0])
    assert sc.boundary_objects() == [] and cfg["boundaryObjects"] == []
    x0, v0, _, _ = sc.get_state()
    block = cfg["fluidBlocks"][0]
    np.testing.assert_array_equal(x0, CF.compact_block(tuple(block["size"]), 1000.0, pos=tuple(block["pos"])))
    assert not v0.any()
    orc = H.OracleSim(sc.params, [])
    orc.set_state(x0, v0)
    for frame in range(5):
        sc.step()  # one frame: 3 substeps
        orc.substeps(3)
        print(f"frame {frame}", flush=True)
        x, v, rho, _ = sc.get_state()
        rx, rv, rrho, _, _ = orc.get_state()
        np.testing.assert_array_equal(x.astype(np.float64), rx)
        np.testing.assert_array_equal(v.astype(np.float64), rv)
        np.testing.assert_array_equal(rho.astype(np.float64), rrho)
    sc.close()
