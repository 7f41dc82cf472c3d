"""GPU mesh signed distances and mesh volume maps (vfs_mesh_*) against the
oracle's exhaustive-search restatement on the same float query points, and
scenes whose boundary objects are OBJ files."""
import json

import numpy as np
import pytest

import harness as H
import vkfluidsim_amd as V
from vkfluidsim_amd import configs as CF

pytestmark = pytest.mark.gpu

MESHES = {
    "box": lambda: CF.box_mesh((0.3, -0.2, 0.5), (0.4, 0.25, 0.6)),
    "icosphere": lambda: CF.icosphere(3, 0.5, (0.1, 0.2, 0.3)),
    "torus": lambda: CF.torus_mesh(1.0, 0.4, 24, 12, (0.0, 0.0, 0.0)),
}


def _query_points(v, f, n, seed):
    """Uniform points in the grown box plus every vertex, edge midpoint and
    triangle centroid (points on the surface)."""
    rng = np.random.default_rng(seed)
    lo, hi = v.min(axis=0), v.max(axis=0)
    pad = 0.3 * (hi - lo)
    pts = [rng.uniform(lo - pad, hi + pad, size=(n, 3))]
    T = v[f.astype(np.int64)].astype(np.float64)
    pts += [v.astype(np.float64), 0.5 * (T[:, 0] + T[:, 1]), T.mean(axis=1)]
    return np.concatenate(pts).astype(np.float32)


@pytest.mark.parametrize("name", sorted(MESHES))
def test_signed_distance_matches_oracle(name):
    v, f = MESHES[name]()
    pts = _query_points(v, f, 20000, 17)
    print(f"case {name}: {len(f)} triangles, {len(pts)} points", flush=True)
    m = V.Mesh.from_arrays(v, f)
    g = m.signed_distance(pts)
    o = H.oracle_mesh_signed_distance(v, f, pts)
    # the distance is the minimum over all triangles: identical bits
    np.testing.assert_array_equal(np.abs(g), np.abs(o))
    off = o != 0.0
    np.testing.assert_array_equal(np.sign(g[off]), np.sign(o[off]))
    assert (o < 0).sum() > 100
    m.close()


@pytest.mark.parametrize("name,res,h", [("box", (12, 10, 14), 0.05), ("icosphere", (14, 14, 14), 0.06),
                                        ("torus", (16, 16, 8), 0.15)])
def test_mesh_volume_map_matches_oracle(name, res, h):
    """Box and icosphere (closed, no obtuse angle) take the generator's
    search-skipping path, the torus (obtuse triangles) the exhaustive one; the
    oracle searches every point."""
    v, f = MESHES[name]()
    print(f"case {name}: res {res}, h {h}", flush=True)
    m = V.Mesh.from_arrays(v, f)
    g = m.volume_map(res, h)
    o = H.oracle_mesh_volume_map(v, f, res, h)
    assert g.box_size == o.box_size and g.box_pos == o.box_pos
    np.testing.assert_array_equal(g.sdf, o.sdf)
    # FP64 sums of up to 4096 terms; device pow may differ from the host's in
    # the last bit of a double, which can move the float result by one ulp
    np.testing.assert_allclose(g.vm, o.vm, rtol=2e-7, atol=1e-12)
    assert np.count_nonzero(g.vm) > 0
    m.close()


def test_box_mesh_grids_match_analytic_box_on_gpu():
    c, e = (0.0, 0.0, 0.0), (0.1, 0.05, 0.08)
    v, f = CF.box_mesh(c, e)
    res, h = (20, 16, 18), 0.02
    g = V.Mesh.from_arrays(v, f).volume_map(res, h)
    a = V.volume_map(V.VFS_SHAPE_BOX, (*c, *e), res, h)
    assert g.box_size == a.box_size and g.box_pos == a.box_pos
    np.testing.assert_allclose(g.sdf, a.sdf, atol=2e-7)
    np.testing.assert_allclose(g.vm, a.vm, rtol=1e-5, atol=1e-9)


def test_scene_with_obj_boundary_objects(tmp_path):
    """A scene in the layout of the reference's obstacle scene: a rotated and
    scaled box OBJ (quads) and an icosphere OBJ, with resource paths; 20
    frames bit-exact against the oracle driven by the scene's own objects."""
    models = tmp_path / "models"
    models.mkdir()
    bv, _ = CF.box_mesh((0.0, 0.0, 0.0), (0.5, 0.5, 0.5))
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    H.write_obj(models / "box.obj", bv, quads)
    sv, sf = CF.icosphere(2, 0.5)
    H.write_obj(models / "ball.obj", sv, sf)
    scene = {
        "name": "Obstacles", "fluidModel": "wcsph",
        "wcsphParameters": {"stiffness": 1000.0, "expoent": 7.0, "viscosityStrenght": 0.01,
                            "outerWallStiffness": 10000.0},
        "simulationParameters": {"gravity": [0.0, -9.81, 0.0], "smoothRadius": 0.2, "timeScale": 1.0,
                                 "iterations": 3, "dt": 0.008333, "targetDensity": 1000.0,
                                 "boundingBox": {"pos": [0, 0, 0], "size": [10, 10, 10]}},
        "fluidBlocks": [{"size": [20, 10, 16], "pos": [0.0, 0.0, 2.0]}],
        "boundaryObjects": [
            {"resourcePath": "models/box.obj", "position": [5.5, 2.9, 6.1],
             "rotation": {"angle": -90.0, "axis": [0, 1, 0]}, "scale": [4, 3, 0.5],
             "volumeMapResolution": [20, 20, 20]},
            {"resourcePath": "models/ball.obj", "position": [3.0, 1.0, 3.0], "volumeMapResolution": [16, 16, 16]},
        ],
    }
    path = tmp_path / "obstacles.json"
    path.write_text(json.dumps(scene))
    s = V.Scene(str(path), str(tmp_path))
    assert s.n == 20 * 10 * 16
    # the loaded objects (grids built on the GPU, transforms from the JSON) drive
    # the oracle, which must then follow the scene bit for bit
    x0, v0, _, _ = s.get_state()
    orc = H.OracleSim(s.params, s.boundary_objects())
    orc.set_state(x0, v0)
    for frame in range(20):
        s.step()
        orc.substeps(int(s.params.iterations))
        if frame % 5 == 4:
            x, v, rho, _ = s.get_state()
            rx, rv, rrho, _, _ = orc.get_state()
            np.testing.assert_array_equal(x.astype(np.float64), rx)
            np.testing.assert_array_equal(v.astype(np.float64), rv)
            np.testing.assert_array_equal(rho.astype(np.float64), rrho)
    x, v, rho, _ = s.get_state()
    assert np.isfinite(x).all() and np.isfinite(v).all() and np.isfinite(rho).all()
    s.close()
