"""Mesh boundary objects on the CPU: the oracle's point-triangle distance,
pseudonormals, signed distance and mesh volume map pinned against brute force,
closed forms and winding numbers; the library's OBJ loader (no GPU needed)."""
import numpy as np
import pytest

import harness as H
import vkfluidsim_amd as V
from vkfluidsim_amd import configs as CF


def _winding(vertices, triangles, points):
    """Generalised winding number (Van Oosterom-Strackee solid angles)."""
    Vd = np.asarray(vertices, np.float64)
    T = Vd[np.asarray(triangles, np.int64)]
    out = np.empty(len(points))
    for k, p in enumerate(np.asarray(points, np.float64)):
        a, b, c = T[:, 0] - p, T[:, 1] - p, T[:, 2] - p
        la, lb, lc = (np.linalg.norm(x, axis=1) for x in (a, b, c))
        num = np.einsum("ij,ij->i", a, np.cross(b, c))
        den = la * lb * lc + np.einsum("ij,ij->i", a, b) * lc + np.einsum("ij,ij->i", b, c) * la + \
            np.einsum("ij,ij->i", c, a) * lb
        out[k] = np.sum(2 * np.arctan2(num, den)) / (4 * np.pi)
    return out


def test_point_triangle_matches_dense_sampling():
    rng = np.random.default_rng(11)
    n = 300
    tris = rng.normal(size=(n, 3, 3))
    tris[: n // 3, 2] = tris[: n // 3, 0] + 3.0 * (tris[: n // 3, 1] - tris[: n // 3, 0]) + 0.05 * rng.normal(
        size=(n // 3, 3))  # obtuse slivers
    pts = rng.normal(size=(n, 3)) * 2.0
    d2, cp, feat = H.oracle_point_triangle(pts, tris)
    N = 240
    s, t = np.meshgrid(np.linspace(0, 1, N + 1), np.linspace(0, 1, N + 1), indexing="ij")
    keep = s + t <= 1 + 1e-12
    s, t = s[keep], t[keep]
    for k in range(n):
        v0, v1, v2 = tris[k]
        e0, e1 = v1 - v0, v2 - v0
        # the closest point is on the triangle and at the reported distance
        A = np.stack([e0, e1], 1)
        st, *_ = np.linalg.lstsq(A, cp[k] - v0, rcond=None)
        assert st.min() >= -1e-9 and st.sum() <= 1 + 1e-9
        assert np.sum((pts[k] - cp[k]) ** 2) == pytest.approx(d2[k], rel=1e-9, abs=1e-12)
        # no sample of the triangle is closer; the nearest sample is within the sampling error
        samples = v0 + s[:, None] * e0 + t[:, None] * e1
        dm = np.min(np.sum((samples - pts[k]) ** 2, axis=1))
        hs = max(np.linalg.norm(e0), np.linalg.norm(e1), np.linalg.norm(v2 - v1)) / N
        assert d2[k] <= dm + 1e-12
        assert np.sqrt(dm) - np.sqrt(d2[k]) <= 2 * hs
        # the feature names where the closest point is
        f = feat[k]
        corners = {0: v0, 1: v1, 2: v2}
        if f in corners:
            np.testing.assert_allclose(cp[k], corners[f], atol=1e-12)
        elif f == 6:
            assert st.min() > -1e-9 and st.sum() < 1 + 1e-9
        else:
            a, b = {3: (v0, v1), 4: (v1, v2), 5: (v0, v2)}[f]
            ab = b - a
            u = np.dot(cp[k] - a, ab) / np.dot(ab, ab)
            np.testing.assert_allclose(cp[k], a + u * ab, atol=1e-9 * (1 + np.linalg.norm(ab)))


def test_cube_pseudonormals():
    v, f = CF.box_mesh((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    face, edge, vert = H.oracle_pseudonormals(v, f)
    for k in range(len(v)):
        np.testing.assert_allclose(vert[k], np.sign(v[k]) / np.sqrt(3.0), atol=1e-6)
    for t in range(len(f)):
        idx = f[t]
        n = face[t]
        assert np.count_nonzero(np.abs(n) > 0.5) == 1
        for e, (a, b) in enumerate([(0, 1), (1, 2), (0, 2)]):
            mid = 0.5 * (v[idx[a]] + v[idx[b]])
            on_axes = np.isclose(np.abs(mid), 1.0)
            if on_axes.sum() == 2:  # a cube edge: average of the two face normals
                want = np.where(on_axes, np.sign(mid), 0.0) / np.sqrt(2.0)
            else:  # a face diagonal: the face normal
                want = n
            np.testing.assert_allclose(edge[t, e], want, atol=1e-6)


def test_box_mesh_distance_matches_closed_form():
    c, e = np.array([0.3, -0.2, 0.5]), np.array([0.4, 0.25, 0.6])
    v, f = CF.box_mesh(c, e)
    rng = np.random.default_rng(3)
    pts = (c + rng.uniform(-2.0, 2.0, size=(4000, 3)) * e).astype(np.float32)
    d = H.oracle_mesh_signed_distance(v, f, pts)
    q = np.abs(pts.astype(np.float64) - c) - e
    want = np.linalg.norm(np.maximum(q, 0), axis=1) + np.minimum(q.max(axis=1), 0)
    np.testing.assert_allclose(d, want, atol=1e-6)


def test_icosphere_distance_close_to_sphere():
    v, f = CF.icosphere(3, 1.0)
    rng = np.random.default_rng(5)
    pts = rng.uniform(-2, 2, size=(3000, 3)).astype(np.float32)
    d = H.oracle_mesh_signed_distance(v, f, pts)
    want = np.linalg.norm(pts.astype(np.float64), axis=1) - 1.0
    # the level-3 icosphere lies within 0.5% of the unit sphere
    assert np.abs(d - want).max() < 0.01
    far = np.abs(want) > 0.02
    assert np.array_equal(np.sign(d[far]), np.sign(want[far]))


def test_signs_match_winding_numbers_on_a_torus():
    """Non-convex closed mesh: the pseudonormal sign against the winding number."""
    v, f = CF.torus_mesh()
    rng = np.random.default_rng(9)
    pts = rng.uniform([-1.6, -1.6, -0.6], [1.6, 1.6, 0.6], size=(1500, 3)).astype(np.float32)
    d = H.oracle_mesh_signed_distance(v, f, pts)
    w = _winding(v, f, pts)
    clear = np.abs(d) > 1e-3
    assert clear.sum() > 1400
    inside = w[clear] > 0.5
    assert np.array_equal(d[clear] < 0, inside)
    assert 100 < inside.sum() < clear.sum() - 100


def test_box_mesh_volume_map_matches_analytic_box():
    """The mesh generator on a 12-triangle box against the analytic-box grids
    (pinned separately against independent quadrature)."""
    c, e = (0.0, 0.0, 0.0), (0.1, 0.05, 0.08)
    v, f = CF.box_mesh(c, e)
    res, h = (10, 9, 11), 0.04
    m = H.oracle_mesh_volume_map(v, f, res, h)
    a = H.oracle_volume_map(0, (*c, *e), res, h)
    assert m.box_size == a.box_size and m.box_pos == a.box_pos
    np.testing.assert_allclose(m.sdf, a.sdf, atol=2e-7)
    np.testing.assert_allclose(m.vm, a.vm, rtol=1e-5, atol=1e-9)
    assert np.count_nonzero(m.vm) > 0


# ----------------------------------------------------------------- OBJ loader
def test_obj_loader_triangulates_and_keeps_the_last_shape(tmp_path):
    # shape 1 (ignored): a triangle; shape 2: a quad, a pentagon, a triangle with
    # v/vt/vn tokens and negative indices
    text = "\n".join([
        "# comment", "o first", "v 0 0 0", "v 1 0 0", "v 0 1 0", "f 1 2 3",
        "o second",
        "v 0 0 1", "v 2 0 1", "v 2 1 1", "v 0 1 1",            # 4..7: a 2 x 1 rectangle
        "v 5 0 0", "v 6 0 0", "v 6.5 1 0", "v 5.5 2 0", "v 4.5 1 0",  # 8..12: pentagon
        "vt 0 0", "vn 0 0 1",
        "f 4/1/1 5/1/1 6/1/1 7/1/1",
        "f 8 9 10 11 12",
        "f -3//1 -2//1 -1//1",
        "g tail",  # a group with no faces after it
    ]) + "\r\n"
    p = tmp_path / "m.obj"
    p.write_text(text)
    m = V.Mesh.from_obj(str(p))
    verts, tris = m.arrays()
    # rectangle: |e02|^2 = 5 = |e13|^2, not smaller -> (0,1,3), (1,2,3);
    # pentagon: fan from its first corner; vertices in first-use order
    order = [4, 5, 7, 6, 8, 9, 10, 11, 12]
    want_file = [(4, 5, 7), (5, 6, 7), (8, 9, 10), (8, 10, 11), (8, 11, 12), (10, 11, 12)]
    file_v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [2, 0, 1], [2, 1, 1], [0, 1, 1], [5, 0, 0],
                       [6, 0, 0], [6.5, 1, 0], [5.5, 2, 0], [4.5, 1, 0]], np.float32)
    first_use = []
    for t in want_file:
        for i in t:
            if i not in first_use:
                first_use.append(i)
    assert sorted(first_use) == sorted(order)
    np.testing.assert_array_equal(verts, file_v[[i - 1 for i in first_use]])
    np.testing.assert_array_equal(tris, [[first_use.index(i) for i in t] for t in want_file])
    nv, nt, lo, hi = m.info()
    assert (nv, nt) == (9, 6)
    np.testing.assert_array_equal(lo, [0, 0, 0])
    np.testing.assert_array_equal(hi, [6.5, 2, 1])
    m.close()


def test_obj_loader_quad_split_takes_the_shorter_diagonal(tmp_path):
    # corners 0..3 with |v2 - v0| < |v3 - v1| -> (0,1,2), (0,2,3)
    p = tmp_path / "q.obj"
    H.write_obj(p, [[0, 0, 0], [3, 0, 0], [1, 1, 0], [0, 1, 0]], [(0, 1, 2, 3)])
    _, tris = V.Mesh.from_obj(str(p)).arrays()
    np.testing.assert_array_equal(tris, [[0, 1, 2], [0, 2, 3]])


def test_obj_loader_errors(tmp_path):
    with pytest.raises(V.VfsError):
        V.Mesh.from_obj(str(tmp_path / "missing.obj"))
    p = tmp_path / "bad.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nf 1 2 7\n")
    with pytest.raises(V.VfsError, match="out of range"):
        V.Mesh.from_obj(str(p))
    p.write_text("v 0 0 0\n")
    with pytest.raises(V.VfsError, match="no faces"):
        V.Mesh.from_obj(str(p))
    with pytest.raises(V.VfsError):
        V.Mesh.from_arrays(np.zeros((3, 3)), np.array([[0, 1, 3]]))


def test_icosphere_generator():
    v, f = CF.icosphere(4, 0.2, (1.0, 0.5, 0.5))
    assert v.shape == (2562, 3) and f.shape == (5120, 3)
    np.testing.assert_allclose(np.linalg.norm(v - np.array([1.0, 0.5, 0.5], np.float32), axis=1), 0.2, rtol=1e-6)
    # closed: every edge in exactly two triangles, in opposite directions
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    assert len({tuple(x) for x in e}) == len(e)
    assert {tuple(x) for x in e} == {tuple(x[::-1]) for x in e}


def test_mesh_properties():
    """closed / acute decide whether the volume-map generator may skip searches."""
    for gen, want in [(lambda: CF.box_mesh((0, 0, 0), (1, 2, 3)), (True, True)),
                      (lambda: CF.icosphere(3, 0.5), (True, True)),
                      (lambda: CF.torus_mesh(), (True, False))]:
        v, f = gen()
        assert V.Mesh.from_arrays(v, f).properties() == want
    v, f = CF.icosphere(1, 1.0)
    assert V.Mesh.from_arrays(v, f[:-1]).properties()[0] is False  # a hole
    assert V.Mesh.from_arrays(v, f[:, [0, 2, 1]]).properties()[0] is True  # inward, still consistent
    g = f.copy()
    g[0] = g[0, [0, 2, 1]]
    assert V.Mesh.from_arrays(v, g).properties()[0] is False  # one flipped triangle
