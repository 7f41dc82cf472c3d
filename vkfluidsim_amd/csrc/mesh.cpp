// Triangle meshes as boundary objects (host side): OBJ loading
// (LoadObjMesh, src/util/mesh_loader.cpp:20-113), pseudonormals
// (MeshPseudonormals::Build, mesh_pseudonormals.cpp:11-103), a BVH for the
// closest-point search, and the C ABI of the GPU mesh SDF / volume-map
// generator (MeshSDF::Build, mesh_sdf.cpp:28-45; GenerateVolumeMap,
// volume_map.cpp:9-54).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "mesh_internal.h"
#include "../../include/vfs_c.h"

using namespace vfs;

namespace {

// ------------------------------------------------------------------ OBJ
// The reference reads OBJ files with tinyobjloader (triangulate = true) and
// keeps the last shape (generic_scene.cpp:93-94).  Restated here: `v` lines
// give positions (parsed as double, stored as float), `f` lines give polygons
// whose corners are `i`, `i/j`, `i//k` or `i/j/k` (1-based, negative = relative
// to the vertices read so far); `o` and `g` lines start a new shape.  Quads
// are split along their shorter diagonal (tinyobjloader 2.x); larger polygons
// as a fan from their first corner.  Only positions matter for the SDF.
struct ObjShape {
    std::vector<uint32_t> tris;  // indices into the file's vertex list
};

int parse_index(const std::string& tok, size_t nv, uint32_t& out) {
    const char* s = tok.c_str();
    char* end = nullptr;
    const long v = std::strtol(s, &end, 10);
    if (end == s) return fail(VFS_ERR_IO, "OBJ: bad face index '" + tok + "'");
    long idx;
    if (v > 0) idx = v - 1;
    else if (v < 0) idx = (long)nv + v;
    else return fail(VFS_ERR_IO, "OBJ: face index 0");
    if (idx < 0 || (size_t)idx >= nv) return fail(VFS_ERR_IO, "OBJ: face index out of range '" + tok + "'");
    out = (uint32_t)idx;
    return 0;
}

int load_obj(const char* path, std::vector<float>& pos, std::vector<uint32_t>& tris) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return fail(VFS_ERR_IO, std::string("cannot open OBJ file '") + path + "'");
    std::vector<float> v;  // file vertex positions
    std::vector<ObjShape> shapes(1);
    std::string line;
    std::vector<uint32_t> poly;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        size_t b = line.find_first_not_of(" \t");
        if (b == std::string::npos || line[b] == '#') continue;
        std::istringstream ls(line.substr(b));
        std::string tag;
        ls >> tag;
        if (tag == "v") {
            double c[3];
            for (double& x : c) {
                std::string t;
                if (!(ls >> t)) return fail(VFS_ERR_IO, "OBJ: vertex with fewer than 3 coordinates");
                x = std::strtod(t.c_str(), nullptr);
            }
            for (double x : c) v.push_back((float)x);
        } else if (tag == "f") {
            poly.clear();
            std::string t;
            while (ls >> t) {
                uint32_t idx;
                VFS_TRY(parse_index(t.substr(0, t.find('/')), v.size() / 3, idx));
                poly.push_back(idx);
            }
            auto& out = shapes.back().tris;
            if (poly.size() < 3) continue;
            if (poly.size() == 4) {
                auto P = [&](int k, int a) { return v[3 * (size_t)poly[k] + a]; };
                float d02 = 0.0f, d13 = 0.0f;
                for (int a = 0; a < 3; a++) {
                    const float e02 = P(2, a) - P(0, a), e13 = P(3, a) - P(1, a);
                    d02 += e02 * e02;
                    d13 += e13 * e13;
                }
                const uint32_t q[6] = {poly[0], poly[1], d02 < d13 ? poly[2] : poly[3],
                                       d02 < d13 ? poly[0] : poly[1], poly[2], poly[3]};
                out.insert(out.end(), q, q + 6);
            } else {
                for (size_t k = 1; k + 1 < poly.size(); k++) {
                    out.push_back(poly[0]);
                    out.push_back(poly[k]);
                    out.push_back(poly[k + 1]);
                }
            }
        } else if (tag == "o" || tag == "g") {
            if (!shapes.back().tris.empty()) shapes.emplace_back();
        }
    }
    const ObjShape* last = nullptr;
    for (const auto& s : shapes)
        if (!s.tris.empty()) last = &s;
    if (!last) return fail(VFS_ERR_IO, std::string("OBJ file '") + path + "' has no faces");
    // the shape's vertices in first-use order
    std::unordered_map<uint32_t, uint32_t> remap;
    pos.clear();
    tris.clear();
    for (uint32_t fi : last->tris) {
        auto it = remap.find(fi);
        if (it == remap.end()) {
            it = remap.emplace(fi, (uint32_t)(pos.size() / 3)).first;
            pos.insert(pos.end(), v.begin() + 3 * (size_t)fi, v.begin() + 3 * (size_t)fi + 3);
        }
        tris.push_back(it->second);
    }
    return 0;
}

// ------------------------------------------------------------ pseudonormals
// float arithmetic of glm: dot = (x x' + y y') + z z', normalize = v * (1 / sqrt(dot))
struct F3 {
    float x, y, z;
};
inline F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float dot(F3 a, F3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline F3 normalize(F3 a) {
    const float r = 1.0f / std::sqrt(dot(a, a));
    return {a.x * r, a.y * r, a.z * r};
}
inline F3 cross(F3 a, F3 b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }

void build_normals(vfs_mesh& m) {
    const size_t nt = m.tris.size() / 3, nv = m.pos.size() / 3;
    auto P = [&](uint32_t k) { return F3{m.pos[3 * (size_t)k], m.pos[3 * (size_t)k + 1], m.pos[3 * (size_t)k + 2]}; };
    std::vector<F3> vn(nv, F3{0.0f, 0.0f, 0.0f}), fn(nt);
    std::unordered_map<uint64_t, F3> en;
    auto ekey = [](uint32_t a, uint32_t b) { return ((uint64_t)std::min(a, b) << 32) | std::max(a, b); };
    auto add = [](F3& acc, F3 v) {
        acc.x += v.x;
        acc.y += v.y;
        acc.z += v.z;
    };
    for (size_t t = 0; t < nt; t++) {
        const uint32_t ia = m.tris[3 * t], ib = m.tris[3 * t + 1], ic = m.tris[3 * t + 2];
        const F3 a = P(ia), b = P(ib), c = P(ic);
        const F3 n = normalize(cross(sub(b, a), sub(c, a)));
        fn[t] = n;
        // incident angles as the reference computes them: acos(|cos|)
        const float al0 = std::acos(std::fabs(dot(normalize(sub(b, a)), normalize(sub(c, a)))));
        const float al1 = std::acos(std::fabs(dot(normalize(sub(a, b)), normalize(sub(c, b)))));
        const float al2 = std::acos(std::fabs(dot(normalize(sub(b, c)), normalize(sub(a, c)))));
        add(vn[ia], F3{al0 * n.x, al0 * n.y, al0 * n.z});
        add(vn[ib], F3{al1 * n.x, al1 * n.y, al1 * n.z});
        add(vn[ic], F3{al2 * n.x, al2 * n.y, al2 * n.z});
        const uint64_t keys[3] = {ekey(ia, ib), ekey(ib, ic), ekey(ia, ic)};
        for (uint64_t k : keys) add(en.emplace(k, F3{0.0f, 0.0f, 0.0f}).first->second, n);
    }
    for (auto& v : vn) v = normalize(v);
    // properties that make the signed distance 1-Lipschitz (see MeshView)
    m.acute = true;
    for (size_t t = 0; t < nt && m.acute; t++) {
        const F3 a = P(m.tris[3 * t]), b = P(m.tris[3 * t + 1]), c = P(m.tris[3 * t + 2]);
        m.acute = dot(normalize(sub(b, a)), normalize(sub(c, a))) >= 0.0f &&
                  dot(normalize(sub(a, b)), normalize(sub(c, b))) >= 0.0f &&
                  dot(normalize(sub(b, c)), normalize(sub(a, c))) >= 0.0f;
    }
    std::unordered_map<uint64_t, int> directed;
    auto dkey = [](uint32_t a, uint32_t b) { return ((uint64_t)a << 32) | b; };
    for (size_t t = 0; t < nt; t++)
        for (int e = 0; e < 3; e++) directed[dkey(m.tris[3 * t + e], m.tris[3 * t + (e + 1) % 3])]++;
    m.closed = true;
    for (const auto& kv : directed) {
        const uint32_t a = (uint32_t)(kv.first >> 32), b = (uint32_t)kv.first;
        const auto it = directed.find(dkey(b, a));
        if (kv.second != 1 || a == b || it == directed.end() || it->second != 1) {
            m.closed = false;
            break;
        }
    }
    m.normals.resize(nt);
    auto put = [](float* d, F3 v) {
        d[0] = v.x;
        d[1] = v.y;
        d[2] = v.z;
    };
    for (size_t t = 0; t < nt; t++) {
        const uint32_t ia = m.tris[3 * t], ib = m.tris[3 * t + 1], ic = m.tris[3 * t + 2];
        TriNormals& tn = m.normals[t];
        put(tn.face, fn[t]);
        put(tn.edge[0], normalize(en.at(ekey(ia, ib))));
        put(tn.edge[1], normalize(en.at(ekey(ib, ic))));
        put(tn.edge[2], normalize(en.at(ekey(ia, ic))));
        put(tn.vert[0], vn[ia]);
        put(tn.vert[1], vn[ib]);
        put(tn.vert[2], vn[ic]);
    }
}

// ------------------------------------------------------------------ BVH
// Median split along the longest axis of the centroid bounds, at most four
// triangles per leaf, children of a node stored next to each other.  Any
// hierarchy gives the reference's result (the search is exhaustive up to
// pruning by distance); this one keeps the depth at log2(n / 4) + 1.
void build_bvh(vfs_mesh& m) {
    const uint32_t nt = (uint32_t)(m.tris.size() / 3);
    std::vector<float> cen(3 * (size_t)nt);
    for (uint32_t t = 0; t < nt; t++)
        for (int a = 0; a < 3; a++) {
            const float s = m.pos[3 * (size_t)m.tris[3 * t] + a] + m.pos[3 * (size_t)m.tris[3 * t + 1] + a] +
                            m.pos[3 * (size_t)m.tris[3 * t + 2] + a];
            cen[3 * (size_t)t + a] = s / 3.0f;
        }
    m.order.resize(nt);
    for (uint32_t t = 0; t < nt; t++) m.order[t] = t;
    m.nodes.clear();
    m.nodes.push_back(MeshNode{{0, 0, 0}, 0u, {0, 0, 0}, nt});
    struct Work {
        uint32_t node, depth;
    };
    std::vector<Work> work{{0u, 1u}};
    m.depth = 1;
    while (!work.empty()) {
        const Work w = work.back();
        work.pop_back();
        MeshNode nd = m.nodes[w.node];
        const uint32_t first = nd.first, count = nd.count;
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t k = first; k < first + count; k++) {
            const uint32_t t = m.order[k];
            for (int c = 0; c < 3; c++)
                for (int a = 0; a < 3; a++) {
                    const float x = m.pos[3 * (size_t)m.tris[3 * t + c] + a];
                    lo[a] = std::min(lo[a], x);
                    hi[a] = std::max(hi[a], x);
                }
            for (int a = 0; a < 3; a++) {
                clo[a] = std::min(clo[a], cen[3 * (size_t)t + a]);
                chi[a] = std::max(chi[a], cen[3 * (size_t)t + a]);
            }
        }
        for (int a = 0; a < 3; a++) {
            nd.lo[a] = lo[a];
            nd.hi[a] = hi[a];
        }
        if (count > 4) {
            int axis = 0;
            for (int a = 1; a < 3; a++)
                if (chi[a] - clo[a] > chi[axis] - clo[axis]) axis = a;
            const uint32_t mid = first + count / 2;
            std::nth_element(m.order.begin() + first, m.order.begin() + mid, m.order.begin() + first + count,
                             [&](uint32_t p, uint32_t q) {
                                 const float cp = cen[3 * (size_t)p + axis], cq = cen[3 * (size_t)q + axis];
                                 return cp < cq || (cp == cq && p < q);
                             });
            const uint32_t child = (uint32_t)m.nodes.size();
            m.nodes.push_back(MeshNode{{0, 0, 0}, first, {0, 0, 0}, mid - first});
            m.nodes.push_back(MeshNode{{0, 0, 0}, mid, {0, 0, 0}, first + count - mid});
            nd.first = child;
            nd.count = 0;
            work.push_back({child + 1, w.depth + 1});
            work.push_back({child, w.depth + 1});
            m.depth = std::max(m.depth, w.depth + 1);
        }
        m.nodes[w.node] = nd;
    }
    for (int a = 0; a < 3; a++) {
        m.lo[a] = m.nodes[0].lo[a];
        m.hi[a] = m.nodes[0].hi[a];
    }
}

int finish_mesh(vfs_mesh* m) {
    const size_t nt = m->tris.size() / 3;
    VFS_CHECK(nt >= 1, VFS_ERR_INVALID, "mesh has no triangles");
    VFS_CHECK(nt < (1u << 28), VFS_ERR_INVALID, "mesh has too many triangles");
    build_normals(*m);
    build_bvh(*m);
    VFS_CHECK(m->depth < (uint32_t)kMeshStack, VFS_ERR_INVALID, "mesh hierarchy too deep");
    return 0;
}

void free_device(vfs_mesh* m) {
    void* p[] = {m->d_nodes, m->d_tri, m->d_pn};
    for (void* q : p)
        if (q) (void)hipFree(q);
    m->d_nodes = nullptr;
    m->d_tri = nullptr;
    m->d_pn = nullptr;
}

// Device copy (first GPU use): nodes, triangle corners and pseudonormals in
// BVH order.
int upload(vfs_mesh* m, MeshView& view) {
    const size_t nt = m->tris.size() / 3;
    if (!m->d_nodes) {
        std::vector<float4> tri(3 * nt);
        std::vector<TriNormals> pn(nt);
        for (size_t k = 0; k < nt; k++) {
            const uint32_t t = m->order[k];
            for (int c = 0; c < 3; c++) {
                const float* p = &m->pos[3 * (size_t)m->tris[3 * t + c]];
                tri[3 * k + c] = make_float4(p[0], p[1], p[2], 0.0f);
            }
            // the mesh index rides in the first corner's w (the search's tie-break)
            std::memcpy(&tri[3 * k].w, &t, sizeof(t));
            pn[k] = m->normals[t];
        }
        if (hipMalloc(&m->d_nodes, m->nodes.size() * sizeof(MeshNode)) != hipSuccess ||
            hipMalloc(&m->d_tri, tri.size() * sizeof(float4)) != hipSuccess ||
            hipMalloc(&m->d_pn, pn.size() * sizeof(TriNormals)) != hipSuccess) {
            free_device(m);
            return fail(VFS_ERR_NOMEM, "mesh upload: hipMalloc failed");
        }
        if (hipMemcpy(m->d_nodes, m->nodes.data(), m->nodes.size() * sizeof(MeshNode), hipMemcpyHostToDevice) !=
                hipSuccess ||
            hipMemcpy(m->d_tri, tri.data(), tri.size() * sizeof(float4), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(m->d_pn, pn.data(), pn.size() * sizeof(TriNormals), hipMemcpyHostToDevice) != hipSuccess) {
            free_device(m);
            return fail(VFS_ERR_HIP, "mesh upload: copy failed");
        }
    }
    view.nodes = m->d_nodes;
    view.tri = m->d_tri;
    view.pn = m->d_pn;
    double d2 = 0.0;
    for (int a = 0; a < 3; a++) {
        view.centre[a] = 0.5f * (m->lo[a] + m->hi[a]);
        const double e = (double)m->hi[a] - (double)m->lo[a];
        d2 += e * e;
    }
    view.diag2 = d2;
    view.lipschitz = (m->closed && m->acute) ? 1u : 0u;
    return 0;
}

}  // namespace

extern "C" {

int vfs_mesh_create(const float* vertices, uint32_t n_vertices, const uint32_t* triangles, uint32_t n_triangles,
                    vfs_mesh** out) {
    VFS_CHECK(out, VFS_ERR_INVALID, "out is null");
    *out = nullptr;
    VFS_CHECK(vertices && triangles, VFS_ERR_INVALID, "null argument");
    VFS_CHECK(n_triangles >= 1, VFS_ERR_INVALID, "mesh has no triangles");
    for (uint64_t k = 0; k < 3ull * n_triangles; k++)
        VFS_CHECK(triangles[k] < n_vertices, VFS_ERR_INVALID, "triangle index out of range");
    auto* m = new vfs_mesh();
    m->pos.assign(vertices, vertices + 3 * (size_t)n_vertices);
    m->tris.assign(triangles, triangles + 3 * (size_t)n_triangles);
    const int r = finish_mesh(m);
    if (r) {
        delete m;
        return r;
    }
    *out = m;
    return 0;
}

int vfs_mesh_load_obj(const char* path, vfs_mesh** out) {
    VFS_CHECK(out, VFS_ERR_INVALID, "out is null");
    *out = nullptr;
    VFS_CHECK(path, VFS_ERR_INVALID, "path is null");
    auto* m = new vfs_mesh();
    int r = load_obj(path, m->pos, m->tris);
    if (r == 0) r = finish_mesh(m);
    if (r) {
        delete m;
        return r;
    }
    *out = m;
    return 0;
}

int vfs_mesh_info(const vfs_mesh* mesh, uint32_t* n_vertices, uint32_t* n_triangles, float* aabb_min,
                  float* aabb_max) {
    VFS_CHECK(mesh, VFS_ERR_INVALID, "mesh is null");
    if (n_vertices) *n_vertices = (uint32_t)(mesh->pos.size() / 3);
    if (n_triangles) *n_triangles = (uint32_t)(mesh->tris.size() / 3);
    for (int a = 0; a < 3; a++) {
        if (aabb_min) aabb_min[a] = mesh->lo[a];
        if (aabb_max) aabb_max[a] = mesh->hi[a];
    }
    return 0;
}

int vfs_mesh_properties(const vfs_mesh* mesh, int* closed, int* acute) {
    VFS_CHECK(mesh, VFS_ERR_INVALID, "mesh is null");
    if (closed) *closed = mesh->closed ? 1 : 0;
    if (acute) *acute = mesh->acute ? 1 : 0;
    return 0;
}

int vfs_mesh_get(const vfs_mesh* mesh, float* vertices, uint32_t* triangles) {
    VFS_CHECK(mesh, VFS_ERR_INVALID, "mesh is null");
    if (vertices) std::memcpy(vertices, mesh->pos.data(), mesh->pos.size() * sizeof(float));
    if (triangles) std::memcpy(triangles, mesh->tris.data(), mesh->tris.size() * sizeof(uint32_t));
    return 0;
}

int vfs_mesh_signed_distance(vfs_mesh* mesh, const float* points, uint32_t n, double* out) {
    VFS_CHECK(mesh, VFS_ERR_INVALID, "mesh is null");
    if (n == 0) return 0;
    VFS_CHECK(points && out, VFS_ERR_INVALID, "null argument");
    MeshView view;
    VFS_TRY(upload(mesh, view));
    float* d_pts = nullptr;
    double* d_out = nullptr;
    VFS_HIP(hipMalloc(&d_pts, (size_t)n * 12));
    if (hipMalloc(&d_out, (size_t)n * 8) != hipSuccess) {
        (void)hipFree(d_pts);
        return fail(VFS_ERR_NOMEM, "hipMalloc failed");
    }
    int r = 0;
    if (hipMemcpy(d_pts, points, (size_t)n * 12, hipMemcpyHostToDevice) != hipSuccess) r = fail(VFS_ERR_HIP, "copy");
    if (r == 0) r = launch_mesh_signed_distance(view, d_pts, n, d_out, nullptr);
    if (r == 0 && hipMemcpy(out, d_out, (size_t)n * 8, hipMemcpyDeviceToHost) != hipSuccess)
        r = fail(VFS_ERR_HIP, "mesh signed distance failed");
    (void)hipFree(d_pts);
    (void)hipFree(d_out);
    return r;
}

int vfs_mesh_volume_map_generate(vfs_mesh* mesh, const uint32_t* res, double support_radius, double tolerance,
                                 double margin, float* sdf_out, float* vm_out, float* box_size, float* box_pos) {
    VFS_CHECK(mesh && res && sdf_out && vm_out && box_size && box_pos, VFS_ERR_INVALID, "null argument");
    VFS_CHECK(res[0] >= 2 && res[1] >= 2 && res[2] >= 2, VFS_ERR_INVALID, "resolution must be >= 2 per axis");
    VFS_CHECK(support_radius > 0.0, VFS_ERR_INVALID, "support radius must be positive");
    const uint64_t total = (uint64_t)res[0] * res[1] * res[2];
    VFS_CHECK(total < (1ull << 31), VFS_ERR_INVALID, "grid too large");
    float dmin[3];
    double step[3];
    grid_domain(mesh->lo, mesh->hi, tolerance, margin, res, box_size, box_pos, dmin, step);
    double gx[16], gw[16];
    gauss_legendre(16, gx, gw);
    MeshView view;
    VFS_TRY(upload(mesh, view));
    float *d_sdf = nullptr, *d_vm = nullptr;
    VFS_HIP(hipMalloc(&d_sdf, total * 4));
    if (hipMalloc(&d_vm, total * 4) != hipSuccess) {
        (void)hipFree(d_sdf);
        return fail(VFS_ERR_NOMEM, "hipMalloc failed");
    }
    int r = launch_mesh_volume_map(view, res, support_radius, tolerance, dmin, step, gx, gw, d_sdf, d_vm, nullptr);
    if (r == 0 && hipDeviceSynchronize() != hipSuccess) r = fail(VFS_ERR_HIP, "mesh volume map kernel failed");
    if (r == 0 && (hipMemcpy(sdf_out, d_sdf, total * 4, hipMemcpyDeviceToHost) != hipSuccess ||
                   hipMemcpy(vm_out, d_vm, total * 4, hipMemcpyDeviceToHost) != hipSuccess))
        r = fail(VFS_ERR_HIP, "copy failed");
    (void)hipFree(d_sdf);
    (void)hipFree(d_vm);
    return r;
}

int vfs_mesh_destroy(vfs_mesh* mesh) {
    if (mesh) {
        free_device(mesh);
        delete mesh;
    }
    return 0;
}

}  // extern "C"
