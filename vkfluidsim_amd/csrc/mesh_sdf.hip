// Signed distance to a triangle mesh and mesh volume maps on the GPU.
//
// The reference evaluates the boundary grids of a mesh on the CPU
// (src/util/mesh_sdf.cpp:28-45, volume_map.cpp:9-54): for every grid node, one
// signed-distance query for the SDF and up to 16^3 more for the volume-map
// quadrature, each a closest-point search over a BVH (mesh_bvh.cpp:272-314)
// with the point-triangle distance of geometry.cpp:37-217 in double and the
// sign of geometry.cpp:239-286 (angle-weighted pseudonormals).  Here one
// thread evaluates one node, sequentially over its quadrature points so that
// the sum order is the reference's, and each query walks the BVH from a small
// per-lane stack.
#include "mesh_internal.h"
#include "volume_map.h"

namespace vfs {

namespace {

enum Entity : int { kV0 = 0, kV1, kV2, kE01, kE12, kE02, kF };

struct TriDistance {
    double d2;
    double s, t;
    int entity;
};

__device__ __forceinline__ double ddot(double ax, double ay, double az, double bx, double by, double bz) {
    return (ax * bx + ay * by) + az * bz;
}

// Squared distance from p to the triangle (v0, v1, v2), the closest point as
// v0 + s e0 + t e1 and the closest feature: Eberly's region classification of
// the minimiser of |v0 + s e0 + t e1 - p|^2 over s, t >= 0, s + t <= 1
// (geometry.cpp:37-217), with its expressions for each region.
__device__ TriDistance tri_distance(const double* p, const double* v0, const double* e0, const double* e1) {
    const double dx = v0[0] - p[0], dy = v0[1] - p[1], dz = v0[2] - p[2];
    const double a00 = ddot(e0[0], e0[1], e0[2], e0[0], e0[1], e0[2]);
    const double a01 = ddot(e0[0], e0[1], e0[2], e1[0], e1[1], e1[2]);
    const double a11 = ddot(e1[0], e1[1], e1[2], e1[0], e1[1], e1[2]);
    const double b0 = ddot(dx, dy, dz, e0[0], e0[1], e0[2]);
    const double b1 = ddot(dx, dy, dz, e1[0], e1[1], e1[2]);
    const double c = ddot(dx, dy, dz, dx, dy, dz);
    const double det = fabs(a00 * a11 - a01 * a01);
    double s = a01 * b1 - a11 * b0;
    double t = a01 * b0 - a00 * b1;

    // closed forms of the boundary cases
    auto at_v0 = [&](TriDistance& r) { r.entity = kV0; r.s = 0.0; r.t = 0.0; r.d2 = c; };
    auto at_v1 = [&](TriDistance& r) { r.entity = kV1; r.s = 1.0; r.t = 0.0; r.d2 = a00 + 2.0 * b0 + c; };
    auto at_v2 = [&](TriDistance& r) { r.entity = kV2; r.s = 0.0; r.t = 1.0; r.d2 = a11 + 2.0 * b1 + c; };
    auto on_e01 = [&](TriDistance& r) {
        r.entity = kE01;
        r.t = 0.0;
        r.s = -b0 / a00;
        r.d2 = b0 * r.s + c;
    };
    auto on_e02 = [&](TriDistance& r) {
        r.entity = kE02;
        r.s = 0.0;
        r.t = -b1 / a11;
        r.d2 = b1 * r.t + c;
    };
    auto quad = [&](double ss, double tt) {
        return ss * (a00 * ss + a01 * tt + 2.0 * b0) + tt * (a01 * ss + a11 * tt + 2.0 * b1) + c;
    };
    // s = 0 edge (v0 .. v2) for the regions left of it
    auto left = [&](TriDistance& r) {
        if (b1 >= 0.0) at_v0(r);
        else if (-b1 >= a11) at_v2(r);
        else on_e02(r);
    };
    // t = 0 edge (v0 .. v1) for the regions below it
    auto below = [&](TriDistance& r) {
        if (b0 >= 0.0) at_v0(r);
        else if (-b0 >= a00) at_v1(r);
        else on_e01(r);
    };

    TriDistance r;
    if (s + t <= det) {
        if (s < 0.0) {
            if (t < 0.0) {
                if (b0 < 0.0) {
                    if (-b0 >= a00) at_v1(r);
                    else on_e01(r);
                } else {
                    left(r);
                }
            } else {
                left(r);
            }
        } else if (t < 0.0) {
            below(r);
        } else {
            const double inv_det = 1.0 / det;
            s *= inv_det;
            t *= inv_det;
            r.entity = kF;
            r.s = s;
            r.t = t;
            r.d2 = quad(s, t);
        }
    } else {
        const double denom = a00 - 2.0 * a01 + a11;
        if (s < 0.0) {
            const double tmp0 = a01 + b0, tmp1 = a11 + b1;
            if (tmp1 > tmp0) {
                const double numer = tmp1 - tmp0;
                if (numer >= denom) {
                    at_v1(r);
                } else {
                    r.entity = kE12;
                    r.s = numer / denom;
                    r.t = 1.0 - r.s;
                    r.d2 = quad(r.s, r.t);
                }
            } else if (tmp1 <= 0.0) {
                at_v2(r);
            } else if (b1 >= 0.0) {
                at_v0(r);
            } else {
                on_e02(r);
            }
        } else if (t < 0.0) {
            const double tmp0 = a01 + b1, tmp1 = a00 + b0;
            if (tmp1 > tmp0) {
                const double numer = tmp1 - tmp0;
                if (numer >= denom) {
                    at_v2(r);
                } else {
                    r.entity = kE12;
                    r.t = numer / denom;
                    r.s = 1.0 - r.t;
                    r.d2 = quad(r.s, r.t);
                }
            } else if (tmp1 <= 0.0) {
                at_v1(r);
            } else if (b0 >= 0.0) {
                at_v0(r);
            } else {
                on_e01(r);
            }
        } else {
            const double numer = a11 + b1 - a01 - b0;
            if (numer <= 0.0) {
                at_v2(r);
            } else if (numer >= denom) {
                at_v1(r);
            } else {
                r.entity = kE12;
                r.s = numer / denom;
                r.t = 1.0 - r.s;
                r.d2 = quad(r.s, r.t);
            }
        }
    }
    if (r.d2 < 0.0) r.d2 = 0.0;  // round-off
    return r;
}

__device__ __forceinline__ double box_d2(const MeshNode& nd, const double* q) {
    double d2 = 0.0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const double lo = nd.lo[a], hi = nd.hi[a];
        const double d = q[a] < lo ? lo - q[a] : (q[a] > hi ? q[a] - hi : 0.0);
        d2 += d * d;
    }
    return d2;
}

struct Closest {
    double d2;
    float p[3];
    int entity;
    uint32_t tri;  // BVH position
    uint32_t id;   // index of the triangle in the mesh
};

// Distance from q to triangle k (BVH order); keeps it in `best` if closer, or
// as close and earlier in the mesh (the first corner's w holds the triangle's
// mesh index), so that exactly equal distances resolve to the lowest index
// whatever order the search visits them in.
__device__ __forceinline__ void try_triangle(const MeshView& m, const double* q, uint32_t k, Closest& best) {
    const float4 a = m.tri[3 * (size_t)k], b = m.tri[3 * (size_t)k + 1], cc = m.tri[3 * (size_t)k + 2];
    const double v0[3] = {a.x, a.y, a.z};
    const double e0[3] = {(double)b.x - v0[0], (double)b.y - v0[1], (double)b.z - v0[2]};
    const double e1[3] = {(double)cc.x - v0[0], (double)cc.y - v0[1], (double)cc.z - v0[2]};
    const TriDistance td = tri_distance(q, v0, e0, e1);
    const uint32_t id = __float_as_uint(a.w);
    if (td.d2 < best.d2 || (td.d2 == best.d2 && id < best.id)) {
        best.d2 = td.d2;
        best.entity = td.entity;
        best.tri = k;
        best.id = id;
        // closest point v0 + s e0 + t e1, stored as float (ClosestPointQueryResult)
        best.p[0] = (float)(v0[0] + td.s * e0[0] + td.t * e1[0]);
        best.p[1] = (float)(v0[1] + td.s * e0[1] + td.t * e1[1]);
        best.p[2] = (float)(v0[2] + td.s * e0[2] + td.t * e1[2]);
    }
}

constexpr uint32_t kNoHint = 0xffffffffu;

// Minimum over all triangles of the computed squared distance.  `hint` (a
// triangle near the query, e.g. the previous query's closest) is tried first
// so that the search starts with a tight bound; among exactly equal
// distances the lowest mesh index is kept.  A subtree is skipped only when its box
// is farther than the best distance plus a slack far above the rounding of
// the triangle distances, so no triangle whose computed distance beats the
// best is ever skipped.
__device__ Closest mesh_closest(const MeshView& m, float qx, float qy, float qz, uint32_t hint = kNoHint) {
    const double q[3] = {qx, qy, qz};
    const double cx = q[0] - m.centre[0], cy = q[1] - m.centre[1], cz = q[2] - m.centre[2];
    const double slack = 1e-12 * (m.diag2 + (cx * cx + cy * cy) + cz * cz);
    Closest best;
    best.d2 = (double)3.402823466e38f;  // ClosestPointQueryResult's initial FLT_MAX
    best.p[0] = best.p[1] = best.p[2] = 0.0f;
    best.entity = kF;
    best.tri = 0;
    best.id = 0xffffffffu;
    if (hint != kNoHint) try_triangle(m, q, hint, best);
    uint32_t stack[kMeshStack];
    int sp = 0;
    uint32_t node = 0;
    while (true) {
        const MeshNode nd = m.nodes[node];
        bool descend = false;
        if (nd.count) {
            for (uint32_t k = nd.first; k < nd.first + nd.count; k++) try_triangle(m, q, k, best);
        } else {
            const uint32_t a = nd.first, b = nd.first + 1u;
            const double da = box_d2(m.nodes[a], q), db = box_d2(m.nodes[b], q);
            // nearer child first (the reference's order, ties to the second)
            const bool a_first = da < db;
            const uint32_t near = a_first ? a : b, far = a_first ? b : a;
            const double dn = a_first ? da : db, df = a_first ? db : da;
            const bool vn = dn < best.d2 + slack, vf = df < best.d2 + slack;
            if (vn) {
                if (vf && sp < kMeshStack) stack[sp++] = far;
                node = near;
                descend = true;
            } else if (vf) {
                node = far;
                descend = true;
            }
        }
        if (descend) continue;
        bool found = false;
        while (sp > 0) {
            node = stack[--sp];
            if (box_d2(m.nodes[node], q) < best.d2 + slack) {
                found = true;
                break;
            }
        }
        if (!found) break;
    }
    return best;
}

// SignedDistanceToMesh (geometry.cpp:239-286): the distance, signed by the
// pseudonormal of the closest feature against (query - closest point) in
// float.  `hint` is updated to the closest triangle.
__device__ double mesh_signed_distance(const MeshView& m, float qx, float qy, float qz, uint32_t& hint) {
    const Closest r = mesh_closest(m, qx, qy, qz, hint);
    hint = r.tri;
    const TriNormals& pn = m.pn[r.tri];
    const float* n;
    switch (r.entity) {
        case kV0: n = pn.vert[0]; break;
        case kV1: n = pn.vert[1]; break;
        case kV2: n = pn.vert[2]; break;
        case kE01: n = pn.edge[0]; break;
        case kE12: n = pn.edge[1]; break;
        case kE02: n = pn.edge[2]; break;
        default: n = pn.face; break;
    }
    const float ux = qx - r.p[0], uy = qy - r.p[1], uz = qz - r.p[2];
    const float dt = (ux * n[0] + uy * n[1]) + uz * n[2];
    const double sign = dt >= 0.0f ? 1.0 : -1.0;
    return sign * sqrt(r.d2);
}

__global__ __launch_bounds__(256) void k_mesh_signed_distance(MeshView m, const float* __restrict__ pts, uint32_t n,
                                                              double* __restrict__ out) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    uint32_t hint = kNoHint;
    out[i] = mesh_signed_distance(m, pts[3 * (size_t)i], pts[3 * (size_t)i + 1], pts[3 * (size_t)i + 2], hint);
}

// One grid node per thread: the SDF value and the volume-map integral
// (volume_map.cpp:23-52, the reference's loop order and expressions).  When
// the signed distance is 1-Lipschitz (MeshView::lipschitz), a quadrature
// point x + xi with d(x) + |xi| < -delta is inside (value 1) and one with
// d(x) - |xi| > radius + delta has value 0, whatever the search would find;
// delta covers the float rounding of x + xi.  Every other point is searched,
// starting from the previous point's closest triangle.
__global__ __launch_bounds__(128) void k_mesh_volume_map(MeshView m, uint32_t rx, uint32_t ry, uint32_t rz, float sr,
                                                         double tolerance, float d0, float d1, float d2, double s0,
                                                         double s1, double s2, GLTable gl,
                                                         float* __restrict__ sdf_out, float* __restrict__ vm_out) {
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t total = (uint64_t)rx * ry * rz;
    if (idx >= total) return;
    float x[3];
    grid_node(idx, rx, ry, d0, d1, d2, s0, s1, s2, x);
    uint32_t hint = kNoHint;
    const double dist = mesh_signed_distance(m, x[0], x[1], x[2], hint);
    sdf_out[idx] = (float)(dist - tolerance);
    if (dist > 2.0 * (double)sr) {
        vm_out[idx] = 0.0f;
        return;
    }
    const double radius = (double)sr;
    const double factor = 8.0 / (M_PI * radius * radius * radius);
    const double w_zero = kernel_w_f64(0.0, radius, factor);
    const double c0 = 0.5 * (double)(sr - (-sr));
    const double c1 = 0.5 * (double)(sr + (-sr));
    const float sr2 = sr * sr;
    const double xmax = fmax(fabs((double)x[0]), fmax(fabs((double)x[1]), fabs((double)x[2])));
    const double delta = 8.0 * 1.1920928955078125e-07 * (xmax + radius) + 1e-12;
    const bool lip = m.lipschitz != 0u;
    double integral = 0.0;
    for (int a = 0; a < 16; a++) {
        const double wi = gl.w[a];
        const float xa = (float)(c0 * gl.x[a] + c1);
        for (int bb = 0; bb < 16; bb++) {
            const double wij = wi * gl.w[bb];
            const float xb = (float)(c0 * gl.x[bb] + c1);
            for (int cc = 0; cc < 16; cc++) {
                const double wijk = wij * gl.w[cc];
                const float xc = (float)(c0 * gl.x[cc] + c1);
                double val;
                if ((xa * xa + xb * xb) + xc * xc > sr2) {
                    val = 0.0;
                } else {
                    const double r = sqrt((double)xa * xa + (double)xb * xb + (double)xc * xc);
                    if (lip && dist + r < -delta) {
                        val = 1.0;
                    } else if (lip && dist - r > radius + delta) {
                        val = 0.0;
                    } else {
                        const double di = mesh_signed_distance(m, x[0] + xa, x[1] + xb, x[2] + xc, hint);
                        if (di <= 0.0)
                            val = 1.0;
                        else if (di < radius)
                            val = kernel_w_f64(di, radius, factor) / w_zero;
                        else
                            val = 0.0;
                    }
                }
                integral += wijk * val;
            }
        }
    }
    integral *= c0 * c0 * c0;
    vm_out[idx] = (float)(0.8 * integral);
}

inline uint32_t blocks_of(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

}  // namespace

int launch_mesh_signed_distance(const MeshView& m, const float* points, uint32_t n, double* out, hipStream_t s) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_mesh_signed_distance, dim3(blocks_of(n, 256)), dim3(256), 0, s, m, points, n, out);
    return check_launch("mesh_signed_distance");
}

int launch_mesh_volume_map(const MeshView& m, const uint32_t* res, double support_radius, double tolerance,
                           const float* dmin, const double* step, const double* gl_x, const double* gl_w,
                           float* sdf_out, float* vm_out, hipStream_t s) {
    GLTable gl;
    for (int i = 0; i < 16; i++) {
        gl.x[i] = gl_x[i];
        gl.w[i] = gl_w[i];
    }
    const uint64_t total = (uint64_t)res[0] * res[1] * res[2];
    hipLaunchKernelGGL(k_mesh_volume_map, dim3(blocks_of(total, 128)), dim3(128), 0, s, m, res[0], res[1], res[2],
                       (float)support_radius, tolerance, dmin[0], dmin[1], dmin[2], step[0], step[1], step[2], gl,
                       sdf_out, vm_out);
    return check_launch("mesh_volume_map");
}

}  // namespace vfs
