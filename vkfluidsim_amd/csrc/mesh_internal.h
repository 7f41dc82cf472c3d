// Triangle meshes as boundary objects: the host mesh (positions, triangles,
// pseudonormals, bounding-volume hierarchy) and its device copy.
//
// Reference: src/util/mesh_loader.cpp:20-113 (OBJ), mesh_pseudonormals.cpp:11-103,
// mesh_bvh.cpp:272-314 and geometry.cpp:37-286 (closest point, signed distance),
// mesh_sdf.cpp:17-45 (grid domain).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "vfs_common.h"

namespace vfs {

// BVH node: box of its triangles; a leaf holds `count` triangles from `first`,
// an inner node (count == 0) has its children at `first` and `first + 1`.
struct MeshNode {
    float lo[3];
    uint32_t first;
    float hi[3];
    uint32_t count;
};

// Pseudonormals of one triangle (BVH order): face, edges 01 / 12 / 02,
// vertices 0 / 1 / 2, as in TriangleClosestEntity.
struct TriNormals {
    float face[3];
    float edge[3][3];
    float vert[3][3];
};

// Device view used by the kernels.  centre/diag2 (root box centre, squared
// diagonal) scale the pruning slack of the closest-point search.
// lipschitz: the mesh is closed and consistently oriented and no triangle
// angle exceeds 90 degrees, so the reference's pseudonormals (weights
// acos|cos|) are the angle-weighted ones and the signed distance they give is
// the true one, 1-Lipschitz in the query point.
struct MeshView {
    const MeshNode* nodes;
    const float4* tri;  // 3 per triangle: v0, v1, v2 (xyz); v0.w holds the triangle's mesh index (bits)
    const TriNormals* pn;
    float centre[3];
    double diag2;
    uint32_t lipschitz;
};

// Grid domain of a boundary object (mesh_sdf.cpp:33-38): the object's box grown
// by tolerance + margin; node spacing as the float the reference computes
// (discretization.cpp:22-23), widened to double.
inline void grid_domain(const float* lo, const float* hi, double tolerance, double margin, const uint32_t* res,
                        float* box_size, float* box_pos, float* dmin, double* step) {
    const double total_margin = tolerance + margin;
    for (int a = 0; a < 3; a++) {
        box_size[a] = (hi[a] - lo[a]) + 2.0f * (float)total_margin;
        box_pos[a] = lo[a] - (float)total_margin;
        dmin[a] = box_pos[a];
        const float dmax = box_pos[a] + box_size[a];
        step[a] = (double)((dmax - dmin[a]) / (float)(res[a] - 1u));
    }
}

constexpr int kMeshStack = 64;  // traversal stack; the build keeps the depth below it

int launch_mesh_signed_distance(const MeshView& m, const float* points, uint32_t n, double* out, hipStream_t s);
int launch_mesh_volume_map(const MeshView& m, const uint32_t* res, double support_radius, double tolerance,
                           const float* dmin, const double* step, const double* gl_x, const double* gl_w,
                           float* sdf_out, float* vm_out, hipStream_t s);
void gauss_legendre(int n, double* x, double* w);

}  // namespace vfs

struct vfs_mesh {
    std::vector<float> pos;      // 3 per vertex
    std::vector<uint32_t> tris;  // 3 per triangle (input order)
    // built
    std::vector<vfs::MeshNode> nodes;
    std::vector<uint32_t> order;  // BVH order -> input triangle
    std::vector<vfs::TriNormals> normals;  // input order
    uint32_t depth = 0;
    float lo[3]{}, hi[3]{};
    bool closed = false;  // every edge in two triangles, traversed in opposite directions
    bool acute = false;   // no triangle angle above 90 degrees
    // device copy
    vfs::MeshNode* d_nodes = nullptr;
    float4* d_tri = nullptr;
    vfs::TriNormals* d_pn = nullptr;
};
