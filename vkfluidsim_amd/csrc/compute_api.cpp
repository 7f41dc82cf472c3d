// C ABI of the compute-layer drop-ins (reference src/compute/*.h) and of the
// analytic-shape volume-map generator (reference src/util/volume_map.cpp).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "mesh_internal.h"
#include "sort_internal.h"

using namespace vfs;

struct vfs_spatial_hash {
    uint32_t n = 0;
    float cell_size = 0.0f;
    uint32_t *keys = nullptr, *indices = nullptr, *offsets = nullptr, *raw_keys = nullptr, *ws = nullptr;
    SortPlan plan{};
};

namespace {

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

void destroy_hash(vfs_spatial_hash* h) {
    if (!h) return;
    uint32_t* ptrs[] = {h->keys, h->indices, h->offsets, h->raw_keys, h->ws};
    for (uint32_t* p : ptrs)
        if (p) (void)hipFree(p);
    delete h;
}

}  // namespace

namespace vfs {

// Newton iteration on the Legendre polynomial P_n (roots) with the standard
// weight formula w = 2 / ((1 - x^2) P_n'(x)^2).
void gauss_legendre(int n, double* x, double* w) {
    for (int i = 0; i < (n + 1) / 2; i++) {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
        double dp = 0.0;
        for (int it = 0; it < 100; it++) {
            double p0 = 1.0, p1 = 0.0;
            for (int j = 1; j <= n; j++) {
                const double pm = p1;
                p1 = p0;
                p0 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * pm) / j;
            }
            dp = n * (z * p0 - p1) / (z * z - 1.0);
            const double prev = z;
            z = prev - p0 / dp;
            if (std::fabs(z - prev) < 1e-16) break;
        }
        x[i] = -z;
        x[n - 1 - i] = z;
        w[i] = w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

}  // namespace vfs

extern "C" {

int vfs_spatial_hash_create(uint32_t n, float cell_size, vfs_spatial_hash** out) {
    VFS_CHECK(out, VFS_ERR_INVALID, "out is null");
    *out = nullptr;
    VFS_CHECK(n < (1u << 30), VFS_ERR_INVALID, "n must be < 2^30");
    VFS_CHECK(cell_size > 0.0f && std::isfinite(cell_size), VFS_ERR_INVALID, "cell_size must be positive");
    auto* h = new vfs_spatial_hash();
    h->n = n;
    h->cell_size = cell_size;
    h->plan = make_sort_plan(n, n ? n - 1 : 0, 8, 16);
    const size_t m = n ? n : 1;
    if (hipMalloc(&h->keys, m * 4) != hipSuccess || hipMalloc(&h->indices, m * 4) != hipSuccess ||
        hipMalloc(&h->offsets, m * 4) != hipSuccess || hipMalloc(&h->raw_keys, m * 4) != hipSuccess ||
        hipMalloc(&h->ws, std::max<size_t>(1, h->plan.workspace_words) * 4) != hipSuccess) {
        destroy_hash(h);
        return fail(VFS_ERR_NOMEM, "hipMalloc failed");
    }
    *out = h;
    return 0;
}

int vfs_spatial_hash_set_cell_size(vfs_spatial_hash* h, float cell_size) {
    VFS_CHECK(h, VFS_ERR_INVALID, "hash is null");
    VFS_CHECK(cell_size > 0.0f && std::isfinite(cell_size), VFS_ERR_INVALID, "cell_size must be positive");
    h->cell_size = cell_size;
    return 0;
}

// SpatialHash::Run (spatial_hash.cpp:52-63): hash -> sort -> offsets(init).
int vfs_spatial_hash_run(vfs_spatial_hash* h, const float* positions_dev, uint32_t stride_floats, void* stream) {
    VFS_CHECK(h, VFS_ERR_INVALID, "hash is null");
    VFS_CHECK(stride_floats == 3 || stride_floats == 4, VFS_ERR_INVALID, "stride must be 3 or 4 floats");
    if (h->n == 0) return 0;
    VFS_CHECK(positions_dev, VFS_ERR_INVALID, "positions is null");
    hipStream_t s = as_stream(stream);
    VFS_TRY(launch_hash_keys(positions_dev, stride_floats, h->n, h->cell_size, h->n, h->raw_keys, s));
    // InitOffsets is folded into the sort's preparation kernel (offsets[k] = n)
    VFS_TRY(sort_keys_iota(h->plan, h->raw_keys, h->keys, h->indices, h->ws, s, nullptr, h->offsets, h->n, h->n));
    VFS_TRY(launch_key_starts(h->keys, h->n, h->offsets, s));
    return 0;
}

int vfs_spatial_hash_buffers(vfs_spatial_hash* h, uint32_t** spatial_keys, uint32_t** spatial_indices,
                             uint32_t** spatial_offsets) {
    VFS_CHECK(h, VFS_ERR_INVALID, "hash is null");
    if (spatial_keys) *spatial_keys = h->keys;
    if (spatial_indices) *spatial_indices = h->indices;
    if (spatial_offsets) *spatial_offsets = h->offsets;
    return 0;
}

int vfs_spatial_hash_destroy(vfs_spatial_hash* h) {
    destroy_hash(h);
    return 0;
}

int vfs_sort_workspace_size(uint32_t n, uint32_t max_value, size_t* bytes) {
    VFS_CHECK(bytes, VFS_ERR_INVALID, "bytes is null");
    VFS_CHECK(n < (1u << 30), VFS_ERR_INVALID, "n must be < 2^30");
    const SortPlan p = make_sort_plan(n, max_value, 8, 16, false);
    *bytes = (p.workspace_words + 2 * (size_t)n) * 4;
    return 0;
}

// GPUCountSort::Run (sort.cpp:107-143): items are reset to 0..n-1 first
// (ClearCounts, sort.slang:20), so on return items holds the permutation and
// keys the sorted keys.
int vfs_sort_pairs(uint32_t* items_dev, uint32_t* keys_dev, uint32_t n, uint32_t max_value, void* workspace,
                   size_t workspace_bytes, void* stream) {
    VFS_CHECK(n < (1u << 30), VFS_ERR_INVALID, "n must be < 2^30");
    if (n == 0) return 0;
    VFS_CHECK(items_dev && keys_dev && workspace, VFS_ERR_INVALID, "null argument");
    const SortPlan p = make_sort_plan(n, max_value, 8, 16, false);
    const size_t need = (p.workspace_words + 2 * (size_t)n) * 4;
    VFS_CHECK(workspace_bytes >= need, VFS_ERR_INVALID,
              "workspace too small: need " + std::to_string(need) + " bytes");
    uint32_t* ws = static_cast<uint32_t*>(workspace);
    uint32_t* out_keys = ws + p.workspace_words;
    uint32_t* out_items = out_keys + n;
    hipStream_t s = as_stream(stream);
    VFS_TRY(sort_keys_iota(p, keys_dev, out_keys, out_items, ws, s, nullptr));
    VFS_TRY(launch_copy_words(out_keys, keys_dev, n, s));
    VFS_TRY(launch_copy_words(out_items, items_dev, n, s));
    return 0;
}

int vfs_scan_workspace_size(uint32_t n, size_t* bytes) {
    VFS_CHECK(bytes, VFS_ERR_INVALID, "bytes is null");
    *bytes = scan_workspace_words(n) * 4;
    return 0;
}

int vfs_scan_exclusive(uint32_t* elements_dev, uint32_t n, void* workspace, size_t workspace_bytes, void* stream) {
    if (n == 0) return 0;
    VFS_CHECK(elements_dev && workspace, VFS_ERR_INVALID, "null argument");
    VFS_CHECK(workspace_bytes >= scan_workspace_words(n) * 4, VFS_ERR_INVALID, "workspace too small");
    return launch_scan_exclusive(elements_dev, n, static_cast<uint32_t*>(workspace), as_stream(stream));
}

int vfs_spatial_offsets(const uint32_t* sorted_keys_dev, uint32_t* offsets_dev, uint32_t n, int init,
                        void* stream) {
    if (n == 0) return 0;
    VFS_CHECK(sorted_keys_dev && offsets_dev, VFS_ERR_INVALID, "null argument");
    if (init) VFS_TRY(launch_fill_value(offsets_dev, n, n, as_stream(stream)));
    return launch_key_starts(sorted_keys_dev, n, offsets_dev, as_stream(stream));
}

int vfs_reorder_f32x3(const uint32_t* sorted_indices_dev, const float* src_dev, float* dst_dev, uint32_t n,
                      void* stream) {
    if (n == 0) return 0;
    VFS_CHECK(sorted_indices_dev && src_dev && dst_dev, VFS_ERR_INVALID, "null argument");
    return launch_reorder_f32x3(sorted_indices_dev, src_dev, dst_dev, n, as_stream(stream));
}

int vfs_volume_map_generate(int shape_kind, const double* shape_params, const uint32_t* res, double support_radius,
                            double tolerance, double margin, float* sdf_out, float* vm_out, float* box_size,
                            float* box_pos, int threads) {
    (void)threads;
    VFS_CHECK(shape_params && res && sdf_out && vm_out && box_size && box_pos, VFS_ERR_INVALID, "null argument");
    VFS_CHECK(shape_kind >= VFS_SHAPE_BOX && shape_kind <= VFS_SHAPE_BOX_WALLS, VFS_ERR_INVALID, "unknown shape");
    VFS_CHECK(res[0] >= 2 && res[1] >= 2 && res[2] >= 2, VFS_ERR_INVALID, "resolution must be >= 2 per axis");
    VFS_CHECK(support_radius > 0.0, VFS_ERR_INVALID, "support radius must be positive");
    const uint64_t total = (uint64_t)res[0] * res[1] * res[2];
    VFS_CHECK(total < (1ull << 31), VFS_ERR_INVALID, "grid too large");
    // Shape AABB (the mesh BVH root box in mesh_sdf.cpp:33-38).
    float lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        const double e = shape_kind == VFS_SHAPE_SPHERE ? shape_params[3] : shape_params[3 + a];
        lo[a] = (float)(shape_params[a] - e);
        hi[a] = (float)(shape_params[a] + e);
    }
    float dmin[3];
    double step[3];
    grid_domain(lo, hi, tolerance, margin, res, box_size, box_pos, dmin, step);
    double gx[16], gw[16];
    gauss_legendre(16, gx, gw);
    float *d_sdf = nullptr, *d_vm = nullptr;
    VFS_HIP(hipMalloc(&d_sdf, total * 4));
    if (hipMalloc(&d_vm, total * 4) != hipSuccess) {
        (void)hipFree(d_sdf);
        return fail(VFS_ERR_NOMEM, "hipMalloc failed");
    }
    int r = launch_volume_map(shape_kind, shape_params, res, support_radius, tolerance, dmin, step, gx, gw, d_sdf,
                              d_vm, nullptr);
    if (r == 0 && hipDeviceSynchronize() != hipSuccess) r = fail(VFS_ERR_HIP, "volume map kernel failed");
    if (r == 0 && (hipMemcpy(sdf_out, d_sdf, total * 4, hipMemcpyDeviceToHost) != hipSuccess ||
                   hipMemcpy(vm_out, d_vm, total * 4, hipMemcpyDeviceToHost) != hipSuccess))
        r = fail(VFS_ERR_HIP, "copy failed");
    (void)hipFree(d_sdf);
    (void)hipFree(d_vm);
    return r;
}

}  // extern "C"
