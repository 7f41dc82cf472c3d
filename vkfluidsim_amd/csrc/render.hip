// Render-buffer handoff: the reference copies its particle buffers into the
// renderer's own buffers every frame (SPHModel::CreateDataBuffers /
// CopyDataBuffers, src/models/model.cpp:149-170; scene_renderer.cpp:52-55).
// Here the render buffers are device allocations in the reference's layout
// (packed float3 positions, velocities, accelerations; float densities) that
// one kernel fills from the simulation's float4 state, plus an event recorded
// after the copy for the renderer's stream to wait on.  A renderer in another
// process maps the same memory through HIP IPC handles; the producer then
// waits for the copy (vfs_render_buffers_synchronize) before it signals that
// process over its own channel.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

#include "sim_internal.h"

using namespace vfs;

namespace {

// x, v (and a) float4 -> packed float3; rho copied.  Thread i owns particle i.
__global__ __launch_bounds__(256) void k_pack_render(const float4* __restrict__ x, const float4* __restrict__ v,
                                                     const float* __restrict__ rho, const float4* __restrict__ a,
                                                     uint32_t n, float* __restrict__ px, float* __restrict__ pv,
                                                     float* __restrict__ pr, float* __restrict__ pa) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const float4 xi = x[i], vi = v[i];
    px[3 * (size_t)i] = xi.x;
    px[3 * (size_t)i + 1] = xi.y;
    px[3 * (size_t)i + 2] = xi.z;
    pv[3 * (size_t)i] = vi.x;
    pv[3 * (size_t)i + 1] = vi.y;
    pv[3 * (size_t)i + 2] = vi.z;
    pr[i] = rho[i];
    if (pa) {
        const float4 ai = a[i];
        pa[3 * (size_t)i] = ai.x;
        pa[3 * (size_t)i + 1] = ai.y;
        pa[3 * (size_t)i + 2] = ai.z;
    }
}

hipStream_t stream_of(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Largest buffer handed to another process: shared allocations stay below 2 GiB.
constexpr uint64_t kIpcMaxBytes = (1ull << 31) - 1;

static_assert(sizeof(hipIpcMemHandle_t) == 64, "vfs_render_ipc holds 64-byte HIP IPC handles");

}  // namespace

extern "C" {

int vfs_render_buffers_create(uint32_t n_particles, int with_accelerations, vfs_render_buffers* out) {
    VFS_CHECK(out, VFS_ERR_INVALID, "out is null");
    std::memset(out, 0, sizeof *out);
    VFS_CHECK(n_particles < (1u << 28), VFS_ERR_INVALID, "n_particles must be below 2^28");
    vfs_render_buffers b{};
    b.n_particles = n_particles;
    const size_t n = n_particles ? n_particles : 1;
    int r = dev_alloc(&b.positions, 3 * n);
    if (r == 0) r = dev_alloc(&b.velocities, 3 * n);
    if (r == 0) r = dev_alloc(&b.densities, n);
    if (r == 0 && with_accelerations) r = dev_alloc(&b.accelerations, 3 * n);
    hipEvent_t ev = nullptr;
    if (r == 0 && hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess)
        r = fail(VFS_ERR_HIP, "render buffers: event creation failed");
    b.ready = ev;
    if (r) {
        (void)vfs_render_buffers_destroy(&b);
        return r;
    }
    *out = b;
    return 0;
}

int vfs_render_buffers_destroy(vfs_render_buffers* b) {
    if (!b) return 0;
    float* ptrs[] = {b->positions, b->velocities, b->densities, b->accelerations};
    int r = 0;
    for (float* p : ptrs) {
        if (!p) continue;
        const hipError_t e = b->opened ? hipIpcCloseMemHandle(p) : hipFree(p);
        if (e != hipSuccess && r == 0) r = fail(VFS_ERR_HIP, std::string("render buffers: ") + hipGetErrorString(e));
    }
    if (b->ready) (void)hipEventDestroy(reinterpret_cast<hipEvent_t>(b->ready));
    std::memset(b, 0, sizeof *b);
    return r;
}

int vfs_sim_copy_data_buffers(vfs_sim* sim, const vfs_render_buffers* b, void* stream) {
    VFS_CHECK(sim && b, VFS_ERR_INVALID, "null argument");
    VFS_CHECK(!sim->batch, VFS_ERR_INVALID, "render buffers take a single-scene simulation");
    VFS_CHECK(b->n_particles == sim->n, VFS_ERR_INVALID, "render buffers hold another particle count");
    VFS_CHECK(b->positions && b->velocities && b->densities, VFS_ERR_INVALID, "render buffers incomplete");
    VFS_CHECK(b->ready, VFS_ERR_INVALID, "render buffers opened from another process are read-only");
    VFS_CHECK(!b->accelerations || sim->accel, VFS_ERR_STATE, "accelerations need VFS_FLAG_STORE_ACCEL");
    hipStream_t s = stream ? stream_of(stream) : sim->stream;
    if (s != sim->stream) VFS_TRY(sim->order_after(s, sim->stream));
    if (sim->n) {
        hipLaunchKernelGGL(k_pack_render, dim3((sim->n + 255u) / 256u), dim3(256), 0, s, sim->x, sim->v, sim->rho,
                           b->accelerations ? sim->accel : nullptr, sim->n, b->positions, b->velocities, b->densities,
                           b->accelerations);
        VFS_TRY(check_launch("pack_render"));
    }
    VFS_HIP(hipEventRecord(reinterpret_cast<hipEvent_t>(b->ready), s));
    // the simulation's next substeps wait for the copy (they overwrite x, v, rho)
    if (s != sim->stream) VFS_HIP(hipStreamWaitEvent(sim->stream, reinterpret_cast<hipEvent_t>(b->ready), 0));
    return 0;
}

int vfs_render_buffers_export(const vfs_render_buffers* b, vfs_render_ipc* out) {
    VFS_CHECK(b && out, VFS_ERR_INVALID, "null argument");
    VFS_CHECK(!b->opened, VFS_ERR_INVALID, "opened render buffers cannot be exported again");
    VFS_CHECK(b->positions && b->velocities && b->densities, VFS_ERR_INVALID, "render buffers incomplete");
    VFS_CHECK(12ull * b->n_particles <= kIpcMaxBytes, VFS_ERR_INVALID,
              "render buffers shared with another process must be smaller than 2 GiB");
    std::memset(out, 0, sizeof *out);
    float* ptrs[] = {b->positions, b->velocities, b->densities, b->accelerations};
    for (int k = 0; k < 4; k++) {
        if (!ptrs[k]) continue;
        hipIpcMemHandle_t h;
        VFS_HIP(hipIpcGetMemHandle(&h, ptrs[k]));
        std::memcpy(out->buffers[k], &h, 64);
    }
    out->n_particles = b->n_particles;
    out->has_accelerations = b->accelerations ? 1u : 0u;
    return 0;
}

int vfs_render_buffers_open(const vfs_render_ipc* in, vfs_render_buffers* out) {
    VFS_CHECK(in && out, VFS_ERR_INVALID, "null argument");
    std::memset(out, 0, sizeof *out);
    VFS_CHECK(in->n_particles < (1u << 28), VFS_ERR_INVALID, "n_particles must be below 2^28");
    vfs_render_buffers b{};
    b.opened = 1;
    b.n_particles = in->n_particles;
    float** ptrs[] = {&b.positions, &b.velocities, &b.densities, &b.accelerations};
    const int count = in->has_accelerations ? 4 : 3;
    int r = 0;
    for (int k = 0; k < count && r == 0; k++) {
        hipIpcMemHandle_t h;
        std::memcpy(&h, in->buffers[k], 64);
        void* p = nullptr;
        const hipError_t e = hipIpcOpenMemHandle(&p, h, hipIpcMemLazyEnablePeerAccess);
        if (e != hipSuccess)
            r = fail(VFS_ERR_HIP, std::string("render buffers: hipIpcOpenMemHandle: ") + hipGetErrorString(e));
        else
            *ptrs[k] = static_cast<float*>(p);
    }
    if (r) {
        (void)vfs_render_buffers_destroy(&b);
        return r;
    }
    *out = b;
    return 0;
}

int vfs_render_buffers_wait(const vfs_render_buffers* b, void* stream) {
    VFS_CHECK(b, VFS_ERR_INVALID, "render buffers are null");
    VFS_CHECK(b->ready, VFS_ERR_STATE, "opened render buffers have no event: their producer synchronizes");
    VFS_HIP(hipStreamWaitEvent(stream_of(stream), reinterpret_cast<hipEvent_t>(b->ready), 0));
    return 0;
}

int vfs_render_buffers_synchronize(const vfs_render_buffers* b) {
    VFS_CHECK(b, VFS_ERR_INVALID, "render buffers are null");
    if (b->ready) VFS_HIP(hipEventSynchronize(reinterpret_cast<hipEvent_t>(b->ready)));
    return 0;
}

int vfs_render_buffers_read(const vfs_render_buffers* b, float* positions, float* velocities, float* densities,
                            float* accelerations) {
    VFS_CHECK(b && b->positions && b->velocities && b->densities, VFS_ERR_INVALID, "render buffers incomplete");
    VFS_CHECK(!accelerations || b->accelerations, VFS_ERR_INVALID, "render buffers hold no accelerations");
    VFS_TRY(vfs_render_buffers_synchronize(b));
    const size_t n = b->n_particles;
    if (n == 0) return 0;
    if (positions) VFS_HIP(hipMemcpy(positions, b->positions, n * 12, hipMemcpyDeviceToHost));
    if (velocities) VFS_HIP(hipMemcpy(velocities, b->velocities, n * 12, hipMemcpyDeviceToHost));
    if (densities) VFS_HIP(hipMemcpy(densities, b->densities, n * 4, hipMemcpyDeviceToHost));
    if (accelerations) VFS_HIP(hipMemcpy(accelerations, b->accelerations, n * 12, hipMemcpyDeviceToHost));
    return 0;
}

}  // extern "C"
