// Trajectory export: particle state frames of a simulation written to a
// binary file for external analysis or rendering (the "record particle and
// object positions, and export to a file" item the reference lists as not
// done, README.md:33).  Layout: see include/vfs_c.h.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "sim_internal.h"

using namespace vfs;

struct vfs_trajectory {
    vfs_sim* sim = nullptr;
    FILE* f = nullptr;
    uint32_t fields = 0;
    uint64_t frames = 0;
    float4* h_x = nullptr;  // pinned staging
    float4* h_v = nullptr;
    float* h_rho = nullptr;
    uint32_t* h_id = nullptr;
    std::vector<float> aos;
};

namespace {

constexpr char kMagic[8] = {'V', 'F', 'S', 'T', 'R', 'A', 'J', '1'};
constexpr uint32_t kVersion = 1;
constexpr long kFramesOffset = 24;  // header field n_frames

struct Header {
    char magic[8];
    uint32_t version;
    uint32_t n_particles;
    uint32_t fields;
    uint32_t n_objects;
    uint64_t n_frames;
    float box_pos[3];
    float box_size[3];
    uint32_t reserved[2];
};
static_assert(sizeof(Header) == 64, "trajectory header is 64 bytes");

void free_traj(vfs_trajectory* t) {
    if (!t) return;
    if (t->f) std::fclose(t->f);
    void* p[] = {t->h_x, t->h_v, t->h_rho, t->h_id};
    for (void* q : p)
        if (q) (void)hipHostFree(q);
    delete t;
}

int write_all(FILE* f, const void* p, size_t bytes) {
    if (bytes && std::fwrite(p, 1, bytes, f) != bytes) return fail(VFS_ERR_IO, "trajectory: write failed");
    return 0;
}

}  // namespace

extern "C" {

int vfs_trajectory_open(vfs_sim* sim, const char* path, uint32_t fields, vfs_trajectory** out) {
    VFS_CHECK(out, VFS_ERR_INVALID, "out is null");
    *out = nullptr;
    VFS_CHECK(sim && path, VFS_ERR_INVALID, "null argument");
    VFS_CHECK(!sim->batch, VFS_ERR_INVALID, "trajectories are recorded from a single-scene simulation");
    VFS_CHECK(fields != 0 && (fields & ~0xFu) == 0, VFS_ERR_INVALID, "fields must be a non-empty set of VFS_TRAJ_*");
    VFS_CHECK(!(fields & VFS_TRAJ_IDS) || sim->pid, VFS_ERR_STATE, "particle ids need VFS_FLAG_TRACK_IDS");
    auto* t = new vfs_trajectory();
    t->sim = sim;
    t->fields = fields;
    const uint32_t n = sim->n;
    const size_t n1 = n ? n : 1;
    int r = 0;
    if ((fields & VFS_TRAJ_POSITIONS) && hipHostMalloc(&t->h_x, n1 * sizeof(float4)) != hipSuccess) r = 1;
    if ((fields & VFS_TRAJ_VELOCITIES) && hipHostMalloc(&t->h_v, n1 * sizeof(float4)) != hipSuccess) r = 1;
    if ((fields & VFS_TRAJ_DENSITIES) && hipHostMalloc(&t->h_rho, n1 * sizeof(float)) != hipSuccess) r = 1;
    if ((fields & VFS_TRAJ_IDS) && hipHostMalloc(&t->h_id, n1 * sizeof(uint32_t)) != hipSuccess) r = 1;
    if (r) {
        free_traj(t);
        return fail(VFS_ERR_NOMEM, "trajectory: pinned staging allocation failed");
    }
    t->aos.resize(3 * n1);
    t->f = std::fopen(path, "wb");
    if (!t->f) {
        free_traj(t);
        return fail(VFS_ERR_IO, std::string("trajectory: cannot create '") + path + "'");
    }
    // header and the objects' world-to-object matrices
    const SimConsts& c = sim->scenes[0].consts;
    Header h{};
    std::memcpy(h.magic, kMagic, 8);
    h.version = kVersion;
    h.n_particles = n;
    h.fields = fields;
    h.n_objects = c.obj_count;
    for (int a = 0; a < 3; a++) {
        h.box_pos[a] = c.bb_lo[a];
        h.box_size[a] = c.bb_hi[a] - c.bb_lo[a];
    }
    std::vector<DevObject> objs(c.obj_count);
    if (c.obj_count &&
        hipMemcpy(objs.data(), sim->objs_dev + c.obj_begin, c.obj_count * sizeof(DevObject), hipMemcpyDeviceToHost) !=
            hipSuccess) {
        free_traj(t);
        return fail(VFS_ERR_HIP, "trajectory: object copy failed");
    }
    r = write_all(t->f, &h, sizeof h);
    for (const auto& o : objs)
        if (r == 0) r = write_all(t->f, o.T, sizeof o.T);
    if (r) {
        free_traj(t);
        return r;
    }
    *out = t;
    return 0;
}

int vfs_trajectory_record(vfs_trajectory* t, uint64_t substep, double time) {
    VFS_CHECK(t && t->f, VFS_ERR_INVALID, "trajectory is null or closed");
    vfs_sim* sim = t->sim;
    const uint32_t n = sim->n;
    hipStream_t s = sim->stream;
    if (n) {
        if (t->h_x) VFS_HIP(hipMemcpyAsync(t->h_x, sim->x, (size_t)n * 16, hipMemcpyDeviceToHost, s));
        if (t->h_v) VFS_HIP(hipMemcpyAsync(t->h_v, sim->v, (size_t)n * 16, hipMemcpyDeviceToHost, s));
        if (t->h_rho) VFS_HIP(hipMemcpyAsync(t->h_rho, sim->rho, (size_t)n * 4, hipMemcpyDeviceToHost, s));
        if (t->h_id) VFS_HIP(hipMemcpyAsync(t->h_id, sim->pid, (size_t)n * 4, hipMemcpyDeviceToHost, s));
    }
    VFS_HIP(hipStreamSynchronize(s));
    VFS_TRY(write_all(t->f, &substep, 8));
    VFS_TRY(write_all(t->f, &time, 8));
    auto aos = [&](const float4* src) {
        for (uint32_t i = 0; i < n; i++) {
            t->aos[3 * (size_t)i] = src[i].x;
            t->aos[3 * (size_t)i + 1] = src[i].y;
            t->aos[3 * (size_t)i + 2] = src[i].z;
        }
        return write_all(t->f, t->aos.data(), (size_t)n * 12);
    };
    if (t->h_x) VFS_TRY(aos(t->h_x));
    if (t->h_v) VFS_TRY(aos(t->h_v));
    if (t->h_rho) VFS_TRY(write_all(t->f, t->h_rho, (size_t)n * 4));
    if (t->h_id) VFS_TRY(write_all(t->f, t->h_id, (size_t)n * 4));
    t->frames++;
    return 0;
}

int vfs_trajectory_close(vfs_trajectory* t) {
    if (!t) return 0;
    int r = 0;
    if (t->f) {
        if (std::fseek(t->f, kFramesOffset, SEEK_SET) != 0 || std::fwrite(&t->frames, 8, 1, t->f) != 1)
            r = fail(VFS_ERR_IO, "trajectory: cannot finish the header");
        if (std::fclose(t->f) != 0 && r == 0) r = fail(VFS_ERR_IO, "trajectory: close failed");
        t->f = nullptr;
    }
    free_traj(t);
    return r;
}

}  // extern "C"
