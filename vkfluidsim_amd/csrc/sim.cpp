// Host orchestration of the WCSPH substep (the C ABI of include/vfs_c.h), for
// one scene (vfs_sim_*) or a batch of independent scenes (vfs_batch_*).
//
// One substep (reference WCSPHWithBoundaryModel::Step, wcsph_with_boundary_model.cpp:81-117,
// with SPHModel::RunSpatialHash, model.cpp:210-223) runs as
//   sort(keys, iota) -> cell ranges -> reorder x, v (and ids)
//   -> boundary volume + density + EOS terms
//   -> pressure + viscosity + wall + integrate + hash of the new positions.
// ExternalAccel is folded into the force pass (a starts at gravity), and the
// keys of the next substep are produced by the force pass, so the standalone
// hash kernel only runs after set_state.  The kernel sequence of a substep is
// captured once into a hipGraph and replayed.
//
// A batch concatenates its scenes in one set of particle arrays, each scene's
// segment padded to a multiple of 64 entries so that no wave spans two
// scenes; a scene's hash keys are offset by its segment start, so one sort of
// all keys sorts every scene inside its own segment (stable, so exactly as the
// scene alone), and one launch per stage serves all scenes, each wave reading
// its scene's constants from a table.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "sim_internal.h"

namespace vfs {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int check_launch(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(VFS_ERR_HIP, std::string("kernel launch ") + what + ": " + hipGetErrorString(e));
    return 0;
}

SimConsts make_consts(const vfs_params& p) {
    SimConsts c{};
    c.gx = p.gravity[0];
    c.gy = p.gravity[1];
    c.gz = p.gravity[2];
    c.h = p.smooth_radius;
    c.h2 = c.h * c.h;
    c.inv_h = 1.0f / c.h;
    c.cell = p.cell_size > 0.0f ? p.cell_size : p.smooth_radius;
    c.rho0 = p.target_density;
    c.B = p.stiffness;
    c.gamma = p.expoent;
    c.gamma_int = (p.expoent >= 0.0f && p.expoent <= 64.0f && std::floor(p.expoent) == p.expoent)
                      ? (int32_t)p.expoent
                      : -1;
    c.nu10 = (2.0f * (3.0f + 2.0f)) * p.viscosity_strenght;
    c.nu = p.viscosity_strenght;
    c.kwall_neg = -p.wall_stiffness;
    c.dt = p.time_scale * p.fixed_dt / (float)p.iterations;
    const float h = c.h;
    const float h3 = (h * h) * h;
    c.mass = c.rho0 * ((0.8f * h3) / 8.0f);
    const float pi_f = 3.14159265358979323846f;
    c.sigma = 8.0f / (pi_f * (float)std::pow((double)h, 3));
    const float grad_scale = 48.0f / (pi_f * (float)std::pow((double)h, 4));
    c.grad_sigma = grad_scale / h;
    c.visc_eps_h2 = 0.01f * c.h2;
    c.rho0_sq = c.rho0 * c.rho0;
    for (int a = 0; a < 3; a++) {
        c.bb_lo[a] = p.bbox_pos[a];
        c.bb_hi[a] = p.bbox_pos[a] + p.bbox_size[a];
    }
    c.n = (uint32_t)p.n_particles;
    c.nbr_cap = kNeighbourCap;
    c.fm = make_fastmod(c.n);
    c.base = 0;
    c.obj_begin = 0;
    c.obj_count = 0;
    return c;
}

LagueConsts make_lague_consts(const vfs_params& p, const vfs_lague_params& lp) {
    LagueConsts l{};
    l.gx = p.gravity[0];
    l.gy = p.gravity[1];
    l.gz = p.gravity[2];
    l.h = p.smooth_radius;
    l.h2 = l.h * l.h;
    l.dt = p.time_scale * p.fixed_dt / (float)p.iterations;
    l.pf = 1.0f / 120.0f;
    l.rho0 = p.target_density;
    l.pm = lp.pressure_multiplier;
    l.npm = lp.near_pressure_multiplier;
    l.visc = lp.viscosity_strenght;
    l.damp = lp.wall_damping_factor;
    // CalcKernelCoefficients (model.cpp:224-233): std::pow(float, int) is computed in double
    const float pi_f = 3.14159265358979323846f;
    const float r = p.smooth_radius;
    l.sp3 = 15.0f / (pi_f * (float)std::pow((double)r, 6));
    l.sp2 = 15.0f / (2.0f * pi_f * (float)std::pow((double)r, 5));
    l.dsp3 = 45.0f / (pi_f * (float)std::pow((double)r, 6));
    l.dsp2 = 15.0f / (pi_f * (float)std::pow((double)r, 5));
    // Poly6 scale (kernels_3d.slang:19) with the shader's PI = 3.1415926 and
    // the integral power r^9 by binary powering in double, rounded once
    double base = (double)std::fabs(r), r9 = 1.0;
    for (unsigned e = 9; e; e >>= 1) {
        if (e & 1u) r9 *= base;
        base *= base;
    }
    l.poly6 = 315.0f / ((64.0f * 3.1415926f) * (float)r9);
    for (int a = 0; a < 3; a++) {
        l.bb_lo[a] = p.bbox_pos[a];
        l.bb_hi[a] = p.bbox_pos[a] + p.bbox_size[a];
    }
    return l;
}

}  // namespace vfs

using namespace vfs;

struct vfs_batch : vfs_sim {};

namespace {

void free_sim(vfs_sim* s) {
    if (!s) return;
    s->drop_graph();
    void* ptrs[] = {s->x,           s->v,       s->xs,         s->vs,       s->bnd,      s->accel,
                    s->rho,         s->pid,     s->pids,       s->keys,     s->sorted_keys, s->perm,
                    s->offsets,     s->cell_start, s->sort_ws, s->objs_dev, s->nbr,      s->nbr_count,
                    s->consts_dev,  s->wave_scene, s->pred, s->ps, s->near_rho, s->check_counters};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (float* g : s->grids)
        if (g) (void)hipFree(g);
    if (s->order_event) (void)hipEventDestroy(s->order_event);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

int validate_params(const vfs_params* p) {
    VFS_CHECK(p, VFS_ERR_INVALID, "params is null");
    VFS_CHECK(p->n_particles >= 0 && p->n_particles < (1 << 28), VFS_ERR_INVALID,
              "n_particles must be in [0, 2^28)");
    VFS_CHECK(p->iterations >= 1, VFS_ERR_INVALID, "iterations must be >= 1");
    // the force pass's reciprocals are exact for divisors in [2^-100, 2^100],
    // which r in (0, h] and r^2 + 0.01 h^2 stay in for h in this range
    VFS_CHECK(p->smooth_radius >= 0x1p-40f && p->smooth_radius <= 0x1p40f, VFS_ERR_INVALID,
              "smooth_radius must lie in [2^-40, 2^40]");
    VFS_CHECK(p->cell_size >= 0.0f && std::isfinite(p->cell_size), VFS_ERR_INVALID, "cell_size must be >= 0");
    const float cell = p->cell_size > 0.0f ? p->cell_size : p->smooth_radius;
    VFS_CHECK(cell >= p->smooth_radius, VFS_ERR_INVALID,
              "cell_size smaller than smooth_radius would miss neighbours of the 27-cell stencil");
    return 0;
}

int validate_objects(const vfs_boundary_object* objects, uint32_t n_objects) {
    VFS_CHECK(n_objects == 0 || objects, VFS_ERR_INVALID, "objects is null");
    for (uint32_t k = 0; k < n_objects; k++) {
        const auto& o = objects[k];
        VFS_CHECK(o.sdf_grid && o.volume_map_grid, VFS_ERR_INVALID, "boundary object grid is null");
        VFS_CHECK(o.resolution[0] >= 2 && o.resolution[1] >= 2 && o.resolution[2] >= 2, VFS_ERR_INVALID,
                  "boundary grid resolution must be >= 2 per axis");
        VFS_CHECK((uint64_t)o.resolution[0] * o.resolution[1] * o.resolution[2] < (1ull << 31), VFS_ERR_INVALID,
                  "boundary grid too large");
        VFS_CHECK(o.box_size[0] > 0 && o.box_size[1] > 0 && o.box_size[2] > 0, VFS_ERR_INVALID,
                  "boundary box size must be positive");
    }
    return 0;
}

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Builds a simulation of n_scenes scenes (objects listed scene after scene,
// objects_per_scene[s] of them for scene s).
template <class T>
int create_impl(const vfs_params* params, uint32_t n_scenes, const vfs_boundary_object* objects,
                const uint32_t* objects_per_scene, bool batch, T** out) {
    VFS_CHECK(out, VFS_ERR_INVALID, "out is null");
    *out = nullptr;
    VFS_CHECK(params, VFS_ERR_INVALID, "params is null");
    VFS_CHECK(n_scenes >= 1, VFS_ERR_INVALID, "at least one scene is needed");
    uint64_t total = 0, total_objects = 0;
    for (uint32_t k = 0; k < n_scenes; k++) {
        VFS_TRY(validate_params(&params[k]));
        VFS_CHECK(params[k].flags == params[0].flags, VFS_ERR_INVALID, "all scenes of a batch need the same flags");
        VFS_CHECK(!batch || !(params[k].flags & VFS_FLAG_CHECK_STATE), VFS_ERR_INVALID,
                  "VFS_FLAG_CHECK_STATE is for single-scene simulations");
        const uint64_t np = (uint64_t)params[k].n_particles;
        total += batch ? (np + 63u) / 64u * 64u : np;
        total_objects += objects_per_scene[k];
    }
    VFS_CHECK(total < (1ull << 28), VFS_ERR_INVALID, "total particle entries must be < 2^28");
    VFS_TRY(validate_objects(objects, (uint32_t)total_objects));

    auto* s = new T();
    s->batch = batch;
    s->flags = params[0].flags;
    s->n = (uint32_t)total;
    auto bail = [&](int r) {
        free_sim(s);
        return r;
    };
    uint32_t base = 0, obj = 0;
    for (uint32_t k = 0; k < n_scenes; k++) {
        SceneSeg sc;
        sc.params = params[k];
        sc.consts = make_consts(params[k]);
        sc.n = (uint32_t)params[k].n_particles;
        sc.base = base;
        sc.padded = batch ? (sc.n + 63u) / 64u * 64u : sc.n;
        sc.cell0 = sc.consts.cell;
        sc.consts.base = base;
        sc.consts.obj_begin = obj;
        sc.consts.obj_count = objects_per_scene[k];
        base += sc.padded;
        obj += objects_per_scene[k];
        s->scenes.push_back(sc);
    }
    if (hipGetDevice(&s->device) != hipSuccess) return bail(fail(VFS_ERR_HIP, "no HIP device"));
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(VFS_ERR_HIP, "hipStreamCreate failed"));
    const uint32_t n = s->n;
    int r = 0;
    r |= dev_alloc(&s->x, n);
    r |= dev_alloc(&s->v, n);
    // the density pass reads candidates in groups of kCandidateGroup and may
    // read up to kCandidateGroup - 1 elements past the last particle
    r |= dev_alloc(&s->xs, (size_t)n + kCandidateGroup);
    r |= dev_alloc(&s->vs, n);
    r |= dev_alloc(&s->bnd, n);
    r |= dev_alloc(&s->rho, n);
    r |= dev_alloc(&s->keys, n);
    r |= dev_alloc(&s->sorted_keys, n);
    r |= dev_alloc(&s->perm, n);
    r |= dev_alloc(&s->offsets, (size_t)n + 1);
    r |= dev_alloc(&s->nbr, (size_t)n * kNeighbourWords);
    r |= dev_alloc(&s->nbr_count, n);
    r |= dev_alloc(&s->cell_start, (size_t)n + 1);
    r |= dev_alloc(&s->check_counters, 2);
    if (s->flags & VFS_FLAG_TRACK_IDS) {
        r |= dev_alloc(&s->pid, n);
        r |= dev_alloc(&s->pids, n);
    }
    if (s->flags & VFS_FLAG_STORE_ACCEL) r |= dev_alloc(&s->accel, n);
    s->plan = make_sort_plan(n, n ? n - 1 : 0, s->tuning.sort_bits, s->tuning.sort_items);
    r |= dev_alloc(&s->sort_ws, s->plan.workspace_words);
    if (batch) {
        r |= dev_alloc(&s->consts_dev, (size_t)n_scenes + 1);
        r |= dev_alloc(&s->wave_scene, launch_waves(n));
    }
    if (r) return bail(VFS_ERR_NOMEM);
    for (uint32_t k = 0; k < (uint32_t)total_objects; k++) {
        const auto& o = objects[k];
        DevObject d{};
        std::memcpy(d.T, o.transform, sizeof(d.T));
        std::memcpy(d.Rm, o.rotation, sizeof(d.Rm));
        for (int a = 0; a < 3; a++) {
            d.box_pos[a] = o.box_pos[a];
            d.box_size[a] = o.box_size[a];
            d.res[a] = o.resolution[a];
        }
        const size_t cnt = (size_t)o.resolution[0] * o.resolution[1] * o.resolution[2];
        float *gs = nullptr, *gv = nullptr;
        if (dev_alloc(&gs, cnt) || dev_alloc(&gv, cnt)) {
            if (gs) (void)hipFree(gs);
            return bail(VFS_ERR_NOMEM);
        }
        s->grids.push_back(gs);
        s->grids.push_back(gv);
        hipError_t e = hipMemcpyAsync(gs, o.sdf_grid, cnt * 4, hipMemcpyHostToDevice, s->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(gv, o.volume_map_grid, cnt * 4, hipMemcpyHostToDevice, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) return bail(fail(VFS_ERR_HIP, std::string("grid upload failed: ") + hipGetErrorString(e)));
        d.sdf = gs;
        d.vm = gv;
        s->objs_host.push_back(d);
    }
    if (total_objects) {
        if (dev_alloc(&s->objs_dev, total_objects)) return bail(VFS_ERR_NOMEM);
        if (hipMemcpyAsync(s->objs_dev, s->objs_host.data(), total_objects * sizeof(DevObject),
                           hipMemcpyHostToDevice, s->stream) != hipSuccess ||
            hipStreamSynchronize(s->stream) != hipSuccess)
            return bail(fail(VFS_ERR_HIP, "object upload failed"));
    }
    if (batch) {
        std::vector<uint32_t> ws(launch_waves(n), n_scenes);
        for (uint32_t k = 0; k < n_scenes; k++)
            for (uint32_t w = s->scenes[k].base / 64u; w < (s->scenes[k].base + s->scenes[k].padded) / 64u; w++)
                ws[w] = k;
        if (hipMemcpyAsync(s->wave_scene, ws.data(), ws.size() * 4, hipMemcpyHostToDevice, s->stream) !=
            hipSuccess)
            return bail(fail(VFS_ERR_HIP, "scene table upload failed"));
        if (s->upload_consts()) return bail(VFS_ERR_HIP);
    }
    // zero state, ids 0..n-1 per scene; padding entries keep key = own index
    // (unique, above every key of their scene, below the next scene's)
    if (n) {
        if (hipMemsetAsync(s->x, 0, (size_t)n * 16, s->stream) != hipSuccess ||
            hipMemsetAsync(s->v, 0, (size_t)n * 16, s->stream) != hipSuccess ||
            hipMemsetAsync(s->rho, 0, (size_t)n * 4, s->stream) != hipSuccess ||
            hipMemsetAsync(s->bnd, 0, (size_t)n * 16, s->stream) != hipSuccess ||
            hipMemsetAsync(s->nbr_count, 0, (size_t)n * 4, s->stream) != hipSuccess)
            return bail(fail(VFS_ERR_HIP, "state clear failed"));
        if (s->accel && hipMemsetAsync(s->accel, 0, (size_t)n * 16, s->stream) != hipSuccess)
            return bail(fail(VFS_ERR_HIP, "state clear failed"));
        if (launch_iota(s->keys, n, s->stream)) return bail(VFS_ERR_HIP);
        for (const auto& sc : s->scenes)
            if (s->pid && launch_iota(s->pid + sc.base, sc.padded, s->stream)) return bail(VFS_ERR_HIP);
        if (hipStreamSynchronize(s->stream) != hipSuccess) return bail(fail(VFS_ERR_HIP, "sync failed"));
    }
    *out = s;
    return 0;
}

int scene_check(vfs_sim* sim, uint32_t scene) {
    VFS_CHECK(sim, VFS_ERR_INVALID, "simulation is null");
    VFS_CHECK(scene < sim->scenes.size(), VFS_ERR_INVALID, "scene index out of range");
    return 0;
}

int set_state_impl(vfs_sim* sim, uint32_t scene, const float* positions, const float* velocities, uint32_t offset,
                   int64_t count) {
    VFS_TRY(scene_check(sim, scene));
    const SceneSeg& sc = sim->scenes[scene];
    if (count < 0) count = (int64_t)sc.n - (int64_t)offset;
    VFS_CHECK(count >= 0 && (uint64_t)offset + (uint64_t)count <= sc.n, VFS_ERR_INVALID, "range out of bounds");
    if (count == 0) return 0;
    VFS_CHECK(positions && velocities, VFS_ERR_INVALID, "positions/velocities null");
    std::vector<float4> hx((size_t)count), hv((size_t)count);
    for (int64_t i = 0; i < count; i++) {
        hx[i] = make_float4(positions[3 * i], positions[3 * i + 1], positions[3 * i + 2], 0.0f);
        hv[i] = make_float4(velocities[3 * i], velocities[3 * i + 1], velocities[3 * i + 2], 0.0f);
    }
    const size_t at = (size_t)sc.base + offset;
    VFS_HIP(hipStreamSynchronize(sim->stream));
    VFS_HIP(hipMemcpyAsync(sim->x + at, hx.data(), (size_t)count * 16, hipMemcpyHostToDevice, sim->stream));
    VFS_HIP(hipMemcpyAsync(sim->v + at, hv.data(), (size_t)count * 16, hipMemcpyHostToDevice, sim->stream));
    VFS_HIP(hipMemsetAsync(sim->rho + at, 0, (size_t)count * 4, sim->stream));
    if (sim->pid) VFS_TRY(launch_iota(sim->pid + sc.base, sc.n, sim->stream));
    VFS_HIP(hipStreamSynchronize(sim->stream));
    sim->keys_valid = false;
    return 0;
}

int set_params_impl(vfs_sim* sim, uint32_t scene, const vfs_params* params) {
    VFS_TRY(scene_check(sim, scene));
    VFS_TRY(validate_params(params));
    SceneSeg& sc = sim->scenes[scene];
    VFS_CHECK(params->n_particles == sc.params.n_particles, VFS_ERR_INVALID, "n_particles cannot change");
    VFS_CHECK(params->flags == sc.params.flags, VFS_ERR_INVALID, "flags cannot change");
    VFS_HIP(hipStreamSynchronize(sim->stream));
    const float old_cell = sc.consts.cell;
    SimConsts c = make_consts(*params);
    c.base = sc.consts.base;
    c.obj_begin = sc.consts.obj_begin;
    c.obj_count = sc.consts.obj_count;
    sc.params = *params;
    sc.cell0 = c.cell;
    c.cell = sc.cell0 * sim->tuning.cell_scale;
    sc.consts = c;
    if (sim->model == 1) sim->lague = make_lague_consts(*params, sim->lague_params);
    if (sc.consts.cell != old_cell) sim->keys_valid = false;
    sim->drop_graph();
    return sim->upload_consts();
}

int get_state_impl(vfs_sim* sim, uint32_t scene, float* positions, float* velocities, float* densities,
                   uint32_t* particle_ids) {
    VFS_TRY(scene_check(sim, scene));
    VFS_CHECK(!particle_ids || sim->pid, VFS_ERR_STATE, "particle ids need VFS_FLAG_TRACK_IDS");
    const SceneSeg& sc = sim->scenes[scene];
    const uint32_t n = sc.n;
    VFS_HIP(hipStreamSynchronize(sim->stream));
    std::vector<float4> tmp(n);
    auto aos = [&](const float4* src, float* dst) -> int {
        VFS_HIP(hipMemcpy(tmp.data(), src + sc.base, (size_t)n * 16, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < n; i++) {
            dst[3 * (size_t)i] = tmp[i].x;
            dst[3 * (size_t)i + 1] = tmp[i].y;
            dst[3 * (size_t)i + 2] = tmp[i].z;
        }
        return 0;
    };
    if (positions && n) VFS_TRY(aos(sim->x, positions));
    if (velocities && n) VFS_TRY(aos(sim->v, velocities));
    if (densities && n) VFS_HIP(hipMemcpy(densities, sim->rho + sc.base, (size_t)n * 4, hipMemcpyDeviceToHost));
    if (particle_ids && n)
        VFS_HIP(hipMemcpy(particle_ids, sim->pid + sc.base, (size_t)n * 4, hipMemcpyDeviceToHost));
    return 0;
}

// Scene-local hash tables: keys and permutation relative to the segment, the
// offsets table with the scene's own "empty" value n.
int get_hash_impl(vfs_sim* sim, uint32_t scene, uint32_t* sorted_keys, uint32_t* sorted_indices,
                  uint32_t* offsets) {
    VFS_TRY(scene_check(sim, scene));
    VFS_CHECK(sim->stepped, VFS_ERR_STATE, "no substep has run yet");
    const SceneSeg& sc = sim->scenes[scene];
    const uint32_t n = sc.n;
    VFS_HIP(hipStreamSynchronize(sim->stream));
    if (!n) return 0;
    if (sorted_keys) {
        VFS_HIP(hipMemcpy(sorted_keys, sim->sorted_keys + sc.base, (size_t)n * 4, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < n; i++) sorted_keys[i] -= sc.base;
    }
    if (sorted_indices) {
        VFS_HIP(hipMemcpy(sorted_indices, sim->perm + sc.base, (size_t)n * 4, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < n; i++) sorted_indices[i] -= sc.base;
    }
    if (offsets) {
        // the reference's offsets table (first index of each key, n if empty)
        // from the bucket starts
        std::vector<uint32_t> cs((size_t)n + 1);
        VFS_HIP(hipMemcpy(cs.data(), sim->cell_start + sc.base, ((size_t)n + 1) * 4, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < n; i++) offsets[i] = cs[i] < cs[i + 1] ? cs[i] - sc.base : n;
    }
    return 0;
}

int get_boundary_impl(vfs_sim* sim, uint32_t scene, float* boundary_density, float* boundary_gradient) {
    VFS_TRY(scene_check(sim, scene));
    const SceneSeg& sc = sim->scenes[scene];
    const uint32_t n = sc.n;
    VFS_HIP(hipStreamSynchronize(sim->stream));
    std::vector<float4> tmp(n);
    if (n) VFS_HIP(hipMemcpy(tmp.data(), sim->bnd + sc.base, (size_t)n * 16, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < n; i++) {
        if (boundary_density) boundary_density[i] = tmp[i].w;
        if (boundary_gradient) {
            boundary_gradient[3 * (size_t)i] = tmp[i].x;
            boundary_gradient[3 * (size_t)i + 1] = tmp[i].y;
            boundary_gradient[3 * (size_t)i + 2] = tmp[i].z;
        }
    }
    return 0;
}

int set_tuning_impl(vfs_sim* sim, const vfs_tuning* t) {
    VFS_CHECK(sim && t, VFS_ERR_INVALID, "null argument");
    VFS_CHECK(valid_sort_config(t->sort_bits, t->sort_items), VFS_ERR_INVALID,
              "sort_bits must be 8, 9 or 10 and sort_items 8, 12 or 16");
    VFS_CHECK(t->block_size == 128 || t->block_size == 256, VFS_ERR_INVALID, "block_size must be 128 or 256");
    VFS_CHECK(t->stage_rows == 0 || t->stage_rows == 1, VFS_ERR_INVALID, "stage_rows must be 0 or 1");
    VFS_CHECK(t->cell_scale >= 1.0f && t->cell_scale <= 4.0f, VFS_ERR_INVALID, "cell_scale must be in [1, 4]");
    VFS_HIP(hipStreamSynchronize(sim->stream));
    if (t->sort_bits != sim->tuning.sort_bits || t->sort_items != sim->tuning.sort_items) {
        const SortPlan plan = make_sort_plan(sim->n, sim->n ? sim->n - 1 : 0, t->sort_bits, t->sort_items);
        uint32_t* ws = nullptr;
        VFS_TRY(dev_alloc(&ws, plan.workspace_words));
        (void)hipFree(sim->sort_ws);
        sim->sort_ws = ws;
        sim->plan = plan;
    }
    const bool cell_changed = t->cell_scale != sim->tuning.cell_scale;
    sim->tuning = *t;
    sim->drop_graph();
    if (cell_changed) VFS_TRY(sim->set_cell_scale(t->cell_scale));
    return 0;
}

// Times every launch configuration on `substeps_per_trial` substeps from the
// current state (restored before each trial and at the end) and keeps the
// fastest.  Cell scales other than the current one are tried only when
// allowed (they change the hash, so results are no longer bit-identical).
int autotune_impl(vfs_sim* sim, uint32_t substeps_per_trial, bool allow_cell_size, vfs_tuning* best,
                  float* best_ms) {
    VFS_CHECK(sim, VFS_ERR_INVALID, "simulation is null");
    VFS_CHECK(substeps_per_trial >= 1, VFS_ERR_INVALID, "substeps_per_trial must be >= 1");
    const vfs_tuning start = sim->tuning;
    vfs_tuning winner = start;
    float winner_ms = 0.0f;
    bool have = false;
    vfs_sim::Snapshot sn;
    hipEvent_t a = nullptr, b = nullptr;
    int rc = sim->snapshot(sn, false);
    if (rc == 0 && (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess))
        rc = fail(VFS_ERR_HIP, "hipEventCreate failed");
    const bool saved_stepped = sim->stepped;
    static const int bits[] = {8, 10};
    static const int items[] = {12, 16};
    static const int blocks[] = {128, 256};
    static const int stage[] = {0, 1};
    std::vector<float> scales{start.cell_scale};
    if (allow_cell_size)
        for (float f : {1.0f, 1.25f, 1.5f})
            if (f != start.cell_scale) scales.push_back(f);
    for (float cs : scales)
        for (int sb : bits)
            for (int si : items)
                for (int bs : blocks)
                    for (int st : stage) {
                        if (rc != 0) break;
                        const vfs_tuning t{sb, si, bs, st, cs};
                        rc = set_tuning_impl(sim, &t);
                        // every trial starts from the same state; one untimed
                        // substep builds the graph and warms the caches
                        if (rc == 0) rc = sim->snapshot(sn, true);
                        if (rc == 0) rc = sim->substeps(1, sim->stream);
                        if (rc != 0) break;
                        (void)hipEventRecord(a, sim->stream);
                        rc = sim->substeps(substeps_per_trial, sim->stream);
                        (void)hipEventRecord(b, sim->stream);
                        if (rc == 0 && hipEventSynchronize(b) != hipSuccess) rc = fail(VFS_ERR_HIP, "timing failed");
                        if (rc != 0) break;
                        float ms = 0.0f;
                        (void)hipEventElapsedTime(&ms, a, b);
                        ms /= (float)substeps_per_trial;
                        if (!have || ms < winner_ms) {
                            winner = t;
                            winner_ms = ms;
                            have = true;
                        }
                    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    // the scene state is left as it was before the trials
    if (rc == 0) rc = set_tuning_impl(sim, &start);
    if (sn.x) {
        const int r2 = sim->snapshot(sn, true);
        if (rc == 0) rc = r2;
    }
    vfs_sim::free_snapshot(sn);
    sim->stepped = saved_stepped;
    if (rc != 0) {
        (void)set_tuning_impl(sim, &start);
        return rc;
    }
    VFS_TRY(set_tuning_impl(sim, &winner));
    if (best) *best = winner;
    if (best_ms) *best_ms = winner_ms;
    return 0;
}

int profile_impl(vfs_sim* sim, uint32_t n, float* ms, uint32_t* launches, const char** names, uint32_t max_stages,
                 uint32_t* n_stages) {
    VFS_CHECK(sim && n_stages, VFS_ERR_INVALID, "null argument");
    static const char* kNames[] = {"predict",   "sort_prepare", "sort_scan", "sort_histogram", "sort_small",
                                   "sort_pass", "sort_merge",   "reorder",   "cell_ranges",    "density",
                                   "forces_integrate"};
    constexpr uint32_t kN = sizeof(kNames) / sizeof(kNames[0]);
    *n_stages = 0;
    if (sim->n == 0 || n == 0) return 0;
    VFS_TRY(sim->ensure_keys(sim->stream));
    EventTimer timer;
    for (uint32_t k = 0; k < n; k++) {
        sim->last_full_sort = !sim->resort_ready;
        VFS_TRY(sim->record_substep(sim->stream, &timer));
        sim->resort_ready = true;
    }
    sim->stepped = true;
    VFS_HIP(hipStreamSynchronize(sim->stream));
    if (timer.err != hipSuccess) return fail(VFS_ERR_HIP, "event creation failed");
    double tot[kN] = {0};
    uint32_t cnt[kN] = {0};
    for (auto& r : timer.recs) {
        float t = 0.0f;
        VFS_HIP(hipEventElapsedTime(&t, r.a, r.b));
        for (uint32_t i = 0; i < kN; i++)
            if (r.name == kNames[i]) {
                tot[i] += t;
                cnt[i]++;
            }
    }
    uint32_t m = 0;
    for (uint32_t i = 0; i < kN && m < max_stages; i++) {
        if (cnt[i] == 0) continue;  // a stage the model does not have
        if (ms) ms[m] = (float)(tot[i] / n);
        if (launches) launches[m] = cnt[i] / n;
        if (names) names[m] = kNames[i];
        m++;
    }
    *n_stages = m;
    return 0;
}

int neighbour_stats_impl(vfs_sim* sim, double* mean, uint32_t* max_count, uint32_t* overflowed) {
    VFS_CHECK(sim, VFS_ERR_INVALID, "simulation is null");
    if (mean) *mean = 0.0;
    if (max_count) *max_count = 0;
    if (overflowed) *overflowed = 0;
    if (sim->n == 0 || !sim->stepped) return 0;
    std::vector<uint32_t> counts(sim->n);
    VFS_HIP(hipMemcpyAsync(counts.data(), sim->nbr_count, (size_t)sim->n * 4, hipMemcpyDeviceToHost, sim->stream));
    VFS_HIP(hipStreamSynchronize(sim->stream));
    uint64_t sum = 0, real = 0;
    uint32_t mx = 0, over = 0;
    for (const auto& sc : sim->scenes)
        for (uint32_t i = sc.base; i < sc.base + sc.n; i++) {
            const uint32_t c = counts[i];
            sum += c;
            mx = std::max(mx, c);
            over += c > kNeighbourCap ? 1u : 0u;
            real++;
        }
    if (mean) *mean = real ? (double)sum / (double)real : 0.0;
    if (max_count) *max_count = mx;
    if (overflowed) *overflowed = over;
    return 0;
}

}  // namespace

extern "C" {

const char* vfs_last_error(void) { return g_last_error.c_str(); }
int vfs_abi_version(void) { return VFS_ABI_VERSION; }

int vfs_device_count(int* count) {
    if (!count) return fail(VFS_ERR_INVALID, "count is null");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    (void)hipGetLastError();
    *count = c;
    return 0;
}

int vfs_dev_malloc(void** ptr, size_t bytes) {
    VFS_CHECK(ptr, VFS_ERR_INVALID, "ptr is null");
    VFS_HIP(hipMalloc(ptr, bytes ? bytes : 1));
    return 0;
}
int vfs_dev_free(void* ptr) {
    if (ptr) VFS_HIP(hipFree(ptr));
    return 0;
}
int vfs_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    if (bytes) VFS_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return 0;
}
int vfs_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    if (bytes) VFS_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return 0;
}
int vfs_memcpy_d2d(void* dst, const void* src, size_t bytes) {
    if (bytes) VFS_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return 0;
}
int vfs_memset_d32(void* dst, uint32_t value, size_t count) {
    if (count) VFS_HIP(hipMemsetD32(reinterpret_cast<hipDeviceptr_t>(dst), (int)value, count));
    return 0;
}
int vfs_stream_create(void** stream) {
    VFS_CHECK(stream, VFS_ERR_INVALID, "stream is null");
    hipStream_t s;
    VFS_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return 0;
}
int vfs_stream_destroy(void* stream) {
    if (stream) VFS_HIP(hipStreamDestroy(as_stream(stream)));
    return 0;
}
int vfs_stream_synchronize(void* stream) {
    VFS_HIP(hipStreamSynchronize(as_stream(stream)));
    return 0;
}

int vfs_sim_create(const vfs_params* params, const vfs_boundary_object* objects, uint32_t n_objects,
                   vfs_sim** out) {
    VFS_CHECK(n_objects == 0 || objects, VFS_ERR_INVALID, "objects is null");
    return create_impl(params, 1, objects, &n_objects, false, out);
}

int vfs_lague_create(const vfs_params* params, const vfs_lague_params* lague, vfs_sim** out) {
    VFS_CHECK(out, VFS_ERR_INVALID, "out is null");
    *out = nullptr;
    // LagueModel's default parameters (lague_model.cpp:26-34)
    const vfs_lague_params lp = lague ? *lague : vfs_lague_params{0.95f, 288.0f, 2.16f, 0.03f};
    uint32_t none = 0;
    vfs_sim* s = nullptr;
    VFS_TRY(create_impl(params, 1, nullptr, &none, false, &s));
    s->model = 1;
    s->lague_params = lp;
    s->lague = make_lague_consts(s->scenes[0].params, lp);
    const uint32_t n = s->n;
    int r = 0;
    r |= dev_alloc(&s->pred, n);
    // the cell walk reads candidates in groups of kCandidateGroup past the last one
    r |= dev_alloc(&s->ps, (size_t)n + kCandidateGroup);
    r |= dev_alloc(&s->near_rho, n);
    if (r) {
        free_sim(s);
        return VFS_ERR_NOMEM;
    }
    if (n && (hipMemsetAsync(s->near_rho, 0, (size_t)n * 4, s->stream) != hipSuccess ||
              hipMemsetAsync(s->pred, 0, (size_t)n * 16, s->stream) != hipSuccess ||
              hipMemsetAsync(s->ps, 0, ((size_t)n + kCandidateGroup) * 16, s->stream) != hipSuccess ||
              hipStreamSynchronize(s->stream) != hipSuccess)) {
        free_sim(s);
        return fail(VFS_ERR_HIP, "state clear failed");
    }
    *out = s;
    return 0;
}

int vfs_lague_set_params(vfs_sim* sim, const vfs_lague_params* lague) {
    VFS_CHECK(sim && lague, VFS_ERR_INVALID, "null argument");
    VFS_CHECK(sim->model == 1, VFS_ERR_INVALID, "not a Lague model simulation");
    VFS_HIP(hipStreamSynchronize(sim->stream));
    sim->lague_params = *lague;
    sim->lague = make_lague_consts(sim->scenes[0].params, *lague);
    sim->drop_graph();
    return 0;
}

int vfs_lague_get_params(const vfs_sim* sim, vfs_lague_params* lague) {
    VFS_CHECK(sim && lague, VFS_ERR_INVALID, "null argument");
    VFS_CHECK(sim->model == 1, VFS_ERR_INVALID, "not a Lague model simulation");
    *lague = sim->lague_params;
    return 0;
}

int vfs_sim_destroy(vfs_sim* sim) {
    if (sim) {
        (void)hipStreamSynchronize(sim->stream);
        free_sim(sim);
    }
    return 0;
}

int vfs_sim_set_state(vfs_sim* sim, const float* positions, const float* velocities, uint32_t offset,
                      int64_t count) {
    return set_state_impl(sim, 0, positions, velocities, offset, count);
}

int vfs_sim_set_params(vfs_sim* sim, const vfs_params* params) { return set_params_impl(sim, 0, params); }

int vfs_sim_get_params(const vfs_sim* sim, vfs_params* params) {
    VFS_CHECK(sim && params, VFS_ERR_INVALID, "null argument");
    *params = sim->scenes[0].params;
    return 0;
}

// Counts of the state after the work queued on the simulation stream.
int check_state_impl(vfs_sim* sim, float margin, uint32_t* non_finite, uint32_t* outside) {
    VFS_CHECK(sim, VFS_ERR_INVALID, "sim is null");
    VFS_CHECK(!sim->batch, VFS_ERR_INVALID, "the state check runs on single-scene simulations");
    VFS_CHECK(margin >= 0.0f, VFS_ERR_INVALID, "margin must be >= 0");
    const SimConsts& c = sim->scenes[0].consts;
    VFS_TRY(launch_check_state(sim->x, sim->v, sim->n, c.bb_lo, c.bb_hi, margin, sim->check_counters, sim->stream));
    uint32_t counts[2] = {0, 0};
    VFS_HIP(hipMemcpyAsync(counts, sim->check_counters, sizeof(counts), hipMemcpyDeviceToHost, sim->stream));
    VFS_HIP(hipStreamSynchronize(sim->stream));
    if (non_finite) *non_finite = counts[0];
    if (outside) *outside = counts[1];
    return 0;
}

int vfs_sim_substeps(vfs_sim* sim, uint32_t n, void* stream) {
    VFS_CHECK(sim, VFS_ERR_INVALID, "sim is null");
    hipStream_t s = stream ? as_stream(stream) : sim->stream;
    VFS_TRY(sim->substeps(n, s));
    if (!(sim->flags & VFS_FLAG_CHECK_STATE) || n == 0) return 0;
    // a particle farther than the box's largest extent outside it has left the domain
    const float* size = sim->scenes[0].params.bbox_size;
    const float margin = std::max(size[0], std::max(size[1], size[2]));
    uint32_t bad = 0, out = 0;
    VFS_TRY(check_state_impl(sim, margin, &bad, &out));
    if (bad || out)
        return fail(VFS_ERR_DIVERGED, "state check after " + std::to_string(n) + " substeps: " +
                                          std::to_string(bad) + " particles with a non-finite position or velocity, " +
                                          std::to_string(out) + " outside the domain");
    return 0;
}

int vfs_sim_check_state(vfs_sim* sim, float margin, uint32_t* non_finite, uint32_t* outside) {
    return check_state_impl(sim, margin, non_finite, outside);
}

int vfs_sim_synchronize(vfs_sim* sim) {
    VFS_CHECK(sim, VFS_ERR_INVALID, "sim is null");
    VFS_HIP(hipStreamSynchronize(sim->stream));
    return 0;
}

int vfs_sim_step(vfs_sim* sim, void* stream) {
    VFS_CHECK(sim, VFS_ERR_INVALID, "sim is null");
    return vfs_sim_substeps(sim, (uint32_t)sim->scenes[0].params.iterations, stream);
}

int vfs_sim_get_state(vfs_sim* sim, float* positions, float* velocities, float* densities,
                      uint32_t* particle_ids) {
    return get_state_impl(sim, 0, positions, velocities, densities, particle_ids);
}

int vfs_sim_get_hash(vfs_sim* sim, uint32_t* sorted_keys, uint32_t* sorted_indices, uint32_t* offsets) {
    return get_hash_impl(sim, 0, sorted_keys, sorted_indices, offsets);
}

int vfs_sim_get_boundary(vfs_sim* sim, float* boundary_density, float* boundary_gradient) {
    VFS_CHECK(!sim || sim->model == 0, VFS_ERR_INVALID, "the Lague model has no boundary objects");
    return get_boundary_impl(sim, 0, boundary_density, boundary_gradient);
}

int vfs_sim_device_buffers(vfs_sim* sim, vfs_device_buffers* out) {
    VFS_CHECK(sim && out, VFS_ERR_INVALID, "null argument");
    out->positions = reinterpret_cast<float*>(sim->x);
    out->velocities = reinterpret_cast<float*>(sim->v);
    out->densities = sim->rho;
    out->accelerations = reinterpret_cast<float*>(sim->accel);
    out->particle_ids = sim->pid;
    out->sorted_keys = sim->sorted_keys;
    out->sorted_indices = sim->perm;
    out->cell_start = sim->cell_start;
    out->n_particles = sim->n;
    return 0;
}

int vfs_sim_get_tuning(const vfs_sim* sim, vfs_tuning* tuning) {
    VFS_CHECK(sim && tuning, VFS_ERR_INVALID, "null argument");
    *tuning = sim->tuning;
    return 0;
}

int vfs_sim_set_tuning(vfs_sim* sim, const vfs_tuning* t) { return set_tuning_impl(sim, t); }

int vfs_sim_autotune(vfs_sim* sim, uint32_t substeps_per_trial, vfs_tuning* best, float* best_ms) {
    return autotune_impl(sim, substeps_per_trial, false, best, best_ms);
}

int vfs_sim_profile_stages(vfs_sim* sim, uint32_t n, float* ms, uint32_t* launches, const char** names,
                           uint32_t max_stages, uint32_t* n_stages) {
    return profile_impl(sim, n, ms, launches, names, max_stages, n_stages);
}

int vfs_sim_neighbour_stats(vfs_sim* sim, double* mean, uint32_t* max_count, uint32_t* overflowed) {
    return neighbour_stats_impl(sim, mean, max_count, overflowed);
}

int vfs_sim_sort_stats(vfs_sim* sim, uint32_t* changed_keys) {
    VFS_CHECK(sim && changed_keys, VFS_ERR_INVALID, "null argument");
    *changed_keys = 0;
    if (sim->n == 0 || !sim->stepped) return 0;
    if (sim->last_full_sort) {
        *changed_keys = sim->n;
        return 0;
    }
    // the re-sort's count of changed keys stays in its state word until the next substep
    VFS_HIP(hipMemcpyAsync(changed_keys, sim->sort_ws + sim->plan.off_rs_state, 4, hipMemcpyDeviceToHost,
                           sim->stream));
    VFS_HIP(hipStreamSynchronize(sim->stream));
    return 0;
}

/* ---- batches ----------------------------------------------------------- */

int vfs_batch_create(const vfs_params* params, uint32_t n_scenes, const vfs_boundary_object* objects,
                     const uint32_t* objects_per_scene, vfs_batch** out) {
    VFS_CHECK(n_scenes >= 1 && objects_per_scene, VFS_ERR_INVALID, "scenes and per-scene object counts needed");
    return create_impl(params, n_scenes, objects, objects_per_scene, true, out);
}

int vfs_batch_destroy(vfs_batch* batch) { return vfs_sim_destroy(batch); }

int vfs_batch_info(const vfs_batch* batch, uint32_t* n_scenes, uint64_t* n_particles, uint32_t* entries) {
    VFS_CHECK(batch, VFS_ERR_INVALID, "batch is null");
    uint64_t total = 0;
    for (const auto& sc : batch->scenes) total += sc.n;
    if (n_scenes) *n_scenes = (uint32_t)batch->scenes.size();
    if (n_particles) *n_particles = total;
    if (entries) *entries = batch->n;
    return 0;
}

int vfs_batch_set_state(vfs_batch* batch, uint32_t scene, const float* positions, const float* velocities) {
    return set_state_impl(batch, scene, positions, velocities, 0, -1);
}

int vfs_batch_set_params(vfs_batch* batch, uint32_t scene, const vfs_params* params) {
    return set_params_impl(batch, scene, params);
}

int vfs_batch_substeps(vfs_batch* batch, uint32_t n, void* stream) { return vfs_sim_substeps(batch, n, stream); }

int vfs_batch_synchronize(vfs_batch* batch) { return vfs_sim_synchronize(batch); }

int vfs_batch_get_state(vfs_batch* batch, uint32_t scene, float* positions, float* velocities, float* densities,
                        uint32_t* particle_ids) {
    return get_state_impl(batch, scene, positions, velocities, densities, particle_ids);
}

int vfs_batch_get_hash(vfs_batch* batch, uint32_t scene, uint32_t* sorted_keys, uint32_t* sorted_indices,
                       uint32_t* offsets) {
    return get_hash_impl(batch, scene, sorted_keys, sorted_indices, offsets);
}

int vfs_batch_get_boundary(vfs_batch* batch, uint32_t scene, float* boundary_density, float* boundary_gradient) {
    return get_boundary_impl(batch, scene, boundary_density, boundary_gradient);
}

int vfs_batch_get_tuning(const vfs_batch* batch, vfs_tuning* tuning) { return vfs_sim_get_tuning(batch, tuning); }

int vfs_batch_set_tuning(vfs_batch* batch, const vfs_tuning* tuning) { return set_tuning_impl(batch, tuning); }

int vfs_batch_autotune(vfs_batch* batch, uint32_t substeps_per_trial, int allow_cell_size, vfs_tuning* best,
                       float* best_ms_per_substep) {
    return autotune_impl(batch, substeps_per_trial, allow_cell_size != 0, best, best_ms_per_substep);
}

int vfs_batch_profile_stages(vfs_batch* batch, uint32_t n, float* ms, uint32_t* launches, const char** names,
                             uint32_t max_stages, uint32_t* n_stages) {
    return profile_impl(batch, n, ms, launches, names, max_stages, n_stages);
}

int vfs_batch_neighbour_stats(vfs_batch* batch, double* mean, uint32_t* max_count, uint32_t* overflowed) {
    return neighbour_stats_impl(batch, mean, max_count, overflowed);
}

}  // extern "C"
