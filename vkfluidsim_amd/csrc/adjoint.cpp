// Differentiable substeps (vfs_adjoint_*): the forward pass keeps the particle
// state every `checkpoint_every` substeps; the backward pass restores each
// checkpoint, recomputes its substeps with the same kernels (hash, sort,
// neighbour lists) while recording what the adjoint needs, and reverses them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "sim_internal.h"

namespace vfs {

int launch_adjoint_substep(const uint32_t* perm, const float4* xs, const float4* vs, const float* rho,
                           const float4* bnd, const uint32_t* nbr, const uint32_t* cnt, const double4* gx_in,
                           const double4* gv_in, double4* abar, double4* gx, double4* gv, double* rbar,
                           double4* gpsi, uint32_t* dup, double4* gx_out, double4* gv_out, double* nu_part,
                           double* nu_bar, uint32_t n, const SimConsts& c, const DevObject* objs, hipStream_t s,
                           StageTimer* timer);
int launch_grid_density(const float* origin, const float* spacing, const uint32_t* dims, const float4* xs,
                        const uint32_t* cell_start, const SimConsts& c, double* out, hipStream_t s);
int launch_grid_loss_grad(const float* origin, const float* spacing, const uint32_t* dims, const float4* x,
                          uint32_t n, const double* rho_g, const float* target, const SimConsts& c, double4* gx,
                          hipStream_t s);
int launch_d_to_f(const double* src, float* dst, uint32_t n, hipStream_t s);
int launch_max_u32(const uint32_t* a, uint32_t n, uint32_t* out, hipStream_t s);
int launch_f3_to_f4(const float* src, float4* dst, uint32_t n, hipStream_t s);
int launch_f4_to_f3(const float4* src, float* dst, uint32_t n, hipStream_t s);
int launch_zero_d4(double4* a, uint32_t n, hipStream_t s);
int launch_d4_to_f3(const double4* src, float* dst, uint32_t n, hipStream_t s);

}  // namespace vfs

struct vfs_adjoint {
    vfs_sim* sim = nullptr;
    uint32_t T = 0, C = 0, n = 0, G = 0;
    float origin[3]{}, spacing[3]{};
    uint32_t dims[3]{};
    // checkpoints of (x, v) at substeps 0, C, 2C, ..., and the final state
    std::vector<float4*> ck_x, ck_v;
    struct Slot {
        uint32_t* perm = nullptr;
        float4 *xs = nullptr, *vs = nullptr, *bnd = nullptr;
        float* rho = nullptr;
        uint32_t *nbr = nullptr, *cnt = nullptr;
    };
    std::vector<Slot> tape;
    // adjoint state in double (the reverse pass accumulates over every substep)
    double4 *gx_a = nullptr, *gv_a = nullptr, *gx_b = nullptr, *gv_b = nullptr;
    double4 *abar = nullptr, *gx = nullptr, *gv = nullptr, *gpsi = nullptr;
    double* rbar = nullptr;
    uint32_t *dup = nullptr, *maxc = nullptr;
    double *nu_bar = nullptr, *nu_part = nullptr;
    double* rho_grid = nullptr;  // sampled density, summed in double
    float* target = nullptr;
    bool forward_done = false;
};

namespace {

void free_adjoint(vfs_adjoint* a) {
    if (!a) return;
    std::vector<void*> ptrs = {a->gx_a, a->gv_a, a->gx_b, a->gv_b, a->abar, a->gx,       a->gv,
                               a->gpsi, a->rbar, a->dup,  a->maxc, a->nu_bar, a->nu_part, a->rho_grid, a->target};
    for (auto* p : a->ck_x) ptrs.push_back(p);
    for (auto* p : a->ck_v) ptrs.push_back(p);
    for (auto& s : a->tape) {
        void* sp[] = {s.perm, s.xs, s.vs, s.bnd, s.rho, s.nbr, s.cnt};
        for (void* p : sp) ptrs.push_back(p);
    }
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete a;
}

// One forward substep writing its sorted state, densities, boundary terms,
// permutation and neighbour lists into a tape slot.  It sorts as the
// simulation's own substeps do: from scratch the first time, then by
// re-sorting the keys that changed bucket.
int record_taped_substep(vfs_sim* sim, const vfs_adjoint::Slot& slot, hipStream_t s, StageTimer* timer) {
    const uint32_t n = sim->n;
    StepBuffers b = sim->buffers();
    b.perm = slot.perm;
    b.xs = slot.xs;
    b.vs = slot.vs;
    b.rho = slot.rho;
    b.bnd = slot.bnd;
    b.nbr = slot.nbr;
    b.nbr_count = slot.cnt;
    if (sim->resort_ready) {
        // re-sort against the last sort's keys and bucket starts (a consistent
        // pair whatever state was restored since); the merge moves x and v
        // into the slot's sorted arrays
        VFS_TRY(resort_keys_iota(sim->plan, sim->keys, sim->sorted_keys, slot.perm, sim->cell_start, sim->sort_ws, s,
                                 timer, &b));
        sim->last_full_sort = false;
    } else {
        VFS_TRY(sort_keys_iota(sim->plan, sim->keys, sim->sorted_keys, slot.perm, sim->sort_ws, s, timer,
                               sim->offsets, (size_t)n + 1, n));
        if (timer) timer->begin(s, "reorder");
        VFS_TRY(launch_reorder_state(b, s));
        if (timer) timer->end(s);
        if (timer) timer->begin(s, "cell_ranges");
        VFS_TRY(launch_cell_start_scan(sim->plan, sim->offsets, sim->cell_start, sim->sort_ws, s));
        if (timer) timer->end(s);
        sim->resort_ready = true;
        sim->last_full_sort = true;
    }
    const SimConsts& c = sim->scenes[0].consts;
    if (timer) timer->begin(s, "density");
    VFS_TRY(launch_density(b, c, nullptr, sim->objs_dev, sim->tuning.block_size, sim->tuning.stage_rows, s));
    if (timer) timer->end(s);
    if (timer) timer->begin(s, "forces_integrate");
    VFS_TRY(launch_forces_integrate(b, c, nullptr, sim->tuning.block_size, s));
    if (timer) timer->end(s);
    return 0;
}

int copy_d2d(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes) VFS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
    return 0;
}

}  // namespace

extern "C" {

int vfs_adjoint_create(vfs_sim* sim, uint32_t n_substeps, uint32_t checkpoint_every, const vfs_density_grid* grid,
                       vfs_adjoint** out) {
    VFS_CHECK(out, VFS_ERR_INVALID, "out is null");
    *out = nullptr;
    VFS_CHECK(sim && grid, VFS_ERR_INVALID, "null argument");
    VFS_CHECK(!sim->batch, VFS_ERR_INVALID, "the adjoint runs on a single-scene simulation");
    VFS_CHECK(sim->model == 0, VFS_ERR_INVALID, "the adjoint differentiates the WCSPH model");
    VFS_CHECK(n_substeps >= 1 && checkpoint_every >= 1, VFS_ERR_INVALID, "substeps and checkpoint interval >= 1");
    VFS_CHECK(sim->tuning.cell_scale == 1.0f, VFS_ERR_INVALID, "the adjoint needs the scene's own cell size");
    VFS_CHECK(sim->n >= 1, VFS_ERR_INVALID, "the adjoint needs at least one particle");
    const uint64_t G = (uint64_t)grid->dims[0] * grid->dims[1] * grid->dims[2];
    VFS_CHECK(G >= 1 && G < (1ull << 31), VFS_ERR_INVALID, "grid dimensions out of range");
    for (int a = 0; a < 3; a++)
        VFS_CHECK(grid->spacing[a] > 0.0f, VFS_ERR_INVALID, "grid spacing must be positive");
    auto* a = new vfs_adjoint();
    a->sim = sim;
    a->T = n_substeps;
    a->C = std::min(checkpoint_every, n_substeps);
    a->n = sim->n;
    a->G = (uint32_t)G;
    for (int k = 0; k < 3; k++) {
        a->origin[k] = grid->origin[k];
        a->spacing[k] = grid->spacing[k];
        a->dims[k] = grid->dims[k];
    }
    const uint32_t n = a->n;
    const uint32_t n_ck = (a->T + a->C - 1) / a->C + 1;  // segment starts + final state
    int r = 0;
    a->ck_x.assign(n_ck, nullptr);
    a->ck_v.assign(n_ck, nullptr);
    for (uint32_t k = 0; k < n_ck; k++) {
        r |= dev_alloc(&a->ck_x[k], n);
        r |= dev_alloc(&a->ck_v[k], n);
    }
    a->tape.resize(a->C);
    for (auto& s : a->tape) {
        r |= dev_alloc(&s.perm, n);
        r |= dev_alloc(&s.xs, (size_t)n + kCandidateGroup);
        r |= dev_alloc(&s.vs, n);
        r |= dev_alloc(&s.bnd, n);
        r |= dev_alloc(&s.rho, n);
        r |= dev_alloc(&s.nbr, (size_t)n * kNeighbourWords);
        r |= dev_alloc(&s.cnt, n);
    }
    r |= dev_alloc(&a->gx_a, n);
    r |= dev_alloc(&a->gv_a, n);
    r |= dev_alloc(&a->gx_b, n);
    r |= dev_alloc(&a->gv_b, n);
    r |= dev_alloc(&a->abar, n);
    r |= dev_alloc(&a->gx, n);
    r |= dev_alloc(&a->gv, n);
    r |= dev_alloc(&a->gpsi, n);
    r |= dev_alloc(&a->rbar, n);
    r |= dev_alloc(&a->dup, n);
    r |= dev_alloc(&a->maxc, 1);
    r |= dev_alloc(&a->nu_bar, 1);
    r |= dev_alloc(&a->nu_part, ((size_t)n * 4 + 255) / 256);  // one per block of the four-lane force pass
    r |= dev_alloc(&a->rho_grid, G);
    r |= dev_alloc(&a->target, G);
    if (r) {
        free_adjoint(a);
        return VFS_ERR_NOMEM;
    }
    *out = a;
    return 0;
}

int vfs_adjoint_destroy(vfs_adjoint* adj) {
    if (adj) {
        (void)hipStreamSynchronize(adj->sim->stream);
        free_adjoint(adj);
    }
    return 0;
}

int vfs_adjoint_forward(vfs_adjoint* a, const float* x0_dev, const float* v0_dev, const float* target_dev,
                        float* density_dev, double* loss) {
    VFS_CHECK(a, VFS_ERR_INVALID, "adjoint is null");
    VFS_CHECK(x0_dev && v0_dev, VFS_ERR_INVALID, "initial state is null");
    vfs_sim* sim = a->sim;
    hipStream_t s = sim->stream;
    const uint32_t n = a->n;
    VFS_HIP(hipStreamSynchronize(s));
    VFS_TRY(launch_f3_to_f4(x0_dev, sim->x, n, s));
    VFS_TRY(launch_f3_to_f4(v0_dev, sim->v, n, s));
    if (n) VFS_HIP(hipMemsetAsync(sim->rho, 0, (size_t)n * 4, s));
    if (sim->pid) VFS_TRY(launch_iota(sim->pid, n, s));
    sim->keys_valid = false;
    uint32_t done = 0, k = 0;
    while (done < a->T) {
        VFS_TRY(copy_d2d(a->ck_x[k], sim->x, (size_t)n * 16, s));
        VFS_TRY(copy_d2d(a->ck_v[k], sim->v, (size_t)n * 16, s));
        const uint32_t steps = std::min(a->C, a->T - done);
        VFS_TRY(sim->substeps(steps, s));
        done += steps;
        k++;
    }
    VFS_TRY(copy_d2d(a->ck_x[k], sim->x, (size_t)n * 16, s));
    VFS_TRY(copy_d2d(a->ck_v[k], sim->v, (size_t)n * 16, s));
    // density field of the final state through its own hash tables
    const SimConsts& c = sim->scenes[0].consts;
    const StepBuffers b = sim->buffers();
    VFS_TRY(sort_keys_iota(sim->plan, sim->keys, sim->sorted_keys, sim->perm, sim->sort_ws, s, nullptr,
                           sim->offsets, (size_t)n + 1, n));
    VFS_TRY(launch_reorder_state(b, s));
    VFS_TRY(launch_cell_start_scan(sim->plan, sim->offsets, sim->cell_start, sim->sort_ws, s));
    VFS_TRY(launch_grid_density(a->origin, a->spacing, a->dims, sim->xs, sim->cell_start, c, a->rho_grid, s));
    if (target_dev)
        VFS_TRY(copy_d2d(a->target, target_dev, (size_t)a->G * 4, s));
    else
        VFS_HIP(hipMemsetAsync(a->target, 0, (size_t)a->G * 4, s));
    if (density_dev) VFS_TRY(launch_d_to_f(a->rho_grid, density_dev, a->G, s));
    std::vector<double> rg(a->G);
    std::vector<float> tg(a->G);
    VFS_HIP(hipMemcpyAsync(rg.data(), a->rho_grid, (size_t)a->G * 8, hipMemcpyDeviceToHost, s));
    VFS_HIP(hipMemcpyAsync(tg.data(), a->target, (size_t)a->G * 4, hipMemcpyDeviceToHost, s));
    VFS_HIP(hipStreamSynchronize(s));
    double sum = 0.0;
    for (uint32_t g = 0; g < a->G; g++) {
        const double d = rg[g] - (double)tg[g];
        sum += d * d;
    }
    if (loss) *loss = sum / (double)a->G;
    sim->stepped = true;
    a->forward_done = true;
    return 0;
}

}  // extern "C"

namespace {

// The reverse pass; timer (may be null) gets an event pair around every
// launch group.
int backward_impl(vfs_adjoint* a, float* grad_x0_dev, float* grad_v0_dev, double* grad_viscosity,
                  StageTimer* timer) {
    VFS_CHECK(a, VFS_ERR_INVALID, "adjoint is null");
    VFS_CHECK(a->forward_done, VFS_ERR_STATE, "vfs_adjoint_forward has not run");
    vfs_sim* sim = a->sim;
    hipStream_t s = sim->stream;
    const uint32_t n = a->n;
    const SimConsts& c = sim->scenes[0].consts;
    const uint32_t n_seg = (a->T + a->C - 1) / a->C;
    // d loss / d x of the final state (its array order), zero velocity adjoint
    if (timer) timer->begin(s, "loss_gradient");
    VFS_TRY(launch_grid_loss_grad(a->origin, a->spacing, a->dims, a->ck_x[n_seg], n, a->rho_grid, a->target, c,
                                  a->gx_a, s));
    if (timer) timer->end(s);
    VFS_TRY(launch_zero_d4(a->gv_a, n, s));
    VFS_HIP(hipMemsetAsync(a->nu_bar, 0, sizeof(double), s));
    double4 *gx_cur = a->gx_a, *gv_cur = a->gv_a, *gx_nxt = a->gx_b, *gv_nxt = a->gv_b;
    for (int seg = (int)n_seg - 1; seg >= 0; seg--) {
        const uint32_t first = (uint32_t)seg * a->C;
        const uint32_t steps = std::min(a->C, a->T - first);
        VFS_TRY(copy_d2d(sim->x, a->ck_x[seg], (size_t)n * 16, s));
        VFS_TRY(copy_d2d(sim->v, a->ck_v[seg], (size_t)n * 16, s));
        sim->keys_valid = false;
        VFS_TRY(sim->ensure_keys(s));
        for (uint32_t k = 0; k < steps; k++) VFS_TRY(record_taped_substep(sim, a->tape[k], s, timer));
        VFS_HIP(hipMemsetAsync(a->maxc, 0, 4, s));
        if (timer) timer->begin(s, "list_check");
        for (uint32_t k = 0; k < steps; k++) VFS_TRY(launch_max_u32(a->tape[k].cnt, n, a->maxc, s));
        if (timer) timer->end(s);
        uint32_t maxc = 0;
        VFS_HIP(hipMemcpyAsync(&maxc, a->maxc, 4, hipMemcpyDeviceToHost, s));
        VFS_HIP(hipStreamSynchronize(s));
        VFS_CHECK(maxc <= kNeighbourCap, VFS_ERR_STATE,
                  "a neighbour list exceeded VFS_NEIGHBOUR_CAPACITY; the adjoint needs complete lists");
        for (int k = (int)steps - 1; k >= 0; k--) {
            const auto& t = a->tape[k];
            VFS_TRY(launch_adjoint_substep(t.perm, t.xs, t.vs, t.rho, t.bnd, t.nbr, t.cnt, gx_cur, gv_cur, a->abar,
                                           a->gx, a->gv, a->rbar, a->gpsi, a->dup, gx_nxt, gv_nxt, a->nu_part,
                                           a->nu_bar, n, c, sim->objs_dev, s, timer));
            std::swap(gx_cur, gx_nxt);
            std::swap(gv_cur, gv_nxt);
        }
    }
    if (grad_x0_dev) VFS_TRY(launch_d4_to_f3(gx_cur, grad_x0_dev, n, s));
    if (grad_v0_dev) VFS_TRY(launch_d4_to_f3(gv_cur, grad_v0_dev, n, s));
    double nb = 0.0;
    VFS_HIP(hipMemcpyAsync(&nb, a->nu_bar, sizeof(double), hipMemcpyDeviceToHost, s));
    // leave the simulation at the forward's final state
    VFS_TRY(copy_d2d(sim->x, a->ck_x[n_seg], (size_t)n * 16, s));
    VFS_TRY(copy_d2d(sim->v, a->ck_v[n_seg], (size_t)n * 16, s));
    sim->keys_valid = false;
    VFS_HIP(hipStreamSynchronize(s));
    if (grad_viscosity) *grad_viscosity = nb;
    return 0;
}

}  // namespace

extern "C" {

int vfs_adjoint_backward(vfs_adjoint* a, float* grad_x0_dev, float* grad_v0_dev, double* grad_viscosity) {
    return backward_impl(a, grad_x0_dev, grad_v0_dev, grad_viscosity, nullptr);
}

int vfs_adjoint_profile_backward(vfs_adjoint* a, float* ms, uint32_t* launches, const char** names,
                                 uint32_t max_stages, uint32_t* n_stages) {
    VFS_CHECK(a && n_stages, VFS_ERR_INVALID, "null argument");
    *n_stages = 0;
    static const char* kNames[] = {"loss_gradient", "sort_prepare", "sort_scan",   "sort_histogram", "sort_small",
                                   "sort_pass",     "sort_merge",   "reorder",     "cell_ranges",    "density",
                                   "forces_integrate", "list_check", "adj_integrate", "adj_forces",
                                   "adj_viscosity_sum", "adj_boundary", "adj_density"};
    constexpr uint32_t kN = sizeof(kNames) / sizeof(kNames[0]);
    EventTimer timer;
    VFS_TRY(backward_impl(a, nullptr, nullptr, nullptr, &timer));  // synchronises the stream
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
        if (cnt[i] == 0) continue;
        if (ms) ms[m] = (float)tot[i];
        if (launches) launches[m] = cnt[i];
        if (names) names[m] = kNames[i];
        m++;
    }
    *n_stages = m;
    return 0;
}

}  // extern "C"
