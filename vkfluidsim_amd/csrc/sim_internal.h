// Internal definition of a simulation (one scene or a batch of scenes) shared
// by the host files of the library.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "sort_internal.h"

namespace vfs {

// One scene's segment of the particle arrays.
struct SceneSeg {
    vfs_params params{};
    SimConsts consts{};
    uint32_t n = 0;       // particles
    uint32_t base = 0;    // first entry
    uint32_t padded = 0;  // entries reserved (n rounded up to 64 in a batch)
    float cell0 = 0.0f;   // the scene's own cell size (cell_size, or smooth_radius)
};

// A HIP event pair around every timed launch group, kept with its stage name
// until the caller has synchronised and read the elapsed times.
struct EventTimer : StageTimer {
    struct Rec {
        std::string name;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    std::string pending;
    hipEvent_t pa = nullptr;
    hipError_t err = hipSuccess;
    void begin(hipStream_t s, const char* name) override {
        pending = name;
        if (hipEventCreate(&pa) != hipSuccess) err = hipErrorUnknown;
        (void)hipEventRecord(pa, s);
    }
    void end(hipStream_t s) override {
        hipEvent_t pb;
        if (hipEventCreate(&pb) != hipSuccess) err = hipErrorUnknown;
        (void)hipEventRecord(pb, s);
        recs.push_back({pending, pa, pb});
    }
    ~EventTimer() override {
        for (auto& r : recs) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
    }
};

}  // namespace vfs

using namespace vfs;

struct vfs_sim {
    std::vector<SceneSeg> scenes;
    bool batch = false;
    uint32_t flags = 0;
    uint32_t n = 0;  // entries of every per-particle array
    hipStream_t stream = nullptr;
    int device = 0;

    float4 *x = nullptr, *v = nullptr, *xs = nullptr, *vs = nullptr, *bnd = nullptr, *accel = nullptr;
    float* rho = nullptr;
    uint32_t *nbr = nullptr, *nbr_count = nullptr;
    uint32_t* check_counters = nullptr;  // state check: non-finite, escaped particles
    uint32_t *pid = nullptr, *pids = nullptr, *keys = nullptr, *sorted_keys = nullptr, *perm = nullptr,
             *offsets = nullptr, *cell_start = nullptr, *sort_ws = nullptr;
    SimConsts* consts_dev = nullptr;  // batch: scene table (+ one empty scene)
    uint32_t* wave_scene = nullptr;   // batch: scene of every wave
    SortPlan plan{};
    std::vector<DevObject> objs_host;
    DevObject* objs_dev = nullptr;
    std::vector<float*> grids;

    // model: 0 = WCSPHWithBoundaryModel, 1 = LagueModel (predictions and near
    // densities in pred / ps / near_rho)
    int model = 0;
    LagueConsts lague{};
    vfs_lague_params lague_params{};
    float4 *pred = nullptr, *ps = nullptr;
    float* near_rho = nullptr;

    bool keys_valid = false;
    bool stepped = false;
    // sorted_keys holds an ascending key array and cell_start its bucket
    // starts (true after the first substep): later substeps re-sort
    bool resort_ready = false;
    bool last_full_sort = false;  // the last substep sorted every key (no re-sort)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    vfs_tuning tuning{8, 16, 128, 1, 1.0f};

    virtual ~vfs_sim() = default;

    StepBuffers buffers() const {
        StepBuffers b{};
        b.x = x;
        b.v = v;
        b.pid = pid;
        b.keys = keys;
        b.sorted_keys = sorted_keys;
        b.perm = perm;
        b.offsets = offsets;
        b.cell_start = cell_start;
        b.xs = xs;
        b.vs = vs;
        b.pids = pids;
        b.rho = rho;
        b.bnd = bnd;
        b.accel = accel;
        b.nbr = nbr;
        b.nbr_count = nbr_count;
        b.pred = pred;
        b.ps = ps;
        b.near_rho = near_rho;
        b.n = n;
        return b;
    }

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }

    // The scene table of a batch (rows in scene order, then an empty scene for
    // the waves past the last segment).
    int upload_consts() {
        if (!batch) return 0;
        std::vector<SimConsts> rows;
        for (const auto& sc : scenes) rows.push_back(sc.consts);
        // waves past the last segment: no particles, keys (never looked up)
        // inside the tables
        SimConsts none = scenes[0].consts;
        none.n = 0;
        none.base = 0;
        none.obj_count = 0;
        rows.push_back(none);
        VFS_HIP(hipMemcpyAsync(consts_dev, rows.data(), rows.size() * sizeof(SimConsts), hipMemcpyHostToDevice,
                               stream));
        VFS_HIP(hipStreamSynchronize(stream));
        return 0;
    }

    // Stable sort of this substep's keys, then the reorder and the bucket
    // starts.  Once a previous substep's sorted keys exist the keys are
    // re-sorted against them and the bucket starts updated from theirs;
    // otherwise a full sort (whose preparation kernel resets the offsets
    // table to n), the reorder's key starts and the reverse min-scan.
    int sort_reorder_ranges(const StepBuffers& b, hipStream_t s, StageTimer* timer) {
        if (resort_ready)  // the merge of the re-sort also reorders the state
            return resort_keys_iota(plan, keys, sorted_keys, perm, cell_start, sort_ws, s, timer, &b);
        VFS_TRY(sort_keys_iota(plan, keys, sorted_keys, perm, sort_ws, s, timer, offsets, (size_t)n + 1, n));
        if (timer) timer->begin(s, "reorder");
        VFS_TRY(launch_reorder_state(b, s));  // + key starts into the offsets table
        if (timer) timer->end(s);
        if (timer) timer->begin(s, "cell_ranges");
        VFS_TRY(launch_cell_start_scan(plan, offsets, cell_start, sort_ws, s));
        if (timer) timer->end(s);
        return 0;
    }

    // LagueModel::Step body (lague_model.cpp:106-122) for one substep.
    int record_lague_substep(hipStream_t s, StageTimer* timer) {
        const StepBuffers b = buffers();
        const SimConsts& c = scenes[0].consts;
        if (timer) timer->begin(s, "predict");
        VFS_TRY(launch_lague_predict(b, c, lague, s));
        if (timer) timer->end(s);
        VFS_TRY(sort_reorder_ranges(b, s, timer));
        if (timer) timer->begin(s, "density");
        VFS_TRY(launch_lague_density(b, c, lague, s));
        if (timer) timer->end(s);
        if (timer) timer->begin(s, "forces_integrate");
        VFS_TRY(launch_lague_forces(b, c, lague, s));
        if (timer) timer->end(s);
        return 0;
    }

    int record_substep(hipStream_t s, StageTimer* timer) {
        if (model == 1) return record_lague_substep(s, timer);
        const StepBuffers b = buffers();
        const SceneTable table{consts_dev, wave_scene};
        const SceneTable* tp = batch ? &table : nullptr;
        VFS_TRY(sort_reorder_ranges(b, s, timer));
        if (timer) timer->begin(s, "density");
        VFS_TRY(launch_density(b, scenes[0].consts, tp, objs_dev, tuning.block_size, tuning.stage_rows, s));
        if (timer) timer->end(s);
        if (timer) timer->begin(s, "forces_integrate");
        VFS_TRY(launch_forces_integrate(b, scenes[0].consts, tp, tuning.block_size, s));
        if (timer) timer->end(s);
        return 0;
    }

    int ensure_keys(hipStream_t s) {
        if (keys_valid || model == 1) return 0;  // the Lague substep hashes its own predictions
        for (const auto& sc : scenes)
            VFS_TRY(launch_hash_keys(reinterpret_cast<const float*>(x + sc.base), 4, sc.n, sc.consts.cell, sc.n,
                                     keys + sc.base, s, sc.base));
        keys_valid = true;
        return 0;
    }

    int ensure_graph() {
        if (exec) return 0;
        VFS_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        const int r = record_substep(stream, nullptr);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(stream, &g);
        if (r != 0) {
            if (g) (void)hipGraphDestroy(g);
            return r;
        }
        if (e != hipSuccess) return fail(VFS_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        graph = g;
        VFS_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        return 0;
    }

    // Cell size = each scene's own cell size times `scale` (1 keeps the
    // reference's hash, and with it bit parity).
    int set_cell_scale(float scale) {
        for (auto& sc : scenes) sc.consts.cell = sc.cell0 * scale;
        keys_valid = false;
        drop_graph();
        return upload_consts();
    }

    // Device copy of the substep input (x, v, ids, keys, densities); used by
    // the autotuners so that every trial starts from the same state.
    struct Snapshot {
        float4 *x = nullptr, *v = nullptr;
        float* rho = nullptr;
        uint32_t *pid = nullptr, *keys = nullptr;
        bool keys_valid = false;
        float cell_scale = 1.0f;
    };

    int snapshot(Snapshot& sn, bool restore) {
        const size_t n4 = (size_t)n * 16, n1 = (size_t)n * 4;
        auto cp = [&](void* dst, const void* src, size_t bytes) {
            return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream);
        };
        if (!restore) {
            if (hipMalloc(&sn.x, n4 ? n4 : 16) != hipSuccess || hipMalloc(&sn.v, n4 ? n4 : 16) != hipSuccess ||
                hipMalloc(&sn.rho, n1 ? n1 : 4) != hipSuccess || hipMalloc(&sn.keys, n1 ? n1 : 4) != hipSuccess ||
                (pid && hipMalloc(&sn.pid, n1 ? n1 : 4) != hipSuccess))
                return fail(VFS_ERR_NOMEM, "snapshot allocation failed");
            sn.keys_valid = keys_valid;
            sn.cell_scale = tuning.cell_scale;
            if (n && (cp(sn.x, x, n4) || cp(sn.v, v, n4) || cp(sn.rho, rho, n1) || cp(sn.keys, keys, n1) ||
                      (pid && cp(sn.pid, pid, n1))))
                return fail(VFS_ERR_HIP, "snapshot copy failed");
        } else {
            if (n && (cp(x, sn.x, n4) || cp(v, sn.v, n4) || cp(rho, sn.rho, n1) || cp(keys, sn.keys, n1) ||
                      (pid && cp(pid, sn.pid, n1))))
                return fail(VFS_ERR_HIP, "snapshot restore failed");
            // saved keys hash the saved cell size
            keys_valid = sn.keys_valid && sn.cell_scale == tuning.cell_scale;
        }
        VFS_HIP(hipStreamSynchronize(stream));
        return 0;
    }

    static void free_snapshot(Snapshot& sn) {
        void* ptrs[] = {sn.x, sn.v, sn.rho, sn.pid, sn.keys};
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
        sn = Snapshot{};
    }

    // Makes `waiter` wait for the work queued so far on `signaller`, through
    // one event per simulation (created on first use, destroyed with it): a
    // wait takes the event's latest record at the time of the call, so later
    // records do not disturb earlier waits, and substep calls on a caller's
    // stream create no events.
    hipEvent_t order_event = nullptr;
    int order_after(hipStream_t waiter, hipStream_t signaller) {
        if (!order_event) VFS_HIP(hipEventCreateWithFlags(&order_event, hipEventDisableTiming));
        VFS_HIP(hipEventRecord(order_event, signaller));
        VFS_HIP(hipStreamWaitEvent(waiter, order_event, 0));
        return 0;
    }

    int substeps(uint32_t count, hipStream_t s) {
        if (n == 0 || count == 0) return 0;
        if (s != stream) VFS_TRY(order_after(s, stream));
        VFS_TRY(ensure_keys(s));
        last_full_sort = false;
        if (!resort_ready) {
            // the first substep sorts from scratch; the graph records the re-sort
            VFS_TRY(record_substep(s, nullptr));
            resort_ready = true;
            last_full_sort = count == 1;
            count--;
        }
        if (!(flags & VFS_FLAG_NO_GRAPH)) {
            if (count) VFS_TRY(ensure_graph());
            for (uint32_t k = 0; k < count; k++) VFS_HIP(hipGraphLaunch(exec, s));
        } else {
            for (uint32_t k = 0; k < count; k++) VFS_TRY(record_substep(s, nullptr));
        }
        // later calls (state reads, set_state) synchronise the simulation stream
        if (s != stream) VFS_TRY(order_after(stream, s));
        stepped = true;
        return 0;
    }
};


namespace vfs {
// Allocation helper (count elements of T, at least one).
template <class T>
int dev_alloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    const hipError_t e = hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
    if (e != hipSuccess) return fail(VFS_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return 0;
}
}  // namespace vfs
