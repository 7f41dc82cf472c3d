// Internal host launchers of the spatial-hash kernels (sort.hip) and the WCSPH
// kernels (wcsph.hip).  Not part of the C ABI.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>

#include "vfs_common.h"

namespace vfs {

constexpr int kMaxPasses = 4;
constexpr int kCellStartItems = 8;

// Optional per-stage event timing (used by vfs_sim_profile_stages).
struct StageTimer {
    virtual void begin(hipStream_t s, const char* name) = 0;
    virtual void end(hipStream_t s) = 0;
    virtual ~StageTimer() = default;
};

// Word offsets into one u32 workspace.
struct SortPlan {
    uint32_t n;
    int bits;        // digit width (8, 9 or 10)
    int items;       // keys per thread (tile = 256 * items)
    int passes;
    uint32_t num_tiles;
    size_t off_tmp[4];
    size_t off_hist, off_counters, off_status;
    uint32_t cs_tiles;                  // tiles of the cell-start scan over n + 1 entries
    size_t off_cs_counter, off_cs_status;
    // re-sort (changed keys sorted and merged into the previous order); the
    // changed keys are sorted with rs_bits-wide digits (fewer passes)
    int rs_bits, rs_passes;
    uint32_t rs_tiles;
    size_t off_rs_nb, off_rs_bkeys, off_rs_bvals, off_rs_bks, off_rs_bis;
    // previous bucket starts and each key's count of smaller changed keys
    // (n + 1 entries each), written by the bucket-start update for the merge
    size_t off_rs_old_cs, off_rs_lb;
    size_t off_rs_tile_counts, off_rs_state;
    size_t zero_begin;
    size_t workspace_words;
};

constexpr int kResortItems = 16;  // keys per thread of the re-sort's count and compaction
// resort = false: a plan for full sorts only (sort_keys_iota), without the
// re-sort's buffers.
SortPlan make_sort_plan(uint32_t n, uint32_t max_value, int bits_per_pass, int items, bool resort = true);
int valid_sort_config(int bits, int items);

// Stable sort of (keys_in, iota): keys_out = sorted keys, perm_out[i] = input
// index of the i-th smallest (key, index).  keys_in is not modified.
// The preparation kernel of the sort also fills fill[0..fill_words) with
// fill_value when fill is not null (the offsets table of the substep).
int sort_keys_iota(const SortPlan& plan, const uint32_t* keys_in, uint32_t* keys_out, uint32_t* perm_out,
                   uint32_t* workspace, hipStream_t s, StageTimer* timer, uint32_t* fill = nullptr,
                   size_t fill_words = 0, uint32_t fill_value = 0);
// The same stable sort when keys_out holds an ascending key array (the
// previous substep's sorted keys) and cell_start its bucket starts: keys
// equal to the previous key at their index keep their relative order, the
// changed ones are radix-sorted on their own (device-side count) and merged;
// cell_start is then updated in place to the bucket starts of the new keys
// (plan.n + 1 entries; no offsets table is written).  With move, the merge
// also reorders the particle state: move->xs[i] = move->x[perm_out[i]], the
// same for v, pid and pred when present (each row written by the thread that
// places its key, so no separate gather over the permutation).
struct StepBuffers;
int resort_keys_iota(const SortPlan& plan, const uint32_t* keys_in, uint32_t* keys_out, uint32_t* perm_out,
                     uint32_t* cell_start, uint32_t* workspace, hipStream_t s, StageTimer* timer,
                     const StepBuffers* move = nullptr);

// keys[i] = key_base + HashCell3D(floor(pos_i / cell)) % table_size
int launch_hash_keys(const float* pos, uint32_t stride, uint32_t n, float cell, uint32_t table_size,
                     uint32_t* keys, hipStream_t s, uint32_t key_base = 0);
// table[sorted_keys[i]] = i where a key starts (CalculateOffsets).
int launch_key_starts(const uint32_t* sorted_keys, uint32_t n, uint32_t* table, hipStream_t s);
int launch_fill_value(uint32_t* a, size_t count, uint32_t value, hipStream_t s);
// cell_start[k] = min(table[k..n]) for k in [0, n] (table has n + 1 entries, table[n] = n).
int launch_cell_start_scan(const SortPlan& plan, const uint32_t* table, uint32_t* cell_start,
                           uint32_t* workspace, hipStream_t s);
int launch_reorder_f32x3(const uint32_t* perm, const float* src, float* dst, uint32_t n, hipStream_t s);
int launch_iota(uint32_t* a, uint32_t n, hipStream_t s);
int launch_copy_words(const uint32_t* src, uint32_t* dst, uint32_t n, hipStream_t s);
size_t scan_workspace_words(uint32_t n);
int launch_scan_exclusive(uint32_t* data, uint32_t n, uint32_t* workspace, hipStream_t s);

// ---- WCSPH substep kernels (wcsph.hip)
struct StepBuffers {
    // state in the current order (input of a substep, output of the previous one)
    float4* x;
    float4* v;
    uint32_t* pid;       // nullable
    uint32_t* keys;      // hash keys of x
    // sorted / reordered
    uint32_t* sorted_keys;
    uint32_t* perm;
    uint32_t* offsets;     // n + 1: reference offsets table (n for empty keys), entry n = n
    uint32_t* cell_start;  // n + 1
    float4* xs;          // n + kCandidateGroup entries; w: p / rho^2 once the density pass ran
    float4* vs;          // w: mass / max(rho, 1e-6) once the density pass ran
    uint32_t* pids;      // nullable
    // per-particle results
    float* rho;
    float4* bnd;         // (grad psi xyz, psi)
    float4* accel;       // nullable
    uint32_t* nbr;       // (nbr_cap + 4) * n: neighbour lists of the substep in chunks of four,
                         // chunk-major (uint4 chunk g of particle i at g * n + i)
    uint32_t* nbr_count; // n: list lengths (> nbr_cap: overflowed)
    // Lague model only (null otherwise): predicted positions in the current
    // order, their sorted copy, near densities
    float4* pred;
    float4* ps;
    float* near_rho;
    uint32_t n;          // entries of every per-particle array (all scenes of a batch, padded)
};

// Candidates per group of the density pass's cell walk (its last group reads
// past the range, so xs is padded by this many elements).
constexpr uint32_t kCandidateGroup = 4;

int launch_reorder_state(const StepBuffers& b, hipStream_t s);
// counters[0] = particles with a non-finite x or v component, counters[1] =
// particles farther than margin outside [lo, hi] (counters cleared first).
int launch_check_state(const float4* x, const float4* v, uint32_t n, const float* lo, const float* hi, float margin,
                       uint32_t* counters, hipStream_t s);
// Lague model stages (one scene): predictions and their hash keys, density
// and near density, pressure + viscosity + integration + collisions.
int launch_lague_predict(const StepBuffers& b, const SimConsts& c, const LagueConsts& lc, hipStream_t s);
int launch_lague_density(const StepBuffers& b, const SimConsts& c, const LagueConsts& lc, hipStream_t s);
int launch_lague_forces(const StepBuffers& b, const SimConsts& c, const LagueConsts& lc, hipStream_t s);
// block: 128 or 256 threads; staged: stage stencil rows of positions in LDS.
// table == nullptr: one scene with constants c over all b.n entries; else a
// batch whose waves read their scene's constants from the table (c unused).
int launch_density(const StepBuffers& b, const SimConsts& c, const SceneTable* table, const DevObject* objs,
                   int block, int staged, hipStream_t s);
int launch_forces_integrate(const StepBuffers& b, const SimConsts& c, const SceneTable* table, int block,
                            hipStream_t s);
// Waves of a padded particle launch of n entries (blocks of up to 256 threads,
// rounded to a multiple of 8 blocks): the size of a batch's wave_scene table.
inline uint32_t launch_waves(uint32_t n) { return ((((n + 255u) / 256u) + 7u) & ~7u) * 4u; }

// Volume-map/SDF grid generator for analytic shapes (FP64 on the device).
int launch_volume_map(int kind, const double* shape_params, const uint32_t* res, double support_radius,
                      double tolerance, const float* dmin, const double* step, const double* gl_x,
                      const double* gl_w, float* sdf_out, float* vm_out, hipStream_t s);

}  // namespace vfs
