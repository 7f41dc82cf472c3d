// Spatial-hash kernels for gfx950: cell-key hashing, a stable onesweep LSD
// radix sort (wave64 ballot ranking staged in LDS, decoupled look-back across
// tiles), cell ranges / reference-format offsets, reorder, and a standalone
// decoupled look-back exclusive scan.
//
// Replaces the reference compute layer (luihabl/VkFluidSim):
//   UpdateSpatialHash        assets/shaders/simulation/spatial_hash/update_spatial_hash.slang:18-28
//   ClearCounts..CopyBack    assets/shaders/simulation/spatial_hash/sort.slang:13-55
//   ScanBlock/ScanCombine    assets/shaders/simulation/spatial_hash/scan.slang:13-83
//   InitOffsets/Calculate..  assets/shaders/simulation/spatial_hash/spatial_offsets.slang:10-31
//   Reorder                  assets/shaders/simulation/spatial_hash/reorder.slang:11-19
// The reference counting sort scatters with atomics, so the order inside a
// bucket is not defined there; this sort is stable (ascending input index
// inside a key), which is the canonical order the oracle uses.
#include "sort_internal.h"

namespace vfs {

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr uint32_t kFlagAggregate = 1u << 30;
constexpr uint32_t kFlagInclusive = 2u << 30;
constexpr uint32_t kValueMask = (1u << 30) - 1u;

__device__ __forceinline__ uint64_t lanemask_lt() {
    const uint32_t lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__device__ __forceinline__ uint32_t ld_acquire_agent(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_release_agent(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive scan of one value per thread over a 256-thread block.
template <int WAVES = kWaves>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_tmp, uint32_t& total) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_tmp[w] = x;
    __syncthreads();
    uint32_t wave_off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < WAVES; i++) {
        const uint32_t t = s_tmp[i];
        if (i < w) wave_off += t;
        tot += t;
    }
    __syncthreads();
    total = tot;
    return wave_off + x - v;
}

// ------------------------------------------------------------------ kernels

// UpdateSpatialHash: key = HashCell3D(floor(x / cell)) % table_size.
__global__ __launch_bounds__(256) void k_hash_keys(const float* __restrict__ pos, uint32_t stride,
                                                   uint32_t n, float cell, FastMod fm, uint32_t key_base,
                                                   uint32_t* __restrict__ keys) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* p = pos + (size_t)i * stride;
    const int32_t cx = floor_to_cell(p[0] / cell);
    const int32_t cy = floor_to_cell(p[1] / cell);
    const int32_t cz = floor_to_cell(p[2] / cell);
    keys[i] = key_base + fastmod(hash_cell(cx, cy, cz), fm);
}

__global__ __launch_bounds__(256) void k_zero_words(uint32_t* __restrict__ a, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        a[i] = 0u;
}

// Digit histograms of every pass in one read of the keys.
template <int BITS>
__global__ __launch_bounds__(256) void k_radix_histogram(const uint32_t* __restrict__ keys, uint32_t n,
                                                         int passes, uint32_t* __restrict__ hist,
                                                         const uint32_t* __restrict__ n_dev) {
    constexpr int RADIX = 1 << BITS;
    __shared__ uint32_t sh[kMaxPasses * RADIX];
    if (n_dev) n = *n_dev;  // element count known on the device only
    for (int i = threadIdx.x; i < passes * RADIX; i += kThreads) sh[i] = 0;
    __syncthreads();
    for (uint32_t idx = blockIdx.x * kThreads + threadIdx.x; idx < n; idx += gridDim.x * kThreads) {
        const uint32_t k = keys[idx];
        for (int p = 0; p < passes; p++) atomicAdd(&sh[p * RADIX + ((k >> (p * BITS)) & (RADIX - 1))], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < passes * RADIX; i += kThreads) {
        const uint32_t c = sh[i];
        if (c) atomicAdd(&hist[i], c);
    }
}

// One onesweep pass over digit `pass`.  Each block claims the next tile in
// launch order, ranks its keys with wave ballots (stable: wave-major, then
// item-major, then lane order = input order), publishes its per-digit counts
// and resolves its global offsets by decoupled look-back over earlier tiles.
template <int BITS, int ITEMS>
__global__ __launch_bounds__(256) void k_radix_pass(const uint32_t* __restrict__ keys_in,
                                                    const uint32_t* __restrict__ vals_in,
                                                    uint32_t* __restrict__ keys_out,
                                                    uint32_t* __restrict__ vals_out, uint32_t n, int shift,
                                                    const uint32_t* __restrict__ hist,
                                                    uint32_t* __restrict__ status,
                                                    uint32_t* __restrict__ tile_counter,
                                                    const uint32_t* __restrict__ n_dev, uint32_t small_cap) {
    constexpr int RADIX = 1 << BITS;
    constexpr int DPT = RADIX / kThreads;  // digits per thread
    constexpr int TILE = kThreads * ITEMS;
    static_assert(DPT >= 1, "radix must be >= 256");

    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_wave[kWaves][RADIX];
    __shared__ uint32_t s_local[RADIX];
    __shared__ int32_t s_global[RADIX];
    __shared__ uint32_t s_keys[TILE];
    __shared__ uint32_t s_vals[TILE];
    __shared__ uint32_t s_tmp[kWaves];

    // a device-side count of at most small_cap keys was sorted by k_resort_rank
    if (n_dev) {
        n = *n_dev;
        if (n <= small_cap) return;
    }
    // Only the first ceil(n / TILE) blocks claim tiles (they get tiles 0, 1, ...
    // in claim order); the others leave without touching the counter or LDS.
    if (blockIdx.x >= (n + (uint32_t)TILE - 1u) / (uint32_t)TILE) return;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid == 0) s_tile = atomicAdd(tile_counter, 1u);
    for (int i = tid; i < kWaves * RADIX; i += kThreads) (&s_wave[0][0])[i] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t base = tile * (uint32_t)TILE;  // < n: tiles past the count are never claimed
    const uint32_t wave_base = base + (uint32_t)w * 64u * ITEMS;

    uint32_t key[ITEMS], val[ITEMS], rank[ITEMS];
#pragma unroll
    for (int it = 0; it < ITEMS; it++) {
        const uint32_t idx = wave_base + (uint32_t)it * 64u + (uint32_t)lane;
        const bool valid = idx < n;
        key[it] = valid ? keys_in[idx] : 0xFFFFFFFFu;
        val[it] = vals_in ? (valid ? vals_in[idx] : 0u) : idx;
    }
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int it = 0; it < ITEMS; it++) {
        const uint32_t idx = wave_base + (uint32_t)it * 64u + (uint32_t)lane;
        const bool valid = idx < n;
        const uint32_t d = (key[it] >> shift) & (RADIX - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < BITS; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        if (valid) {
            const uint32_t before = (uint32_t)__popcll(peers & lt);
            const uint32_t prior = s_wave[w][d];
            rank[it] = prior + before;
            if (before == 0) s_wave[w][d] = prior + (uint32_t)__popcll(peers);
        }
    }
    __syncthreads();

    // Per digit: exclusive offsets across waves and the tile total.
    uint32_t tcount[DPT];
    uint32_t tsum = 0;
#pragma unroll
    for (int dd = 0; dd < DPT; dd++) {
        const int d = tid * DPT + dd;
        uint32_t s = 0;
#pragma unroll
        for (int ww = 0; ww < kWaves; ww++) {
            const uint32_t c = s_wave[ww][d];
            s_wave[ww][d] = s;
            s += c;
        }
        tcount[dd] = s;
        tsum += s;
        // Publish the aggregate before looking back so that later tiles progress.
        st_release_agent(&status[(size_t)tile * RADIX + d], (tile == 0 ? kFlagInclusive : kFlagAggregate) | s);
    }
    // Local (in-tile) digit starts.
    uint32_t tot;
    uint32_t lrun = block_exclusive_scan(tsum, s_tmp, tot);
    // Global digit starts from the pass histogram.
    uint32_t hv[DPT];
    uint32_t hsum = 0;
#pragma unroll
    for (int dd = 0; dd < DPT; dd++) {
        hv[dd] = hist[tid * DPT + dd];
        hsum += hv[dd];
    }
    uint32_t grun = block_exclusive_scan(hsum, s_tmp, tot);

#pragma unroll
    for (int dd = 0; dd < DPT; dd++) {
        const int d = tid * DPT + dd;
        uint32_t excl = 0;
        if (tile > 0) {
            int64_t t = (int64_t)tile - 1;
            while (true) {
                const uint32_t s = ld_acquire_agent(&status[(size_t)t * RADIX + d]);
                const uint32_t flag = s & ~kValueMask;
                if (flag == 0) continue;
                excl += s & kValueMask;
                if (flag == kFlagInclusive) break;
                t--;
            }
            st_release_agent(&status[(size_t)tile * RADIX + d], kFlagInclusive | (excl + tcount[dd]));
        }
        s_local[d] = lrun;
        s_global[d] = (int32_t)(grun + excl) - (int32_t)lrun;
        lrun += tcount[dd];
        grun += hv[dd];
    }
    __syncthreads();

#pragma unroll
    for (int it = 0; it < ITEMS; it++) {
        const uint32_t idx = wave_base + (uint32_t)it * 64u + (uint32_t)lane;
        if (idx < n) {
            const uint32_t d = (key[it] >> shift) & (RADIX - 1);
            const uint32_t lpos = s_local[d] + s_wave[w][d] + rank[it];
            s_keys[lpos] = key[it];
            s_vals[lpos] = val[it];
        }
    }
    __syncthreads();

    const uint32_t tile_valid = (n - base) < (uint32_t)TILE ? (n - base) : (uint32_t)TILE;
    for (uint32_t i = tid; i < tile_valid; i += kThreads) {
        const uint32_t k = s_keys[i];
        const uint32_t d = (k >> shift) & (RADIX - 1);
        const uint32_t dst = (uint32_t)(s_global[d] + (int32_t)i);
        keys_out[dst] = k;
        vals_out[dst] = s_vals[i];
    }
}

// Stable sort of at most kRankSort changed (key, value) pairs (count read on
// the device) by counting: the entry at compacted position p (the compaction
// keeps index order, so position order is value order) goes to rank
// #{q : (key_q, q) < (key_p, p)}.  Each block copies the changed keys to LDS
// and ranks kRankKeys entries; every entry's comparisons are split into
// kRankWaves * kRankSplit equal shares of q (lane l of wave w takes entry
// l % kRankKeys and share w * kRankSplit + l / kRankKeys, reading four keys per
// 16-byte LDS load that its lanes share), and the first wave adds the shares
// up and scatters.  A larger count leaves at once (the multi-tile radix passes
// sort it).  The work is count^2 comparisons over count / kRankKeys blocks.
constexpr uint32_t kRankSort = 16384;
constexpr int kRankWaves = 16;
constexpr uint32_t kRankSplit = 4;  // 16 entries per block (2 and 1 measured slower)
constexpr uint32_t kRankKeys = 64u / kRankSplit;
static_assert(kRankKeys * kRankSplit == 64u, "lanes split evenly over the entries");
__global__ __launch_bounds__(kRankWaves * 64) void k_resort_rank(const uint32_t* __restrict__ keys_in,
                                                                 const uint32_t* __restrict__ vals_in,
                                                                 uint32_t* __restrict__ keys_out,
                                                                 uint32_t* __restrict__ vals_out,
                                                                 const uint32_t* __restrict__ n_dev) {
    __shared__ uint4 s_keys4[kRankSort / 4u];
    __shared__ uint32_t s_rank[kRankWaves][64];
    const uint32_t n = *n_dev;
    const uint32_t first = blockIdx.x * kRankKeys;
    if (n > kRankSort || first >= n) return;
    uint32_t* s_keys = reinterpret_cast<uint32_t*>(s_keys4);
    for (uint32_t i = threadIdx.x; i < n; i += kRankWaves * 64u) s_keys[i] = keys_in[i];
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint32_t p = first + lane % kRankKeys;
    const uint32_t share = w * kRankSplit + lane / kRankKeys;
    const uint32_t kp = p < n ? s_keys[p] : 0xFFFFFFFFu;
    const uint64_t mine = ((uint64_t)kp << 32) | p;
    // shares of a multiple of four keys, so that every 16-byte read is aligned
    const uint32_t chunk = ((n + kRankWaves * kRankSplit - 1u) / (kRankWaves * kRankSplit) + 3u) & ~3u;
    const uint32_t q0 = min(n, share * chunk), q1 = min(n, q0 + chunk);
    uint32_t rank = 0;
    uint32_t q = q0;
    for (; q + 4u <= q1; q += 4u) {
        const uint4 k4 = s_keys4[q >> 2];
        rank += ((((uint64_t)k4.x << 32) | q) < mine) ? 1u : 0u;
        rank += ((((uint64_t)k4.y << 32) | (q + 1u)) < mine) ? 1u : 0u;
        rank += ((((uint64_t)k4.z << 32) | (q + 2u)) < mine) ? 1u : 0u;
        rank += ((((uint64_t)k4.w << 32) | (q + 3u)) < mine) ? 1u : 0u;
    }
    for (; q < q1; q++) rank += ((((uint64_t)s_keys[q] << 32) | q) < mine) ? 1u : 0u;
    s_rank[w][lane] = rank;
    __syncthreads();
    if (w == 0 && lane < kRankKeys && p < n) {
        uint32_t r = 0;
#pragma unroll
        for (int i = 0; i < kRankWaves; i++)
#pragma unroll
            for (uint32_t h = 0; h < kRankSplit; h++) r += s_rank[i][h * kRankKeys + lane];
        keys_out[r] = kp;
        vals_out[r] = vals_in[p];
    }
}

// Reference offsets table from sorted keys: T[k] = first i with sorted[i] == k
// (CalculateOffsets, spatial_offsets.slang:19-31).  T must hold the empty
// sentinel for absent keys beforehand (InitOffsets) when a full table is wanted.
__global__ __launch_bounds__(256) void k_key_starts(const uint32_t* __restrict__ sk, uint32_t n,
                                                    uint32_t* __restrict__ table) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = sk[i];
    if ((i == 0 || sk[i - 1] != k) && k < n) table[k] = i;
}

// Exclusive min over one value per thread of a 256-thread block (identity ~0u).
__device__ __forceinline__ uint32_t block_exclusive_min(uint32_t v, uint32_t* s_tmp, uint32_t& total) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x = min(x, y);
    }
    const uint32_t ex_wave = __shfl_up(x, 1, 64);
    if (lane == 63) s_tmp[w] = x;
    __syncthreads();
    uint32_t wave_off = 0xFFFFFFFFu, tot = 0xFFFFFFFFu;
#pragma unroll
    for (int i = 0; i < kWaves; i++) {
        const uint32_t t = s_tmp[i];
        if (i < w) wave_off = min(wave_off, t);
        tot = min(tot, t);
    }
    __syncthreads();
    total = tot;
    return lane == 0 ? wave_off : min(wave_off, ex_wave);
}

// Cell ranges from the offsets table: cell_start[k] = min(T[k..m-1]) over the
// m = n + 1 entries (T[n] = n), i.e. the start of the next present key, so
// bucket k is [cell_start[k], cell_start[k+1]).  A reverse inclusive min-scan
// in one pass: tiles are claimed from the top of the table down and resolve
// their prefix by decoupled look-back over the tiles above them.
template <int ITEMS>
__global__ __launch_bounds__(256) void k_cell_start_scan(const uint32_t* __restrict__ table, uint32_t m,
                                                         uint32_t* __restrict__ cell_start,
                                                         uint32_t* __restrict__ status,
                                                         uint32_t* __restrict__ tile_counter) {
    constexpr int TILE = kThreads * ITEMS;
    __shared__ uint32_t s_data[TILE];
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_tmp[kWaves];
    __shared__ uint32_t s_prefix;
    const int tid = threadIdx.x;
    if (tid == 0) s_tile = atomicAdd(tile_counter, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const int64_t top = (int64_t)m - (int64_t)tile * TILE;  // exclusive upper index of this tile
    // element e of the tile is table index top - 1 - e
    for (int e = tid; e < TILE; e += kThreads) {
        const int64_t idx = top - 1 - e;
        s_data[e] = idx >= 0 ? table[idx] : 0xFFFFFFFFu;
    }
    __syncthreads();
    uint32_t local[ITEMS];
    uint32_t run = 0xFFFFFFFFu;
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        run = min(run, s_data[tid * ITEMS + j]);
        local[j] = run;
    }
    uint32_t total;
    const uint32_t excl = block_exclusive_min(run, s_tmp, total);
    if (tid < 64) {
        // Wave-parallel look-back: lane l inspects tile (base - l); a window
        // resolves once every tile up to the nearest inclusive one is ready.
        const int lane = tid;
        const uint32_t tv = total & kValueMask;
        if (lane == 0) st_release_agent(&status[tile], (tile == 0 ? kFlagInclusive : kFlagAggregate) | tv);
        uint32_t prefix = kValueMask;
        if (tile > 0) {
            int64_t base = (int64_t)tile - 1;
            while (true) {
                const int64_t t = base - lane;
                const uint32_t sv = t >= 0 ? ld_acquire_agent(&status[t]) : (kFlagInclusive | kValueMask);
                const uint32_t flag = sv & ~kValueMask;
                const uint64_t incl = __ballot(flag == kFlagInclusive);
                const uint64_t wait = __ballot(flag == 0);
                const int first = incl ? __ffsll((unsigned long long)incl) - 1 : 64;
                const uint64_t upto = first >= 63 ? ~0ull : ((2ull << first) - 1ull);
                if (wait & upto) continue;
                uint32_t v = lane <= first ? (sv & kValueMask) : kValueMask;
#pragma unroll
                for (int o = 32; o >= 1; o >>= 1) v = min(v, (uint32_t)__shfl_xor(v, o, 64));
                prefix = min(prefix, v);
                if (first < 64) break;
                base -= 64;
            }
            if (lane == 0) st_release_agent(&status[tile], kFlagInclusive | min(prefix, tv));
        }
        if (lane == 0) s_prefix = prefix;
    }
    __syncthreads();
    const uint32_t pre = min(s_prefix, excl);
#pragma unroll
    for (int j = 0; j < ITEMS; j++) s_data[tid * ITEMS + j] = min(pre, local[j]);
    __syncthreads();
    for (int e = tid; e < TILE; e += kThreads) {
        const int64_t idx = top - 1 - e;
        if (idx >= 0) cell_start[idx] = s_data[e];
    }
}

// Zeroes one region and fills another with a value (per-substep preparation).
__global__ __launch_bounds__(256) void k_prepare(uint32_t* __restrict__ zero, size_t zero_words,
                                                 uint32_t* __restrict__ fill, size_t fill_words,
                                                 uint32_t fill_value) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < zero_words; i += stride) zero[i] = 0u;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < fill_words; i += stride)
        fill[i] = fill_value;
}

__global__ __launch_bounds__(256) void k_reorder_f32x3(const uint32_t* __restrict__ perm,
                                                       const float* __restrict__ src,
                                                       float* __restrict__ dst, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = perm[i];
    dst[3 * (size_t)i + 0] = src[3 * (size_t)s + 0];
    dst[3 * (size_t)i + 1] = src[3 * (size_t)s + 1];
    dst[3 * (size_t)i + 2] = src[3 * (size_t)s + 2];
}

__global__ __launch_bounds__(256) void k_iota(uint32_t* __restrict__ a, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

__global__ __launch_bounds__(256) void k_copy_words(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                                    uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// Decoupled look-back exclusive scan (in place).  64-bit status words:
// high half = flag (1 aggregate, 2 inclusive), low half = value mod 2^32.
template <int ITEMS>
__global__ __launch_bounds__(256) void k_scan_lookback(uint32_t* __restrict__ data, uint32_t n,
                                                       unsigned long long* __restrict__ status,
                                                       uint32_t* __restrict__ tile_counter) {
    constexpr int TILE = kThreads * ITEMS;
    __shared__ uint32_t s_data[TILE];
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_tmp[kWaves];
    __shared__ uint32_t s_prefix;
    const int tid = threadIdx.x;
    if (tid == 0) s_tile = atomicAdd(tile_counter, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t base = tile * (uint32_t)TILE;
    for (int i = tid; i < TILE; i += kThreads) {
        const uint32_t idx = base + i;
        s_data[i] = idx < n ? data[idx] : 0u;
    }
    __syncthreads();
    uint32_t local[ITEMS];
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        local[j] = s_data[tid * ITEMS + j];
        sum += local[j];
    }
    uint32_t total;
    const uint32_t excl_thread = block_exclusive_scan(sum, s_tmp, total);
    if (tid == 0) {
        const unsigned long long agg = ((unsigned long long)(tile == 0 ? 2u : 1u) << 32) | total;
        __hip_atomic_store(&status[tile], agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t prefix = 0;
        if (tile > 0) {
            int64_t t = (int64_t)tile - 1;
            while (true) {
                const unsigned long long s =
                    __hip_atomic_load(&status[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const uint32_t flag = (uint32_t)(s >> 32);
                if (flag == 0) continue;
                prefix += (uint32_t)s;
                if (flag == 2) break;
                t--;
            }
            __hip_atomic_store(&status[tile], (2ull << 32) | (uint32_t)(prefix + total), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        s_prefix = prefix;
    }
    __syncthreads();
    uint32_t run = s_prefix + excl_thread;
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        s_data[tid * ITEMS + j] = run;
        run += local[j];
    }
    __syncthreads();
    for (int i = tid; i < TILE; i += kThreads) {
        const uint32_t idx = base + i;
        if (idx < n) data[idx] = s_data[i];
    }
}

// ------------------------------------------------------------ re-sort
// Between substeps few particles change cell, so the new keys mostly equal
// the previous substep's sorted keys at the same index ("unchanged").  The
// previous sorted keys ascend, so the unchanged particles already form a run
// sorted by (key, index); the changed ones (set B, compacted in index order)
// are sorted by the radix kernels with a device-side count, and both runs are
// merged by position:
//   unchanged j:  out = (j - #B before j) + #{b in B : (key_b, b) < (key_j, j)}
//   B entry r:    out = (c - #B before c) + r, with c = clamp(b, L, U) and
//                 [L, U) the previous bucket of key_b (previous cell starts)
// which is exactly the stable sort, for any number of changed keys.

// Changed keys per tile.  Block 0 also clears the digit histograms and tile
// counters of the radix passes that sort the changed keys.
template <int ITEMS>
__global__ __launch_bounds__(256) void k_resort_count(const uint32_t* __restrict__ keys,
                                                      const uint32_t* __restrict__ old_keys, uint32_t n,
                                                      uint32_t* __restrict__ tile_counts,
                                                      uint32_t* __restrict__ clear, uint32_t clear_words) {
    constexpr int TILE = kThreads * ITEMS;
    __shared__ uint32_t s_tmp[kWaves];
    if (blockIdx.x == 0)
        for (uint32_t i = threadIdx.x; i < clear_words; i += kThreads) clear[i] = 0u;
    const uint32_t base = blockIdx.x * (uint32_t)TILE;
    // every load of the thread is issued before the first compare (indices
    // past n read the last entry and count nothing)
    uint32_t a[ITEMS], b[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        const uint32_t idx = min(base + (uint32_t)(j * kThreads + (int)threadIdx.x), n - 1u);
        a[j] = keys[idx];
        b[j] = old_keys[idx];
    }
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        const uint32_t idx = base + (uint32_t)(j * kThreads + (int)threadIdx.x);
        c += (idx < n && a[j] != b[j]) ? 1u : 0u;
    }
    uint32_t total;
    (void)block_exclusive_scan(c, s_tmp, total);
    if (threadIdx.x == 0) tile_counts[blockIdx.x] = total;
}

// Exclusive count nb of changed keys before each index (tile prefix from the
// tile counts, then a block scan), the changed keys and indices compacted in
// index order, and on the last tile the total in state[0] and nb[n].  For the
// radix passes over the compacted keys it adds their digit histograms
// (`passes` digits of `bits` bits) to hist and clears the look-back status of
// the pass tiles (radix_tile keys each, status rows of `radix` words per
// tile, pass_words per pass) that its changed keys fall in; later tiles
// leave without reading any status.
template <int ITEMS>
__global__ __launch_bounds__(256) void k_resort_compact(const uint32_t* __restrict__ keys,
                                                        const uint32_t* __restrict__ old_keys, uint32_t n,
                                                        const uint32_t* __restrict__ tile_counts,
                                                        uint32_t* __restrict__ nb, uint32_t* __restrict__ bkeys,
                                                        uint32_t* __restrict__ bvals, uint32_t* __restrict__ state,
                                                        uint32_t* __restrict__ hist, int passes, int bits,
                                                        uint32_t* __restrict__ status, uint32_t radix_tile,
                                                        size_t pass_words) {
    constexpr int TILE = kThreads * ITEMS;
    __shared__ uint8_t s_flag[TILE];
    __shared__ uint32_t s_data[TILE];
    __shared__ uint32_t s_tmp[kWaves];
    __shared__ uint32_t s_hist[kMaxPasses << 10];
    const uint32_t radix = 1u << bits, mask = radix - 1u;
    for (uint32_t i = threadIdx.x; i < (uint32_t)passes * radix; i += kThreads) s_hist[i] = 0u;
    const int tid = threadIdx.x;
    const uint32_t tile = blockIdx.x;
    const uint32_t base = tile * (uint32_t)TILE;
    uint32_t prefix;
    // the thread's keys (entries tid + j * kThreads of the tile), every load
    // issued before the first compare
    uint32_t kk[ITEMS];
    {
        uint32_t ko[ITEMS];
#pragma unroll
        for (int j = 0; j < ITEMS; j++) {
            const uint32_t idx = min(base + (uint32_t)(j * kThreads + tid), n - 1u);
            kk[j] = keys[idx];
            ko[j] = old_keys[idx];
        }
        // the tile's prefix while those loads are in flight: every block sums
        // the counts of the tiles before it (tiles^2 / 2 reads of an
        // L2-resident array: 2M at 8M particles; at the 2^28-particle limit
        // about 0.3% of a substep)
        uint32_t before = 0;
        for (uint32_t t = (uint32_t)tid; t < tile; t += kThreads) before += tile_counts[t];
        (void)block_exclusive_scan(before, s_tmp, prefix);
#pragma unroll
        for (int j = 0; j < ITEMS; j++) {
            const uint32_t idx = base + (uint32_t)(j * kThreads + tid);
            s_flag[j * kThreads + tid] = (idx < n && kk[j] != ko[j]) ? 1 : 0;
        }
    }
    __syncthreads();
    uint32_t local[ITEMS];
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        local[j] = s_flag[tid * ITEMS + j];
        sum += local[j];
    }
    uint32_t total;
    uint32_t run = prefix + block_exclusive_scan(sum, s_tmp, total);
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        s_data[tid * ITEMS + j] = run;
        run += local[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        const int i = j * kThreads + tid;
        const uint32_t idx = base + (uint32_t)i;
        if (idx >= n) continue;
        const uint32_t e = s_data[i];
        nb[idx] = e;
        if (s_flag[i]) {
            const uint32_t k = kk[j];
            bkeys[e] = k;
            bvals[e] = idx;
            for (int ps = 0; ps < passes; ps++) atomicAdd(&s_hist[ps * radix + ((k >> (ps * bits)) & mask)], 1u);
        }
    }
    if (total) {
        // status rows of the pass tiles holding changed keys prefix .. prefix + total - 1
        const uint32_t t0 = prefix / radix_tile, t1 = (prefix + total - 1u) / radix_tile;
        const uint32_t words = (t1 - t0 + 1u) * radix;
        for (int ps = 0; ps < passes; ps++)
            for (uint32_t i = tid; i < words; i += kThreads) status[ps * pass_words + (size_t)t0 * radix + i] = 0u;
    }
    if (tile == gridDim.x - 1u && tid == 0) {
        nb[n] = prefix + total;
        state[0] = prefix + total;
    }
    __syncthreads();
    for (uint32_t i = tid; i < (uint32_t)passes * radix; i += kThreads) {
        const uint32_t c = s_hist[i];
        if (c) atomicAdd(&hist[i], c);
    }
}

// First index in [lo, hi) of the sorted changed keys with a key >= k (hi if none).
__device__ __forceinline__ uint32_t changed_lower_bound(const uint32_t* __restrict__ bks, uint32_t k, uint32_t lo,
                                                        uint32_t hi) {
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (bks[mid] < k)
            lo = mid + 1u;
        else
            hi = mid;
    }
    return lo;
}

// First index in [lo, hi) where pred is false, pred being true on a prefix of
// the range; the whole wave calls it with the same arguments.  Each round the
// 64 lanes test the last index of 64 equal chunks, and the count of true tests
// picks the chunk holding the answer: a range of c entries takes about
// log64(c) + 1 dependent loads instead of log2(c).
template <class Pred>
__device__ __forceinline__ uint32_t wave_partition_point(uint32_t lo, uint32_t hi, Pred&& pred) {
    const uint32_t lane = threadIdx.x & 63u;
    while (hi - lo > 64u) {
        const uint32_t step = (hi - lo + 63u) >> 6;
        const uint32_t probe = lo + (lane + 1u) * step - 1u;
        const bool t = probe < hi && pred(probe);
        lo += (uint32_t)__popcll(__ballot(t)) * step;  // a true test is below hi, so lo <= hi
        hi = min(hi, lo + step);
    }
    const uint32_t probe = lo + lane;
    const bool t = probe < hi && pred(probe);
    return lo + (uint32_t)__popcll(__ballot(t));
}

// Output positions of both runs: threads below the changed count place the
// sorted changed entry of their index (previous bucket starts in old_cs), every
// unchanged index adds the changed entries that precede it: lb[key] of them
// have a smaller key (written by k_resort_cell_starts), and those with the
// same key and a lower index sit at the front of [lb[key], lb[key + 1]) (a
// search only when that range is not empty).  MOVE: each placed entry also
// moves its particle's rows to the sorted place (Reorder, reorder.slang:11-19,
// as a scatter), x and v (and ids, predictions when present) from mv's
// current-order arrays into its sorted ones.
template <bool MOVE>
__global__ __launch_bounds__(256) void k_resort_merge(const uint32_t* __restrict__ keys,
                                                      const uint32_t* __restrict__ nb, uint32_t n,
                                                      const uint32_t* __restrict__ state,
                                                      const uint32_t* __restrict__ bks,
                                                      const uint32_t* __restrict__ bis,
                                                      const uint32_t* __restrict__ old_cs,
                                                      const uint32_t* __restrict__ lb,
                                                      uint32_t* __restrict__ keys_out,
                                                      uint32_t* __restrict__ perm_out, StepBuffers mv) {
    const uint32_t g = blockIdx.x * (uint32_t)kThreads + threadIdx.x;
    if (g >= n) return;
    const uint32_t count = state[0];
    if (g < count) {
        const uint32_t kb = bks[g], ib = bis[g];
        const uint32_t lo = old_cs[kb], hi = old_cs[kb + 1u];
        const uint32_t c = min(max(ib, lo), hi);
        const uint32_t pos = (c - nb[c]) + g;
        if (pos < n) {  // always, for an ascending previous key array and its bucket starts
            keys_out[pos] = kb;
            perm_out[pos] = ib;
            if (MOVE) {
                mv.xs[pos] = mv.x[ib];
                mv.vs[pos] = mv.v[ib];
                if (mv.pid) mv.pids[pos] = mv.pid[ib];
                if (mv.ps) mv.ps[pos] = mv.pred[ib];
            }
        }
    }
    const uint32_t e0 = nb[g];
    if (nb[g + 1u] != e0) return;  // changed: placed by its sorted entry
    const uint32_t kj = keys[g];
    // the row of an unchanged entry is read before its place is searched
    float4 xg, vg, pg;
    uint32_t ig = 0;
    if (MOVE) {
        xg = mv.x[g];
        vg = mv.v[g];
        if (mv.pid) ig = mv.pid[g];
        if (mv.ps) pg = mv.pred[g];
    }
    uint32_t lo = lb[kj], hi = lb[kj + 1u];  // kj < n: lb has n + 1 entries
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (bis[mid] < g)
            lo = mid + 1u;
        else
            hi = mid;
    }
    const uint32_t out = (g - e0) + lo;
    if (out < n) {
        keys_out[out] = kj;
        perm_out[out] = g;
        if (MOVE) {
            mv.xs[out] = xg;
            mv.vs[out] = vg;
            if (mv.pid) mv.pids[out] = ig;
            if (mv.ps) mv.ps[out] = pg;
        }
    }
}

// Bucket starts of the re-sorted keys, in place over the previous ones: the
// entries with a key below k are the unchanged entries at previous indices
// below cs[k] (cs[k] - nb[cs[k]] of them) and the changed keys below k (a
// lower bound in the sorted changed keys).  The lower bounds of the block's
// keys lie between those of its first key and of the key after its last, so
// waves 0 and 1 find those two among all changed keys (a wave-wide search)
// and every thread searches only that window.  The previous start goes to
// old_cs and the lower bound to lb, for the merge that runs after it.
// m = table size + 1 entries.
__global__ __launch_bounds__(256) void k_resort_cell_starts(uint32_t* __restrict__ cs, const uint32_t* __restrict__ nb,
                                                            const uint32_t* __restrict__ bks,
                                                            const uint32_t* __restrict__ state, uint32_t m,
                                                            uint32_t* __restrict__ old_cs, uint32_t* __restrict__ lb) {
    __shared__ uint32_t s_window[2];
    const uint32_t k0 = blockIdx.x * (uint32_t)kThreads;
    const uint32_t k = k0 + threadIdx.x;
    const uint32_t count = state[0];
    uint32_t below = 0u, c = 0u;
    if (k < m) {
        c = cs[k];
        below = c - nb[c];
    }
    const uint32_t wave = threadIdx.x >> 6;
    if (wave < 2u) {
        const uint32_t kp = wave == 0u ? k0 : k0 + (uint32_t)kThreads;
        const uint32_t r = wave_partition_point(0u, count, [&](uint32_t i) { return bks[i] < kp; });
        if ((threadIdx.x & 63u) == 0u) s_window[wave] = r;
    }
    __syncthreads();
    if (k >= m) return;
    const uint32_t l = changed_lower_bound(bks, k, s_window[0], s_window[1]);
    old_cs[k] = c;
    lb[k] = l;
    cs[k] = below + l;
}

inline uint32_t grid_for(uint64_t n, uint32_t block = 256) {
    return (uint32_t)((n + block - 1) / block);
}

int bit_width(uint32_t v) {
    int b = 0;
    while (v) {
        b++;
        v >>= 1;
    }
    return b;
}

}  // namespace

// ------------------------------------------------------------ host launchers

SortPlan make_sort_plan(uint32_t n, uint32_t max_value, int bits_per_pass, int items, bool resort) {
    SortPlan p{};
    p.n = n;
    p.bits = bits_per_pass;
    p.items = items;
    const int kb = bit_width(max_value);
    p.passes = kb == 0 ? 1 : (kb + bits_per_pass - 1) / bits_per_pass;
    const uint32_t tile = (uint32_t)kThreads * (uint32_t)items;
    p.num_tiles = n == 0 ? 0 : (n + tile - 1) / tile;
    const size_t radix = (size_t)1 << bits_per_pass;
    // layout: tmp keys/vals x2, hist, counters, status
    size_t off = 0;
    auto take = [&](size_t words) {
        const size_t o = off;
        off += (words + 63) & ~(size_t)63;
        return o;
    };
    p.off_tmp[0] = take(n);
    p.off_tmp[1] = take(n);
    p.off_tmp[2] = take(n);
    p.off_tmp[3] = take(n);
    // a plan without re-sort gives its buffers no words (they are never used)
    const size_t rs_n = resort ? (size_t)n : 0;
    p.off_rs_nb = take(resort ? rs_n + 1 : 0);
    p.off_rs_bkeys = take(rs_n);
    p.off_rs_bvals = take(rs_n);
    p.off_rs_bks = take(rs_n);
    p.off_rs_bis = take(rs_n);
    p.off_rs_old_cs = take(resort ? rs_n + 1 : 0);
    p.off_rs_lb = take(resort ? rs_n + 1 : 0);
    p.rs_bits = 10;
    p.rs_passes = kb == 0 ? 1 : (kb + p.rs_bits - 1) / p.rs_bits;
    const size_t rs_radix = resort ? (size_t)1 << p.rs_bits : 0;
    p.off_hist = take(kMaxPasses * std::max(radix, rs_radix));
    p.off_counters = take(kMaxPasses);
    p.off_status = take(std::max((size_t)p.passes * radix, (size_t)p.rs_passes * rs_radix) * p.num_tiles);
    const uint32_t cs_tile = (uint32_t)kThreads * kCellStartItems;
    p.cs_tiles = (uint32_t)(((uint64_t)n + 1 + cs_tile - 1) / cs_tile);
    p.off_cs_counter = take(1);
    p.off_cs_status = take(p.cs_tiles);
    const uint32_t rs_tile = (uint32_t)kThreads * kResortItems;
    p.rs_tiles = (n == 0 || !resort) ? 0 : (n + rs_tile - 1) / rs_tile;
    p.off_rs_tile_counts = take(p.rs_tiles);
    p.off_rs_state = take(2);
    p.zero_begin = p.off_hist;
    p.workspace_words = off;
    return p;
}

int valid_sort_config(int bits, int items) {
    return (bits == 8 || bits == 9 || bits == 10) && (items == 8 || items == 12 || items == 16);
}

template <int BITS, int ITEMS>
static void launch_pass(const uint32_t* ki, const uint32_t* vi, uint32_t* ko, uint32_t* vo, uint32_t n, int shift,
                        const uint32_t* hist, uint32_t* status, uint32_t* counter, uint32_t tiles,
                        const uint32_t* n_dev, uint32_t small_cap, hipStream_t s) {
    hipLaunchKernelGGL((k_radix_pass<BITS, ITEMS>), dim3(tiles), dim3(kThreads), 0, s, ki, vi, ko, vo, n, shift,
                       hist, status, counter, n_dev, small_cap);
}

template <int BITS>
static void launch_pass_items(int items, const uint32_t* ki, const uint32_t* vi, uint32_t* ko, uint32_t* vo,
                              uint32_t n, int shift, const uint32_t* hist, uint32_t* status, uint32_t* counter,
                              uint32_t tiles, const uint32_t* n_dev, uint32_t small_cap, hipStream_t s) {
    if (items == 8)
        launch_pass<BITS, 8>(ki, vi, ko, vo, n, shift, hist, status, counter, tiles, n_dev, small_cap, s);
    else if (items == 12)
        launch_pass<BITS, 12>(ki, vi, ko, vo, n, shift, hist, status, counter, tiles, n_dev, small_cap, s);
    else
        launch_pass<BITS, 16>(ki, vi, ko, vo, n, shift, hist, status, counter, tiles, n_dev, small_cap, s);
}

// Stable radix sort of (keys_in, vals_in or iota) into keys_out / vals_out;
// n_dev: element count read on the device (at most p.n).
static int radix_sort(const SortPlan& p, int bits, int passes, const uint32_t* keys_in, const uint32_t* vals_in,
                      uint32_t* keys_out, uint32_t* vals_out, const uint32_t* n_dev, uint32_t* workspace,
                      hipStream_t s, StageTimer* timer, bool have_histograms = false, uint32_t small_cap = 0) {
    const uint32_t n = p.n;
    const size_t radix = (size_t)1 << bits;
    uint32_t* hist = workspace + p.off_hist;
    uint32_t* counters = workspace + p.off_counters;
    uint32_t* status = workspace + p.off_status;
    if (!have_histograms) {
        if (timer) timer->begin(s, "sort_histogram");
        const uint32_t hblocks = std::max<uint32_t>(1u, std::min<uint32_t>(grid_for(n, 256u * 16u), 1024u));
        if (bits == 8)
            hipLaunchKernelGGL(k_radix_histogram<8>, dim3(hblocks), dim3(256), 0, s, keys_in, n, passes, hist,
                               n_dev);
        else if (bits == 9)
            hipLaunchKernelGGL(k_radix_histogram<9>, dim3(hblocks), dim3(256), 0, s, keys_in, n, passes, hist,
                               n_dev);
        else
            hipLaunchKernelGGL(k_radix_histogram<10>, dim3(hblocks), dim3(256), 0, s, keys_in, n, passes, hist,
                               n_dev);
        VFS_TRY(check_launch("sort_histogram"));
        if (timer) timer->end(s);
    }
    uint32_t* tmpk[2] = {workspace + p.off_tmp[0], workspace + p.off_tmp[2]};
    uint32_t* tmpv[2] = {workspace + p.off_tmp[1], workspace + p.off_tmp[3]};
    for (int pass = 0; pass < passes; pass++) {
        const uint32_t* ki = pass == 0 ? keys_in : tmpk[(pass - 1) & 1];
        const uint32_t* vi = pass == 0 ? vals_in : tmpv[(pass - 1) & 1];
        const bool last = pass == passes - 1;
        uint32_t* ko = last ? keys_out : tmpk[pass & 1];
        uint32_t* vo = last ? vals_out : tmpv[pass & 1];
        uint32_t* st = status + (size_t)pass * p.num_tiles * radix;
        if (timer) timer->begin(s, "sort_pass");
        if (bits == 8)
            launch_pass_items<8>(p.items, ki, vi, ko, vo, n, pass * bits, hist + pass * radix, st,
                                 counters + pass, p.num_tiles, n_dev, small_cap, s);
        else if (bits == 9)
            launch_pass_items<9>(p.items, ki, vi, ko, vo, n, pass * bits, hist + pass * radix, st,
                                 counters + pass, p.num_tiles, n_dev, small_cap, s);
        else
            launch_pass_items<10>(p.items, ki, vi, ko, vo, n, pass * bits, hist + pass * radix, st,
                                  counters + pass, p.num_tiles, n_dev, small_cap, s);
        VFS_TRY(check_launch("sort_pass"));
        if (timer) timer->end(s);
    }
    return 0;
}

// Preparation (zeroed sort workspace, filled offsets table), then the radix
// sort of all keys, or (cell_start != nullptr) the re-sort against the
// previous sorted keys in keys_out and their bucket starts.
static int sort_impl(const SortPlan& p, const uint32_t* keys_in, uint32_t* keys_out, uint32_t* perm_out,
                     uint32_t* cell_start, uint32_t* workspace, hipStream_t s, StageTimer* timer,
                     uint32_t* fill, size_t fill_words, uint32_t fill_value, const StepBuffers* move) {
    const uint32_t n = p.n;
    if (n == 0) return 0;
    if (!cell_start) {
        const size_t zero_words = p.workspace_words - p.zero_begin;
        if (timer) timer->begin(s, "sort_prepare");
        hipLaunchKernelGGL(k_prepare, dim3(std::min<uint32_t>(grid_for(std::max(zero_words, fill_words)), 2048u)),
                           dim3(256), 0, s, workspace + p.zero_begin, zero_words, fill, fill ? fill_words : 0,
                           fill_value);
        VFS_TRY(check_launch("sort_prepare"));
        if (timer) timer->end(s);
        return radix_sort(p, p.bits, p.passes, keys_in, nullptr, keys_out, perm_out, nullptr, workspace, s, timer);
    }
    // re-sort: the count and compaction kernels clear what the radix passes
    // over the changed keys need and build their histograms

    uint32_t* nb = workspace + p.off_rs_nb;
    uint32_t* bkeys = workspace + p.off_rs_bkeys;
    uint32_t* bvals = workspace + p.off_rs_bvals;
    uint32_t* bks = workspace + p.off_rs_bks;
    uint32_t* bis = workspace + p.off_rs_bis;
    uint32_t* counts = workspace + p.off_rs_tile_counts;
    uint32_t* state = workspace + p.off_rs_state;
    if (timer) timer->begin(s, "sort_scan");
    const uint32_t clear_words = (uint32_t)(p.off_counters + kMaxPasses - p.off_hist);  // histograms + tile counters
    hipLaunchKernelGGL(k_resort_count<kResortItems>, dim3(p.rs_tiles), dim3(256), 0, s, keys_in, keys_out, n, counts,
                       workspace + p.off_hist, clear_words);
    VFS_TRY(check_launch("sort_scan"));
    const size_t rs_radix = (size_t)1 << p.rs_bits;
    hipLaunchKernelGGL(k_resort_compact<kResortItems>, dim3(p.rs_tiles), dim3(256), 0, s, keys_in, keys_out, n,
                       counts, nb, bkeys, bvals, state, workspace + p.off_hist, p.rs_passes, p.rs_bits,
                       workspace + p.off_status, (uint32_t)(kThreads * p.items), (size_t)p.num_tiles * rs_radix);
    VFS_TRY(check_launch("sort_scan"));
    if (timer) timer->end(s);
    // up to kRankSort changed keys are ranked by counting, more by the radix passes
    static_assert(kMaxPasses * 1024 >= 3 * 1024, "histograms of the 10-bit re-sort digits fit");
    if (timer) timer->begin(s, "sort_small");
    hipLaunchKernelGGL(k_resort_rank, dim3((std::min(n, kRankSort) + kRankKeys - 1u) / kRankKeys), dim3(kRankWaves * 64), 0, s, bkeys,
                       bvals, bks, bis, state);
    VFS_TRY(check_launch("sort_small"));
    if (timer) timer->end(s);
    VFS_TRY(radix_sort(p, p.rs_bits, p.rs_passes, bkeys, bvals, bks, bis, state, workspace, s, timer, true,
                       kRankSort));
    // the bucket starts first: they also leave the previous starts and each
    // key's count of smaller changed keys for the merge
    uint32_t* old_cs = workspace + p.off_rs_old_cs;
    uint32_t* lb = workspace + p.off_rs_lb;
    if (timer) timer->begin(s, "cell_ranges");
    hipLaunchKernelGGL(k_resort_cell_starts, dim3(grid_for((uint64_t)n + 1)), dim3(256), 0, s, cell_start, nb, bks,
                       state, n + 1u, old_cs, lb);
    VFS_TRY(check_launch("cell_ranges"));
    if (timer) timer->end(s);
    if (timer) timer->begin(s, "sort_merge");
    if (move)
        hipLaunchKernelGGL(k_resort_merge<true>, dim3(grid_for(n)), dim3(256), 0, s, keys_in, nb, n, state, bks,
                           bis, old_cs, lb, keys_out, perm_out, *move);
    else
        hipLaunchKernelGGL(k_resort_merge<false>, dim3(grid_for(n)), dim3(256), 0, s, keys_in, nb, n, state, bks,
                           bis, old_cs, lb, keys_out, perm_out, StepBuffers{});
    VFS_TRY(check_launch("sort_merge"));
    if (timer) timer->end(s);
    return 0;
}

int sort_keys_iota(const SortPlan& p, const uint32_t* keys_in, uint32_t* keys_out, uint32_t* perm_out,
                   uint32_t* workspace, hipStream_t s, StageTimer* timer, uint32_t* fill, size_t fill_words,
                   uint32_t fill_value) {
    return sort_impl(p, keys_in, keys_out, perm_out, nullptr, workspace, s, timer, fill, fill_words, fill_value,
                     nullptr);
}

int resort_keys_iota(const SortPlan& p, const uint32_t* keys_in, uint32_t* keys_out, uint32_t* perm_out,
                     uint32_t* cell_start, uint32_t* workspace, hipStream_t s, StageTimer* timer,
                     const StepBuffers* move) {
    return sort_impl(p, keys_in, keys_out, perm_out, cell_start, workspace, s, timer, nullptr, 0, 0, move);
}

int launch_hash_keys(const float* pos, uint32_t stride, uint32_t n, float cell, uint32_t table_size,
                     uint32_t* keys, hipStream_t s, uint32_t key_base) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_hash_keys, dim3(grid_for(n)), dim3(256), 0, s, pos, stride, n, cell,
                       make_fastmod(table_size), key_base, keys);
    return check_launch("hash_keys");
}

int launch_key_starts(const uint32_t* sorted_keys, uint32_t n, uint32_t* table, hipStream_t s) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_key_starts, dim3(grid_for(n)), dim3(256), 0, s, sorted_keys, n, table);
    return check_launch("key_starts");
}

int launch_fill_value(uint32_t* a, size_t count, uint32_t value, hipStream_t s) {
    if (count == 0) return 0;
    hipLaunchKernelGGL(k_prepare, dim3(std::min<uint32_t>(grid_for(count), 2048u)), dim3(256), 0, s, nullptr,
                       (size_t)0, a, count, value);
    return check_launch("fill_value");
}

int launch_cell_start_scan(const SortPlan& p, const uint32_t* table, uint32_t* cell_start, uint32_t* workspace,
                           hipStream_t s) {
    if (p.cs_tiles == 0) return 0;
    hipLaunchKernelGGL(k_cell_start_scan<kCellStartItems>, dim3(p.cs_tiles), dim3(256), 0, s, table, p.n + 1,
                       cell_start, workspace + p.off_cs_status, workspace + p.off_cs_counter);
    return check_launch("cell_start_scan");
}

int launch_reorder_f32x3(const uint32_t* perm, const float* src, float* dst, uint32_t n, hipStream_t s) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_reorder_f32x3, dim3(grid_for(n)), dim3(256), 0, s, perm, src, dst, n);
    return check_launch("reorder_f32x3");
}

int launch_iota(uint32_t* a, uint32_t n, hipStream_t s) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_iota, dim3(grid_for(n)), dim3(256), 0, s, a, n);
    return check_launch("iota");
}

int launch_copy_words(const uint32_t* src, uint32_t* dst, uint32_t n, hipStream_t s) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_copy_words, dim3(grid_for(n)), dim3(256), 0, s, src, dst, n);
    return check_launch("copy_words");
}

size_t scan_workspace_words(uint32_t n) {
    const uint32_t tile = kThreads * 8;
    const size_t tiles = (n + tile - 1) / tile;
    return 64 + 2 * tiles;
}

int launch_scan_exclusive(uint32_t* data, uint32_t n, uint32_t* workspace, hipStream_t s) {
    if (n == 0) return 0;
    const uint32_t tile = kThreads * 8;
    const uint32_t tiles = (n + tile - 1) / tile;
    const size_t words = scan_workspace_words(n);
    hipLaunchKernelGGL(k_zero_words, dim3(std::min<uint32_t>(grid_for(words), 1024u)), dim3(256), 0, s, workspace,
                       words);
    VFS_TRY(check_launch("scan_prepare"));
    uint32_t* counter = workspace;
    auto* status = reinterpret_cast<unsigned long long*>(workspace + 64);
    hipLaunchKernelGGL(k_scan_lookback<8>, dim3(tiles), dim3(256), 0, s, data, n, status, counter);
    return check_launch("scan_lookback");
}

}  // namespace vfs
