// WCSPH substep kernels for gfx950: reorder of the particle state, boundary
// volume + density + equation of state, and the fused pressure + viscosity +
// wall + symplectic-Euler update that also hashes the new positions for the
// next substep.  Plus the analytic-shape volume-map generator.
//
// Arithmetic follows the reference shaders (luihabl/VkFluidSim,
// assets/shaders/simulation/wcsph_with_boundary_model.slang and
// kernels/kernels_3d.slang) operation by operation; the library is compiled
// with -ffp-contract=off so that every float operation rounds once, as in the
// CPU oracle; a vector divided by a scalar, and r / h in the smoothing kernels,
// are multiplied by the IEEE reciprocal of the divisor, other scalar quotients
// are IEEE divisions (the oracle's canonical reading of the shader arithmetic).
// Neighbour enumeration visits the 27 hashed cells in the reference order (z
// outer, y, x inner; spatial_hash_3d.slang:3-10) and each bucket in ascending
// sorted index, so sums are accumulated in the reference order.  When the
// three x-adjacent cells of a row hash to consecutive keys their buckets are
// contiguous in sorted order and are walked as one range.
#include "volume_map.h"
#include "wcsph_math.h"

namespace vfs {

namespace {

// The cell walk appends in-range neighbours to a per-lane LDS queue and
// moves its complete chunks of four into the neighbour list (one 16-byte
// store each) once any active lane holds kQueue of them, so the list stores
// run with most lanes busy; the unstored remainder (up to three) moves to the
// front of the queue.  Queued neighbours keep their arrival order (the
// reference order).  Candidates are read kUnroll at
// a time (all loads issued before the tests); every candidate writes the next
// queue slot and only advances the count when it is in range, so the test has
// no branch.  A queue can therefore hold up to kQueue - 1 + kUnroll entries
// before the flush check.
constexpr int kUnroll = (int)kCandidateGroup;
constexpr int kQueue = 8;
constexpr int kQueueCap = kQueue + kUnroll;
static_assert(kQueueCap % 4 == 0, "the flush reads the three slots after the last complete chunk");

// Row (dz, dy) of the 27-cell stencil: the buckets of cells (cx-1..cx+1, cy+dy,
// cz+dz) as up to three [begin, end) ranges of sorted indices, merged into one
// when their keys are consecutive (then the buckets are adjacent).
struct RowRanges {
    uint32_t b0, e0, b1, e1, b2, e2;
    int nr;
};

__device__ __forceinline__ RowRanges row_ranges(int32_t cx, int32_t cy, int32_t cz, int dy, int dz,
                                                const SimConsts& c, const uint32_t* __restrict__ cs) {
    const uint32_t k0 = cell_key(cx, cy, cz, -1, dy, dz, c);
    const uint32_t k1 = cell_key(cx, cy, cz, 0, dy, dz, c);
    const uint32_t k2 = cell_key(cx, cy, cz, 1, dy, dz, c);
    RowRanges r;
    if (k1 == k0 + 1u && k2 == k1 + 1u) {
        r.b0 = cs[k0];
        r.e0 = cs[k2 + 1u];
        r.b1 = r.e1 = r.b2 = r.e2 = 0u;
        r.nr = 1;
    } else {
        r.b0 = cs[k0];
        r.e0 = cs[k0 + 1u];
        r.b1 = cs[k1];
        r.e1 = cs[k1 + 1u];
        r.b2 = cs[k2];
        r.e2 = cs[k2 + 1u];
        r.nr = 3;
    }
    return r;
}

// Optional wave-level LDS staging of a stencil row.  The 64 particles of a
// wave lie in a few consecutive cells of one x-row of cells, so their ranges
// for row (dy, dz) fall in one short span of sorted indices: the wave copies
// the span to its LDS slice with coalesced loads and the lanes read their
// candidates from there.  Rows whose span exceeds the slice read global
// memory.  No block barrier: the copy and the reads are ordered within the
// wave (LDS executes one wave's instructions in issue order).
constexpr uint32_t kWaveStage = 192;  // three coalesced loads per lane

// Stages row (dy, dz) for the wave when the lanes' ranges fit one slice: the
// span runs from the first lane's lowest begin to the last lane's highest end
// (sorted keys ascend over the lanes, so this is the common case) and a wave
// vote confirms that every lane's ranges lie inside.  Returns the wave-uniform
// decision; L is the sorted index of slice entry 0.
__device__ __forceinline__ bool stage_row(float4* __restrict__ st, const RowRanges& r, bool active,
                                          const float4* __restrict__ xs, uint32_t& L) {
    uint32_t lo = 0xFFFFFFFFu, hi = 0u;
    auto span = [&](uint32_t jb, uint32_t je) {
        if (je > jb) {
            lo = min(lo, jb);
            hi = max(hi, je);
        }
    };
    span(r.b0, r.e0);
    if (r.nr == 3) {
        span(r.b1, r.e1);
        span(r.b2, r.e2);
    }
    const uint32_t wl = (uint32_t)__shfl(lo, 0, 64);
    const uint32_t wh = (uint32_t)__shfl(hi, 63, 64);
    const bool fits = !active || (lo >= wl && hi <= wh);
    const bool staged = __all(fits) && wh > wl && wh - wl <= kWaveStage;
    if (staged) {
        // every lane's reads of the previous row precede these writes
        __builtin_amdgcn_wave_barrier();
        const uint32_t lane = threadIdx.x & 63u, n = wh - wl;
        float4 a0, a1, a2;
        if (lane < n) a0 = xs[wl + lane];
        if (lane + 64u < n) a1 = xs[wl + lane + 64u];
        if (lane + 128u < n) a2 = xs[wl + lane + 128u];
        if (lane < n) st[lane] = a0;
        if (lane + 64u < n) st[lane + 64u] = a1;
        if (lane + 128u < n) st[lane + 128u] = a2;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    L = wl;
    return staged;
}

// Visits candidates jb..je-1 in order with kUnroll loads in flight per lane:
// whole groups first (every candidate valid), then one masked tail group whose
// extra loads read neighbouring particles (or the padding past the array) and
// whose validity enters the count as a bit, without a branch per candidate.
// hit(j, q, valid) appends j to the queue when valid and in range.
template <class Load, class Hit, class Flush>
__device__ __forceinline__ void walk_range(uint32_t jb, uint32_t je, Load&& load, const int& cnt, Hit&& hit,
                                           Flush&& flush) {
    uint32_t j0 = jb;
    for (; j0 + (uint32_t)kUnroll <= je; j0 += (uint32_t)kUnroll) {
        float4 q[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) q[u] = load(j0, u);
#pragma unroll
        for (int u = 0; u < kUnroll; u++) hit(j0 + (uint32_t)u, q[u], true);
        if (__any(cnt >= kQueue)) flush();
    }
    if (j0 < je) {
        float4 q[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) q[u] = load(j0, u);
#pragma unroll
        for (int u = 0; u < kUnroll; u++) hit(j0 + (uint32_t)u, q[u], j0 + (uint32_t)u < je);
    }
    if (__any(cnt >= kQueue)) flush();
}

// Reorder (reorder.slang:11-19) of x, v (and ids) by the sort permutation, and
// CalculateOffsets (spatial_offsets.slang:19-31) into the offsets table when
// one is given.
__global__ __launch_bounds__(256) void k_reorder_state(StepBuffers b, uint32_t nblocks) {
    const uint32_t i = xcd_block(nblocks) * 256u + threadIdx.x;
    if (i >= b.n) return;
    const uint32_t s = b.perm[i];
    b.xs[i] = b.x[s];
    b.vs[i] = b.v[s];
    if (b.pid) b.pids[i] = b.pid[s];
    if (b.ps) b.ps[i] = b.pred[s];
    if (!b.offsets) return;
    const uint32_t k = b.sorted_keys[i];
    if (i == 0 || b.sorted_keys[i - 1] != k) b.offsets[k] = i;
}

// The 27-cell stencil of the cell of pos as nine rows in the reference
// order (z outer, y inner); inactive lanes read no cell starts and get empty
// rows.
template <class Visit>
__device__ __forceinline__ void stencil_rows(const float* pos, bool active, const SimConsts& c,
                                             const uint32_t* __restrict__ cs, Visit&& visit) {
    const int32_t cx = floor_to_cell(pos[0] / c.cell);
    const int32_t cy = floor_to_cell(pos[1] / c.cell);
    const int32_t cz = floor_to_cell(pos[2] / c.cell);
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++) {
            const RowRanges r = active ? row_ranges(cx, cy, cz, dy, dz, c, cs) : RowRanges{0u, 0u, 0u, 0u, 0u, 0u, 1};
            visit(r);
        }
}

// Shared memory of the list builder: the per-lane neighbour queues and the
// wave slices of staged rows (padded for the masked tail group's reads).
template <int BLOCK, bool STAGED>
struct ListShared {
    uint32_t queue[kQueueCap][BLOCK];
    float4 stage[STAGED ? BLOCK / 64 : 1][kWaveStage + kUnroll];
};

// This substep's neighbour list of particle i (position pos) over the sorted
// positions pts: the cell walk finds the neighbours j with r^2 <= h2 (self
// included, once per visit of its bucket) in reference order and stores them
// in chunks of four, chunk-major (chunk g of particle i at g * n + i), so that
// list reads and writes are coalesced and each store fills 16 bytes.  Returns
// the count; a list longer than nbr_cap stores its count only.  Every lane of
// the wave calls it (wave votes inside); inactive lanes list nothing.
template <int BLOCK, bool STAGED>
__device__ __forceinline__ uint32_t build_list(ListShared<BLOCK, STAGED>& sh, const float4* __restrict__ pts,
                                               const float* pos, float h2, bool active, uint32_t i,
                                               const SimConsts& c, const uint32_t* __restrict__ cs,
                                               uint4* __restrict__ list, uint32_t n) {
    const int tid = threadIdx.x;
    auto& s_j = sh.queue;
    int cnt = 0;          // queued neighbours, the unstored part of the last chunk first
    uint32_t chunks = 0;  // complete chunks of four moved out of the queue
    auto hit = [&](uint32_t j, const float4& q, bool valid) {
        const float ox = q.x - pos[0], oy = q.y - pos[1], oz = q.z - pos[2];
        const float r2 = dot3(ox, oy, oz, ox, oy, oz);
        s_j[cnt][tid] = j;
        cnt += (int)((uint32_t)valid & (uint32_t)!(r2 > h2));  // no branch on valid
    };
    // complete chunks into the list in order; the remainder's three slots move
    // to the front whole (slots past the remainder are overwritten before use)
    auto flush = [&]() {
        const int full = cnt >> 2;
        for (int g = 0; g < full; g++) {
            const uint4 v = make_uint4(s_j[4 * g][tid], s_j[4 * g + 1][tid], s_j[4 * g + 2][tid], s_j[4 * g + 3][tid]);
            list[(size_t)min(chunks, c.nbr_cap >> 2) * n + i] = v;  // past the capacity: the spare row
            chunks++;
        }
        const int rem = cnt & 3;
        if (full > 0) {
            const uint32_t r0 = s_j[4 * full][tid], r1 = s_j[4 * full + 1][tid], r2 = s_j[4 * full + 2][tid];
            s_j[0][tid] = r0;
            s_j[1][tid] = r1;
            s_j[2][tid] = r2;
        }
        cnt = rem;
    };
    stencil_rows(pos, active, c, cs, [&](const RowRanges& r) {
        uint32_t L = 0;
        bool staged = false;
        if (STAGED) staged = stage_row(sh.stage[STAGED ? tid >> 6 : 0], r, active, pts, L);
        if (active) {
            auto visit = [&](uint32_t jb, uint32_t je) {
                if (STAGED && staged)
                    walk_range(
                        jb, je, [&](uint32_t j0, int u) { return sh.stage[STAGED ? tid >> 6 : 0][j0 - L + u]; },
                        cnt, hit, flush);
                else
                    walk_range(jb, je, [&](uint32_t j0, int u) { return (pts + j0)[u]; }, cnt, hit, flush);
            };
            visit(r.b0, r.e0);
            if (r.nr == 3) {
                visit(r.b1, r.e1);
                visit(r.b2, r.e2);
            }
        }
    });
    if (!active) return 0;
    flush();
    // the last, partial chunk; its unused entries hold the particle's own index
    const uint32_t listed = 4u * chunks + (uint32_t)cnt;
    if (cnt > 0 && listed <= c.nbr_cap) {
        const uint32_t a0 = s_j[0][tid];
        const uint32_t a1 = cnt > 1 ? s_j[1][tid] : i;
        const uint32_t a2 = cnt > 2 ? s_j[2][tid] : i;
        list[(size_t)chunks * n + i] = make_uint4(a0, a1, a2, i);
    }
    return listed;
}

// Calls add(q) for the positions q (from r) of the first `listed` entries of
// the list `own` (particle i's column) in order, every lane busy until its own
// count.  One chunk's positions are in flight while the previous chunk is
// summed; chunk reads past the last one are clamped to it, entries past
// `listed` in the last chunk are loaded and ignored.  The lane's own list may
// just have been written (same-address accesses of a wave are ordered).
template <class Add>
__device__ __forceinline__ void sum_list(const uint4* __restrict__ own, uint32_t listed, uint32_t n,
                                         __amdgpu_buffer_rsrc_t r, Add&& add) {
    const uint32_t groups = (listed + 3u) >> 2;
    const uint32_t last = groups ? groups - 1u : 0u;
    uint4 en = own[(size_t)min(1u, last) * n];
    float4 q0, q1, q2, q3;
    {
        const uint4 e = own[0];
        q0 = load_f4(r, e.x), q1 = load_f4(r, e.y), q2 = load_f4(r, e.z), q3 = load_f4(r, e.w);
    }
    for (uint32_t gi = 0; gi < groups; gi++) {
        const float4 p0 = q0, p1 = q1, p2 = q2, p3 = q3;
        q0 = load_f4(r, en.x), q1 = load_f4(r, en.y), q2 = load_f4(r, en.z), q3 = load_f4(r, en.w);
        en = own[(size_t)min(gi + 2u, last) * n];
        const uint32_t t = 4u * gi;
        add(p0);
        if (t + 1u < listed) add(p1);
        if (t + 2u < listed) add(p2);
        if (t + 3u < listed) add(p3);
    }
}

// Boundary volume, CalculateDensities (slang:153-197), the per-particle EOS
// terms the force pass reads for every neighbour, and this substep's neighbour
// list (build_list).  The density sum then walks the particle's own list.  A
// particle with more than nbr_cap neighbours stores its count only; its
// density, and its forces in the force pass, walk the cells again.
template <int BLOCK, bool STAGED, bool BATCH>
__global__ __launch_bounds__(BLOCK) void k_density(StepBuffers b, SimConsts c0, SceneTable table,
                                                   const DevObject* __restrict__ objs, uint32_t nblocks) {
    __shared__ ListShared<BLOCK, STAGED> sh;
    const uint32_t i = xcd_block(nblocks) * (uint32_t)BLOCK + threadIdx.x;
    const SimConsts c = scene_consts<BATCH>(c0, table, i);
    const bool active = i - c.base < c.n;
    const float4 p4 = b.xs[active ? i : 0u];
    const float pos[3] = {p4.x, p4.y, p4.z};
    float psi = 0.0f, g[3] = {0.0f, 0.0f, 0.0f};
    if (active) boundary_volume(objs + c.obj_begin, c.obj_count, pos, c, psi, g);

    uint4* __restrict__ list = reinterpret_cast<uint4*>(b.nbr);
    const uint32_t listed = build_list<BLOCK, STAGED>(sh, b.xs, pos, c.h2, active, i, c, b.cell_start, list, b.n);
    if (!active) return;
    b.nbr_count[i] = listed;

    float density = 0.0f;
    auto add = [&](const float4& q) {
        const float ox = q.x - pos[0], oy = q.y - pos[1], oz = q.z - pos[2];
        density += c.mass * kernel_w(sqrtf(dot3(ox, oy, oz, ox, oy, oz)), c);
    };
    if (listed <= c.nbr_cap) {
        sum_list(list + i, listed, b.n, make_rsrc(b.xs, b.n * 16u), add);
    } else {
        const float4* __restrict__ xs = b.xs;
        stencil_rows(pos, true, c, b.cell_start, [&](const RowRanges& r) {
            auto visit = [&](uint32_t jb, uint32_t je) {
                for (uint32_t j = jb; j < je; j++) {
                    const float4 q = xs[j];
                    const float ox = q.x - pos[0], oy = q.y - pos[1], oz = q.z - pos[2];
                    if (!(dot3(ox, oy, oz, ox, oy, oz) > c.h2)) add(q);
                }
            };
            visit(r.b0, r.e0);
            if (r.nr == 3) {
                visit(r.b1, r.e1);
                visit(r.b2, r.e2);
            }
        });
    }
    density += c.rho0 * psi;
    b.rho[i] = density;
    const float p = eos(density, c);
    const float pd = density > 0.0f ? p / (density * density) : 0.0f;
    const float mrho = c.mass / fmaxf(density, 1e-6f);
    // the force pass reads a neighbour's EOS terms with its position and velocity
    reinterpret_cast<float*>(b.xs)[4 * (size_t)i + 3] = pd;
    reinterpret_cast<float*>(b.vs)[4 * (size_t)i + 3] = mrho;
    b.bnd[i] = make_float4(g[0], g[1], g[2], psi);
}

struct ForceAcc {
    float pfx, pfy, pfz, vfx, vfy, vfz;
};

// Pressure (slang:205-262) and viscosity (slang:273-324) terms of one pair:
// neighbour position and p/rho^2 in q, velocity and mass/max(rho, 1e-6) in vj.
__device__ __forceinline__ void force_pair(ForceAcc& f, const float4& p4, const float4& v4, float pdi, const float4& q,
                                           const float4& vj, const SimConsts& c) {
    const float xx = p4.x - q.x, xy = p4.y - q.y, xz = p4.z - q.z;
    const float r2 = dot3(xx, xy, xz, xx, xy, xz);
    const float r = sqrtf(r2);
    float nx, ny, nz;
    if (r > 0.0f) {
        const float inv_r = rcp_rn(r);  // r in (0, h]: IEEE 1 / r
        nx = xx * inv_r;
        ny = xy * inv_r;
        nz = xz * inv_r;
    } else {
        nx = 0.0f;
        ny = 1.0f;
        nz = 0.0f;
    }
    const float G = kernel_dw(r, c);
    const float s = (c.mass * (pdi + q.w)) * G;
    f.pfx += s * nx;
    f.pfy += s * ny;
    f.pfz += s * nz;
    const float ux = v4.x - vj.x, uy = v4.y - vj.y, uz = v4.z - vj.z;
    const float sv = vj.w * dot3(ux, uy, uz, xx, xy, xz);
    const float inv_den = rcp_rn(r2 + c.visc_eps_h2);  // in [0.01 h^2, 1.01 h^2]: IEEE 1 / (.)
    f.vfx += (sv * (nx * G)) * inv_den;
    f.vfy += (sv * (ny * G)) * inv_den;
    f.vfz += (sv * (nz * G)) * inv_den;
}

// CalculatePressureAccel (slang:205-271) + CalculateViscousAccel (slang:273-334)
// over this substep's neighbour list (one walk, separate accumulators), then
// WallAcceleration and UpdatePositions (slang:336-360), then UpdateSpatialHash
// of the new position.  Particles whose list overflowed walk their cells.
// DEEP: each list chunk's indices are loaded two chunks ahead of its
// neighbours' gathers, so that no iteration waits for an index load and then
// for the gathers it feeds.  It costs registers, so it is launched where the
// particles fill about a wave per SIMD and latency, not occupancy, bounds the
// pass (the same pairs in the same order: results are identical).
template <int BLOCK, bool BATCH, bool DEEP = false>
__global__ __launch_bounds__(BLOCK) void k_forces_integrate(StepBuffers b, SimConsts c0, SceneTable table,
                                                            uint32_t nblocks) {
    const uint32_t i = xcd_block(nblocks) * (uint32_t)BLOCK + threadIdx.x;
    const SimConsts c = scene_consts<BATCH>(c0, table, i);
    if (i - c.base >= c.n) return;
    const float4 p4 = b.xs[i];
    const float4 v4 = b.vs[i];
    const float di = b.rho[i];
    const float pdi = p4.w;
    const float pi = eos(di, c);
    const auto xr = make_rsrc(b.xs, b.n * 16u);
    const auto vr = make_rsrc(b.vs, b.n * 16u);
    ForceAcc f{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};

    const uint32_t count = b.nbr_count[i];
    if (count <= c.nbr_cap) {
        // chunks of four neighbours; entries past count in the last chunk hold
        // the particle's own index, loaded and ignored; self entries are skipped
        // as in the reference loops.  The next chunk's neighbours are loaded
        // while the current chunk's pairs are evaluated (reads past the last
        // chunk are clamped to it).
        const uint4* __restrict__ list = reinterpret_cast<const uint4*>(b.nbr) + i;
        const uint32_t groups = (count + 3u) >> 2;
        uint4 e = groups ? list[0] : make_uint4(0u, 0u, 0u, 0u);
        uint4 e1 = make_uint4(0u, 0u, 0u, 0u);
        if (DEEP && groups) e1 = list[(size_t)min(1u, groups - 1u) * b.n];
        float4 q0 = load_f4(xr, e.x), w0 = load_f4(vr, e.x);
        float4 q1 = load_f4(xr, e.y), w1 = load_f4(vr, e.y);
        float4 q2 = load_f4(xr, e.z), w2 = load_f4(vr, e.z);
        float4 q3 = load_f4(xr, e.w), w3 = load_f4(vr, e.w);
        for (uint32_t g = 0; g < groups; g++) {
            uint4 en;
            if (DEEP) {
                en = e1;
                e1 = list[(size_t)min(g + 2u, groups - 1u) * b.n];
            } else {
                en = list[(size_t)min(g + 1u, groups - 1u) * b.n];
            }
            const float4 n0 = load_f4(xr, en.x), m0 = load_f4(vr, en.x);
            const float4 n1 = load_f4(xr, en.y), m1 = load_f4(vr, en.y);
            const float4 n2 = load_f4(xr, en.z), m2 = load_f4(vr, en.z);
            const float4 n3 = load_f4(xr, en.w), m3 = load_f4(vr, en.w);
            const uint32_t t = 4u * g;
            if (e.x != i) force_pair(f, p4, v4, pdi, q0, w0, c);
            if (t + 1u < count && e.y != i) force_pair(f, p4, v4, pdi, q1, w1, c);
            if (t + 2u < count && e.z != i) force_pair(f, p4, v4, pdi, q2, w2, c);
            if (t + 3u < count && e.w != i) force_pair(f, p4, v4, pdi, q3, w3, c);
            e = en;
            q0 = n0, q1 = n1, q2 = n2, q3 = n3;
            w0 = m0, w1 = m1, w2 = m2, w3 = m3;
        }
    } else {
        // overflow: the 27 cells in reference order, pairs evaluated as found
        const int32_t cx = floor_to_cell(p4.x / c.cell);
        const int32_t cy = floor_to_cell(p4.y / c.cell);
        const int32_t cz = floor_to_cell(p4.z / c.cell);
        auto visit = [&](uint32_t jb, uint32_t je) {
            for (uint32_t j = jb; j < je; j++) {
                if (j == i) continue;
                const float4 q = load_f4(xr, j);
                const float xx = p4.x - q.x, xy = p4.y - q.y, xz = p4.z - q.z;
                const float r2 = dot3(xx, xy, xz, xx, xy, xz);
                if (r2 > c.h2) continue;
                force_pair(f, p4, v4, pdi, q, load_f4(vr, j), c);
            }
        };
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++) {
                const RowRanges r = row_ranges(cx, cy, cz, dy, dz, c, b.cell_start);
                visit(r.b0, r.e0);
                if (r.nr == 3) {
                    visit(r.b1, r.e1);
                    visit(r.b2, r.e2);
                }
            }
    }

    const float pdb = di > 0.0f ? pi / c.rho0_sq : 0.0f;
    const float sb = (pdi + pdb) * c.rho0;
    const float4 bg = b.bnd[i];
    f.pfx += sb * bg.x;
    f.pfy += sb * bg.y;
    f.pfz += sb * bg.z;

    float ax = c.gx + (-f.pfx), ay = c.gy + (-f.pfy), az = c.gz + (-f.pfz);
    ax += c.nu10 * f.vfx;
    ay += c.nu10 * f.vfy;
    az += c.nu10 * f.vfz;
    // WallAcceleration
    const float dfx = fminf(p4.x - c.bb_lo[0], 0.0f) - fminf(c.bb_hi[0] - p4.x, 0.0f);
    const float dfy = fminf(p4.y - c.bb_lo[1], 0.0f) - fminf(c.bb_hi[1] - p4.y, 0.0f);
    const float dfz = fminf(p4.z - c.bb_lo[2], 0.0f) - fminf(c.bb_hi[2] - p4.z, 0.0f);
    const bool wall = fabsf(dfx) > 1e-9f || fabsf(dfy) > 1e-9f || fabsf(dfz) > 1e-9f;
    ax += wall ? c.kwall_neg * dfx : 0.0f;
    ay += wall ? c.kwall_neg * dfy : 0.0f;
    az += wall ? c.kwall_neg * dfz : 0.0f;
    // symplectic Euler
    const float nvx = v4.x + ax * c.dt, nvy = v4.y + ay * c.dt, nvz = v4.z + az * c.dt;
    const float nxx = p4.x + nvx * c.dt, nxy = p4.y + nvy * c.dt, nxz = p4.z + nvz * c.dt;
    b.x[i] = make_float4(nxx, nxy, nxz, 0.0f);
    b.v[i] = make_float4(nvx, nvy, nvz, 0.0f);
    if (b.accel) b.accel[i] = make_float4(ax, ay, az, 0.0f);
    if (b.pid) b.pid[i] = b.pids[i];
    b.keys[i] = c.base + fastmod(hash_cell(floor_to_cell(nxx / c.cell), floor_to_cell(nxy / c.cell),
                                           floor_to_cell(nxz / c.cell)),
                                 c.fm);
}

// ---------------------------------------------------------- Lague model
// SpikyPow2 / SpikyPow3 / their gradients / Poly6 (kernels_3d.slang:16-60).
__device__ __forceinline__ float spiky2(float d, const LagueConsts& l) {
    const float v = l.h - d;
    return d < l.h ? (v * v) * l.sp2 : 0.0f;
}
__device__ __forceinline__ float spiky3(float d, const LagueConsts& l) {
    const float v = l.h - d;
    return d < l.h ? ((v * v) * v) * l.sp3 : 0.0f;
}
__device__ __forceinline__ float grad_spiky2(float d, const LagueConsts& l) {
    const float v = l.h - d;
    return d <= l.h ? (-v) * l.dsp2 : 0.0f;
}
__device__ __forceinline__ float grad_spiky3(float d, const LagueConsts& l) {
    const float v = l.h - d;
    return d <= l.h ? ((-v) * v) * l.dsp3 : 0.0f;
}
__device__ __forceinline__ float poly6(float d, const LagueConsts& l) {
    const float w = l.h * l.h - d * d;
    return d < l.h ? ((w * w) * w) * l.poly6 : 0.0f;
}

// ExternalForces (lague_model.slang:211-221) and the hash key of the
// prediction, which the spatial hash of this substep sorts (model.cpp:210-222).
__global__ __launch_bounds__(256) void k_lague_predict(StepBuffers b, SimConsts c, LagueConsts l, uint32_t nblocks) {
    const uint32_t i = xcd_block(nblocks) * 256u + threadIdx.x;
    if (i >= b.n) return;
    const float4 x = b.x[i];
    const float4 v0 = b.v[i];
    const float vx = v0.x + l.gx * l.dt, vy = v0.y + l.gy * l.dt, vz = v0.z + l.gz * l.dt;
    const float px = x.x + vx * l.pf, py = x.y + vy * l.pf, pz = x.z + vz * l.pf;
    b.v[i] = make_float4(vx, vy, vz, 0.0f);
    b.pred[i] = make_float4(px, py, pz, 0.0f);
    b.keys[i] = c.base + fastmod(hash_cell(floor_to_cell(px / c.cell), floor_to_cell(py / c.cell),
                                           floor_to_cell(pz / c.cell)),
                                 c.fm);
}

// Walks the 27-cell stencil of the sorted predictions in the reference's
// order, calling hit(j, q, valid) for every candidate.
template <class Hit>
__device__ __forceinline__ void lague_walk(const float4& p, bool active, const StepBuffers& b, const SimConsts& c,
                                           Hit&& hit) {
    const int32_t cx = floor_to_cell(p.x / c.cell), cy = floor_to_cell(p.y / c.cell),
                  cz = floor_to_cell(p.z / c.cell);
    const float4* __restrict__ ps = b.ps;
    const int cnt = 0;
    auto none = [] {};
    auto load = [&](uint32_t j0, int u) { return (ps + j0)[u]; };
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++) {
            if (!active) continue;
            const RowRanges r = row_ranges(cx, cy, cz, dy, dz, c, b.cell_start);
            walk_range(r.b0, r.e0, load, cnt, hit, none);
            if (r.nr == 3) {
                walk_range(r.b1, r.e1, load, cnt, hit, none);
                walk_range(r.b2, r.e2, load, cnt, hit, none);
            }
        }
}

// CalculateDensities (lague_model.slang:60-97, 222-230): density and near
// density of the predicted positions, self included, summed over this
// substep's neighbour list of the predictions (build_list), which the force
// pass walks again.  The force pass reads a neighbour's pressure with its
// prediction (ps.w) and its near pressure and density reciprocals from b.bnd
// (unused by this model): the same values the reference derives per pair.
template <bool STAGED>
__global__ __launch_bounds__(128) void k_lague_density(StepBuffers b, SimConsts c, LagueConsts l, uint32_t nblocks) {
    __shared__ ListShared<128, STAGED> sh;
    const uint32_t i = xcd_block(nblocks) * 128u + threadIdx.x;
    const bool active = i < b.n;
    const float4 p = b.ps[active ? i : 0u];
    const float pos[3] = {p.x, p.y, p.z};
    uint4* __restrict__ list = reinterpret_cast<uint4*>(b.nbr);
    const uint32_t listed = build_list<128, STAGED>(sh, b.ps, pos, l.h2, active, i, c, b.cell_start, list, b.n);
    if (!active) return;
    b.nbr_count[i] = listed;
    float density = 0.0f, near = 0.0f;
    auto add = [&](const float4& q) {
        const float ox = q.x - p.x, oy = q.y - p.y, oz = q.z - p.z;
        const float d = sqrtf(dot3(ox, oy, oz, ox, oy, oz));
        density += spiky2(d, l);
        near += spiky3(d, l);
    };
    if (listed <= c.nbr_cap) {
        sum_list(list + i, listed, b.n, make_rsrc(b.ps, b.n * 16u), add);
    } else {
        lague_walk(p, true, b, c, [&](uint32_t, const float4& q, bool valid) {
            const float ox = q.x - p.x, oy = q.y - p.y, oz = q.z - p.z;
            if (valid && !(dot3(ox, oy, oz, ox, oy, oz) > l.h2)) add(q);
        });
    }
    b.rho[i] = density;
    b.near_rho[i] = near;
    reinterpret_cast<float*>(b.ps)[4 * (size_t)i + 3] = (density - l.rho0) * l.pm;
    b.bnd[i] = make_float4(l.npm * near, 1.0f / density, 1.0f / near, 0.0f);
}

// CalculatePressureForces (lague_model.slang:99-160, 231-238) with the
// velocities from before this pass, over the density pass's neighbour list
// (self skipped; a list that overflowed walks the cells again), then
// UpdatePositions with ResolveCollisions (slang:162-200, 240-250).
__global__ __launch_bounds__(128) void k_lague_forces(StepBuffers b, SimConsts c, LagueConsts l, uint32_t nblocks) {
    const uint32_t i = xcd_block(nblocks) * 128u + threadIdx.x;
    if (i >= b.n) return;
    const float4 p = b.ps[i];
    const float4 vel = b.vs[i];
    const float4 own = b.bnd[i];
    const float pressure = p.w;
    const float near_pressure = own.x;
    float pfx = 0.0f, pfy = 0.0f, pfz = 0.0f, vfx = 0.0f, vfy = 0.0f, vfz = 0.0f;
    // one in-range neighbour j != i: prediction q, near pressure and density
    // reciprocals nb, velocity vj
    auto pair = [&](const float4& q, const float4& nb, const float4& vj) {
        const float ox = q.x - p.x, oy = q.y - p.y, oz = q.z - p.z;
        const float r2 = dot3(ox, oy, oz, ox, oy, oz);
        const float shared = (pressure + q.w) * 0.5f;
        const float shared_near = (near_pressure + nb.x) * 0.5f;
        const float d = sqrtf(r2);
        float dx = 0.0f, dy = 1.0f, dz = 0.0f;
        if (d > 0.0f) {
            const float inv = 1.0f / d;
            dx = ox * inv;
            dy = oy * inv;
            dz = oz * inv;
        }
        const float dd = grad_spiky2(d, l), ndd = grad_spiky3(d, l);
        const float inv_nd = nb.y, inv_nnd = nb.z;
        pfx += ((dx * dd) * shared) * inv_nd;
        pfy += ((dy * dd) * shared) * inv_nd;
        pfz += ((dz * dd) * shared) * inv_nd;
        pfx += ((dx * ndd) * shared_near) * inv_nnd;
        pfy += ((dy * ndd) * shared_near) * inv_nnd;
        pfz += ((dz * ndd) * shared_near) * inv_nnd;
        const float w = poly6(d, l);
        vfx += (vj.x - vel.x) * w;
        vfy += (vj.y - vel.y) * w;
        vfz += (vj.z - vel.z) * w;
    };
    const uint32_t count = b.nbr_count[i];
    if (count <= c.nbr_cap) {
        const auto pr = make_rsrc(b.ps, b.n * 16u), br = make_rsrc(b.bnd, b.n * 16u), vr = make_rsrc(b.vs, b.n * 16u);
        const uint4* __restrict__ list = reinterpret_cast<const uint4*>(b.nbr) + i;
        const uint32_t groups = (count + 3u) >> 2;
        for (uint32_t g = 0; g < groups; g++) {
            const uint4 e = list[(size_t)g * b.n];
            const uint32_t js[4] = {e.x, e.y, e.z, e.w};
            float4 q[4], nb[4], vj[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                q[k] = load_f4(pr, js[k]);
                nb[k] = load_f4(br, js[k]);
                vj[k] = load_f4(vr, js[k]);
            }
            const uint32_t t = 4u * g;
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (t + (uint32_t)k < count && js[k] != i) pair(q[k], nb[k], vj[k]);
        }
    } else {
        lague_walk(p, true, b, c, [&](uint32_t j, const float4& q, bool valid) {
            const float ox = q.x - p.x, oy = q.y - p.y, oz = q.z - p.z;
            const float r2 = dot3(ox, oy, oz, ox, oy, oz);
            if (!valid || j == i || r2 > l.h2) return;
            pair(q, b.bnd[j], b.vs[j]);
        });
    }
    const float inv_density = own.y;
    const float ax = pfx * inv_density + vfx * l.visc;
    const float ay = pfy * inv_density + vfy * l.visc;
    const float az = pfz * inv_density + vfz * l.visc;
    float v3[3] = {vel.x + ax * l.dt, vel.y + ay * l.dt, vel.z + az * l.dt};
    const float4 x4 = b.xs[i];
    float x3[3] = {x4.x + v3[0] * l.dt, x4.y + v3[1] * l.dt, x4.z + v3[2] * l.dt};
#pragma unroll
    for (int a = 0; a < 3; a++)
        if (x3[a] < l.bb_lo[a]) {
            x3[a] = l.bb_lo[a];
            v3[a] = (-v3[a]) * l.damp;
        }
#pragma unroll
    for (int a = 0; a < 3; a++)
        if (x3[a] > l.bb_hi[a]) {
            x3[a] = l.bb_hi[a];
            v3[a] = (-v3[a]) * l.damp;
        }
    b.x[i] = make_float4(x3[0], x3[1], x3[2], 0.0f);
    b.v[i] = make_float4(v3[0], v3[1], v3[2], 0.0f);
    if (b.pid) b.pid[i] = b.pids[i];
}

// ---------------------------------------------------------- volume maps
struct ShapeArgs {
    int kind;
    double p[6];
};

__device__ double shape_sdf(const ShapeArgs& s, float qx, float qy, float qz) {
    const double px = qx, py = qy, pz = qz;
    if (s.kind == VFS_SHAPE_SPHERE) {
        const double dx = px - s.p[0], dy = py - s.p[1], dz = pz - s.p[2];
        return sqrt(dx * dx + dy * dy + dz * dz) - s.p[3];
    }
    const double ax = fabs(px - s.p[0]) - s.p[3], ay = fabs(py - s.p[1]) - s.p[4], az = fabs(pz - s.p[2]) - s.p[5];
    const double ox = fmax(ax, 0.0), oy = fmax(ay, 0.0), oz = fmax(az, 0.0);
    const double outside = sqrt(ox * ox + oy * oy + oz * oz);
    const double inside = fmin(fmax(ax, fmax(ay, az)), 0.0);
    const double d = outside + inside;
    return s.kind == VFS_SHAPE_BOX_WALLS ? -d : d;
}

// One grid node per thread: SDF value (mesh_sdf.cpp:28-45) and the volume map
// integral of volume_map.cpp:9-54 for an analytic shape.
__global__ __launch_bounds__(128) void k_volume_map(ShapeArgs sh, uint32_t rx, uint32_t ry, uint32_t rz, float sr,
                                                    double tolerance, float d0, float d1, float d2, double s0,
                                                    double s1, double s2, GLTable gl, float* __restrict__ sdf_out,
                                                    float* __restrict__ vm_out) {
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t total = (uint64_t)rx * ry * rz;
    if (idx >= total) return;
    float x[3];
    grid_node(idx, rx, ry, d0, d1, d2, s0, s1, s2, x);
    const double dist = shape_sdf(sh, x[0], x[1], x[2]);
    sdf_out[idx] = (float)(dist - tolerance);
    auto sdf = [&](float qx, float qy, float qz) { return shape_sdf(sh, qx, qy, qz); };
    vm_out[idx] = (float)volume_map_integral(sdf, x, dist, sr, gl);
}

inline uint32_t grid_for(uint64_t n, uint32_t block) { return (uint32_t)((n + block - 1) / block); }

// Particle kernels: blocks padded to a multiple of 8 (xcd_block).
inline uint32_t particle_blocks(uint32_t n, uint32_t block) { return (grid_for(n, block) + 7u) & ~7u; }

}  // namespace

// State check: particles with a non-finite position or velocity component
// (counters[0]) and particles farther than `margin` outside [lo, hi] on any
// axis (counters[1]); one atomic per wave.  The first kernel clears the
// counters (a kernel, not a memset, so the pair can sit in a captured graph).
__global__ void k_check_clear(uint32_t* counters) {
    if (threadIdx.x < 2) counters[threadIdx.x] = 0u;
}

__global__ __launch_bounds__(256) void k_check_state(const float4* __restrict__ x, const float4* __restrict__ v,
                                                     uint32_t n, float3 lo, float3 hi, float margin,
                                                     uint32_t* __restrict__ counters) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    bool bad = false, out = false;
    if (i < n) {
        const float4 p = x[i], q = v[i];
        bad = !(isfinite(p.x) && isfinite(p.y) && isfinite(p.z) && isfinite(q.x) && isfinite(q.y) && isfinite(q.z));
        out = !bad && (p.x < lo.x - margin || p.y < lo.y - margin || p.z < lo.z - margin || p.x > hi.x + margin ||
                       p.y > hi.y + margin || p.z > hi.z + margin);
    }
    const uint32_t nb = (uint32_t)__popcll(__ballot(bad)), no = (uint32_t)__popcll(__ballot(out));
    if ((threadIdx.x & 63u) == 0u) {
        if (nb) atomicAdd(&counters[0], nb);
        if (no) atomicAdd(&counters[1], no);
    }
}

int launch_check_state(const float4* x, const float4* v, uint32_t n, const float* lo, const float* hi, float margin,
                       uint32_t* counters, hipStream_t s) {
    hipLaunchKernelGGL(k_check_clear, dim3(1), dim3(64), 0, s, counters);
    VFS_TRY(check_launch("check_clear"));
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_check_state, dim3((n + 255u) / 256u), dim3(256), 0, s, x, v, n,
                       make_float3(lo[0], lo[1], lo[2]), make_float3(hi[0], hi[1], hi[2]), margin, counters);
    return check_launch("check_state");
}

int launch_reorder_state(const StepBuffers& b, hipStream_t s) {
    if (b.n == 0) return 0;
    const uint32_t nb = particle_blocks(b.n, 256);
    hipLaunchKernelGGL(k_reorder_state, dim3(nb), dim3(256), 0, s, b, nb);
    return check_launch("reorder_state");
}

int launch_density(const StepBuffers& b, const SimConsts& c, const SceneTable* table, const DevObject* objs,
                   int block, int staged, hipStream_t s) {
    if (b.n == 0) return 0;
    const uint32_t bs = block == 128 ? 128u : 256u;
    const uint32_t nb = particle_blocks(b.n, bs);
    const SceneTable t = table ? *table : SceneTable{nullptr, nullptr};
#define VFS_DENSITY(BS, ST, BA) \
    hipLaunchKernelGGL((k_density<BS, ST, BA>), dim3(nb), dim3(BS), 0, s, b, c, t, objs, nb)
    if (table) {
        if (bs == 128) {
            if (staged) VFS_DENSITY(128, true, true); else VFS_DENSITY(128, false, true);
        } else {
            if (staged) VFS_DENSITY(256, true, true); else VFS_DENSITY(256, false, true);
        }
    } else {
        if (bs == 128) {
            if (staged) VFS_DENSITY(128, true, false); else VFS_DENSITY(128, false, false);
        } else {
            if (staged) VFS_DENSITY(256, true, false); else VFS_DENSITY(256, false, false);
        }
    }
#undef VFS_DENSITY
    return check_launch("density");
}

// Single scenes up to this many particles (about two waves per SIMD of the
// 1,024) take the force pass with list indices loaded two chunks ahead.
constexpr uint32_t kForcesDeepMax = 131072;

int launch_forces_integrate(const StepBuffers& b, const SimConsts& c, const SceneTable* table, int block,
                            hipStream_t s) {
    if (b.n == 0) return 0;
    const uint32_t bs = block == 128 ? 128u : 256u;
    const uint32_t nb = particle_blocks(b.n, bs);
    const SceneTable t = table ? *table : SceneTable{nullptr, nullptr};
    if (table) {
        if (bs == 128)
            hipLaunchKernelGGL((k_forces_integrate<128, true>), dim3(nb), dim3(128), 0, s, b, c, t, nb);
        else
            hipLaunchKernelGGL((k_forces_integrate<256, true>), dim3(nb), dim3(256), 0, s, b, c, t, nb);
    } else if (b.n <= kForcesDeepMax) {
        if (bs == 128)
            hipLaunchKernelGGL((k_forces_integrate<128, false, true>), dim3(nb), dim3(128), 0, s, b, c, t, nb);
        else
            hipLaunchKernelGGL((k_forces_integrate<256, false, true>), dim3(nb), dim3(256), 0, s, b, c, t, nb);
    } else {
        if (bs == 128)
            hipLaunchKernelGGL((k_forces_integrate<128, false>), dim3(nb), dim3(128), 0, s, b, c, t, nb);
        else
            hipLaunchKernelGGL((k_forces_integrate<256, false>), dim3(nb), dim3(256), 0, s, b, c, t, nb);
    }
    return check_launch("forces_integrate");
}

int launch_lague_predict(const StepBuffers& b, const SimConsts& c, const LagueConsts& lc, hipStream_t s) {
    if (b.n == 0) return 0;
    const uint32_t nb = particle_blocks(b.n, 256);
    hipLaunchKernelGGL(k_lague_predict, dim3(nb), dim3(256), 0, s, b, c, lc, nb);
    return check_launch("lague_predict");
}

int launch_lague_density(const StepBuffers& b, const SimConsts& c, const LagueConsts& lc, hipStream_t s) {
    if (b.n == 0) return 0;
    const uint32_t nb = particle_blocks(b.n, 128);
    hipLaunchKernelGGL(k_lague_density<true>, dim3(nb), dim3(128), 0, s, b, c, lc, nb);
    return check_launch("lague_density");
}

int launch_lague_forces(const StepBuffers& b, const SimConsts& c, const LagueConsts& lc, hipStream_t s) {
    if (b.n == 0) return 0;
    const uint32_t nb = particle_blocks(b.n, 128);
    hipLaunchKernelGGL(k_lague_forces, dim3(nb), dim3(128), 0, s, b, c, lc, nb);
    return check_launch("lague_forces");
}

int launch_volume_map(int kind, const double* shape_params, const uint32_t* res, double support_radius,
                      double tolerance, const float* dmin, const double* step, const double* gl_x,
                      const double* gl_w, float* sdf_out, float* vm_out, hipStream_t s) {
    ShapeArgs sh;
    sh.kind = kind;
    for (int i = 0; i < 6; i++) sh.p[i] = shape_params[i];
    GLTable gl;
    for (int i = 0; i < 16; i++) {
        gl.x[i] = gl_x[i];
        gl.w[i] = gl_w[i];
    }
    const uint64_t total = (uint64_t)res[0] * res[1] * res[2];
    hipLaunchKernelGGL(k_volume_map, dim3(grid_for(total, 128)), dim3(128), 0, s, sh, res[0], res[1], res[2],
                       (float)support_radius, tolerance, dmin[0], dmin[1], dmin[2], step[0], step[1], step[2], gl,
                       sdf_out, vm_out);
    return check_launch("volume_map");
}

}  // namespace vfs
