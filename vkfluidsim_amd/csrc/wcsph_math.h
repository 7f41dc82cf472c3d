// Device arithmetic shared by the substep kernels (wcsph.hip) and their
// adjoint (adjoint.hip): smoothing kernels, equation of state, cell keys,
// grid sampling, boundary volume, launch-shape helpers.  Header-only; every
// function is internal to the translation unit that includes it.
#pragma once

#include "sort_internal.h"

namespace vfs {

namespace {

// dot(a, b) as fma(az, bz, fma(ay, by, ax * bx)).
__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    return __builtin_fmaf(az, bz, __builtin_fmaf(ay, by, ax * bx));
}

// 1 / x rounded to nearest, as the IEEE division 1.0f / x, for x in
// [kRcpMin, kRcpMax]: the hardware reciprocal (within one ulp) refined by one
// Newton-Raphson step with fused residual, y + y (1 - x y) (Markstein's
// reciprocal iteration).  Three instructions instead of the division's
// eleven; tests/test_gpu_arith.py compares it with 1.0f / x for every float in
// that range.  Callers keep their operands inside it (sim.cpp bounds the
// smoothing radius so that the force terms' divisors do).
constexpr float kRcpMin = 0x1p-100f, kRcpMax = 0x1p100f;
__device__ __forceinline__ float rcp_rn(float x) {
    const float y = __builtin_amdgcn_rcpf(x);
    const float e = __builtin_fmaf(-x, y, 1.0f);
    return __builtin_fmaf(e, y, y);
}

// CubicSpline, kernels_3d.slang:62-76
__device__ __forceinline__ float kernel_w(float r, const SimConsts& c) {
    const float q = r * c.inv_h;
    if (q <= 1.0f) {
        if (q > 0.5f) {
            const float t = 1.0f - q;
            return (c.sigma * 2.0f) * ((t * t) * t);
        }
        const float q2 = q * q;
        const float q3 = q2 * q;
        return c.sigma * ((6.0f * q3 - 6.0f * q2) + 1.0f);
    }
    return 0.0f;
}

// GradCubicSpline, kernels_3d.slang:78-91 (grad_sigma carries the extra 1/h)
__device__ __forceinline__ float kernel_dw(float r, const SimConsts& c) {
    const float q = r * c.inv_h;
    if ((r > 1e-9f) && (q <= 1.0f)) {
        if (q > 0.5f) {
            const float t = 1.0f - q;
            return (-c.grad_sigma) * (t * t);
        }
        return (c.grad_sigma * q) * (3.0f * q - 2.0f);
    }
    return 0.0f;
}

// (rho / rho0)^gamma: binary powering in double for an integral exponent,
// rounded once to float; general exponents use the double pow.
__device__ __forceinline__ float eos_pow(float x, const SimConsts& c) {
    if (c.gamma_int >= 0) {
        unsigned e = (unsigned)c.gamma_int;
        double base = (double)x, result = 1.0;
        while (e) {
            if (e & 1u) result *= base;
            base *= base;
            e >>= 1;
        }
        return (float)result;
    }
    return (float)pow((double)x, (double)c.gamma);
}

// PressureFromDensityEOS, slang:199-203
__device__ __forceinline__ float eos(float rho, const SimConsts& c) {
    rho = fmaxf(rho, c.rho0);
    return c.B * (eos_pow(rho / c.rho0, c) - 1.0f);
}

__device__ __forceinline__ uint32_t cell_key(int32_t cx, int32_t cy, int32_t cz, int dx, int dy, int dz,
                                             const SimConsts& c) {
    return c.base + fastmod(hash_cell((int32_t)((uint32_t)cx + (uint32_t)dx),
                                      (int32_t)((uint32_t)cy + (uint32_t)dy),
                                      (int32_t)((uint32_t)cz + (uint32_t)dz)),
                            c.fm);
}

// Constants of the scene of entry i: the launch argument for a single scene,
// the wave's row of the scene table for a batch (i >> 6 is uniform over the
// wave, so the row is read with scalar loads).
template <bool BATCH>
__device__ __forceinline__ SimConsts scene_consts(const SimConsts& c0, const SceneTable& t, uint32_t i) {
    if (!BATCH) return c0;
    const uint32_t s = __builtin_amdgcn_readfirstlane(t.wave_scene[i >> 6]);
    return t.consts[s];
}

// SampleGrid, slang:51-66, base cell clamped to [0, n-2], in three parts: the
// cell and its local coordinates, the eight corner values, the trilinear sum.
__device__ __forceinline__ uint32_t grid_cell(const uint32_t* res, const float* uvw, float* chi) {
    uint32_t ci[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float x = ((float)res[a] - 1.0f) * uvw[a];
        const float f = floorf(x);
        const uint32_t hi = res[a] >= 2u ? res[a] - 2u : 0u;
        uint32_t cc;
        if (f >= 0.0f)
            cc = f <= (float)hi ? (uint32_t)f : hi;
        else
            cc = 0u;
        ci[a] = cc;
        chi[a] = 2.0f * (x - (float)cc) - 1.0f;
    }
    return ci[0] + res[0] * ci[1] + (res[0] * res[1]) * ci[2];
}

// corners in the order (0,0,0),(0,1,0),(0,0,1),(0,1,1),(1,0,0),(1,1,0),(1,0,1),(1,1,1)
__device__ __forceinline__ void grid_corners(const float* __restrict__ grid, const uint32_t* res, uint32_t b0,
                                             float* g) {
    const uint32_t sx = 1u, sy = res[0], sz = res[0] * res[1];
    g[0] = grid[b0];
    g[1] = grid[b0 + sy];
    g[2] = grid[b0 + sz];
    g[3] = grid[b0 + sy + sz];
    g[4] = grid[b0 + sx];
    g[5] = grid[b0 + sx + sy];
    g[6] = grid[b0 + sx + sz];
    g[7] = grid[b0 + sx + sy + sz];
}

__device__ __forceinline__ float grid_interp(const float* g, const float* chi) {
    const float mx = 1.0f + (-1.0f) * chi[0], px = 1.0f + chi[0];
    const float my = 1.0f + (-1.0f) * chi[1], py = 1.0f + chi[1];
    const float mz = 1.0f + (-1.0f) * chi[2], pz = 1.0f + chi[2];
    float val = 0.0f;
    val += g[0] * ((mx * my) * mz);
    val += g[1] * ((mx * py) * mz);
    val += g[2] * ((mx * my) * pz);
    val += g[3] * ((mx * py) * pz);
    val += g[4] * ((px * my) * mz);
    val += g[5] * ((px * py) * mz);
    val += g[6] * ((px * my) * pz);
    val += g[7] * ((px * py) * pz);
    return val / 8.0f;
}

__device__ __forceinline__ void to_uvw(const DevObject& o, const float* p, float* uvw) {
    uvw[0] = (p[0] - o.box_pos[0]) / o.box_size[0];
    uvw[1] = (p[1] - o.box_pos[1]) / o.box_size[1];
    uvw[2] = (p[2] - o.box_pos[2]) / o.box_size[2];
}

__device__ __forceinline__ void mat_vec(const float* M, const float* v, float* out) {
#pragma unroll
    for (int r = 0; r < 3; r++) out[r] = ((M[r] * v[0] + M[4 + r] * v[1]) + M[8 + r] * v[2]) + M[12 + r] * 1.0f;
}

// CalculateBoundaryVolume for one particle, slang:102-151.
__device__ void boundary_volume(const DevObject* __restrict__ objs, uint32_t n_obj, const float* pos,
                                const SimConsts& c, float& psi, float* grad) {
    psi = 0.0f;
    grad[0] = grad[1] = grad[2] = 0.0f;
    for (uint32_t oi = 0; oi < n_obj; oi++) {
        const DevObject& o = objs[oi];
        float lp[3];
        mat_vec(o.T, pos, lp);
        bool outside = false;
#pragma unroll
        for (int a = 0; a < 3; a++)
            outside |= (lp[a] < o.box_pos[a]) || (lp[a] > o.box_pos[a] + o.box_size[a]);
        if (outside) continue;
        float uvw[3], chi[3], gs[8], g8[8];
        to_uvw(o, lp, uvw);
        const uint32_t cell = grid_cell(o.res, uvw, chi);
        grid_corners(o.sdf, o.res, cell, gs);
        const float dist = grid_interp(gs, chi);
        if (!(dist >= 0.0f && dist < c.h)) continue;
        grid_corners(o.vm, o.res, cell, g8);
        const float volume = grid_interp(g8, chi);
        if (!(volume > 0.0f)) continue;
        // GradientVector, slang:68-85; samples in the particle's grid cell
        // reuse its SDF corners
        const float eps = 0.1f * c.h;
        auto sdf_at = [&](const float* q) {
            float u[3], ch[3];
            to_uvw(o, q, u);
            const uint32_t cq = grid_cell(o.res, u, ch);
            if (cq == cell) return grid_interp(gs, ch);
            grid_corners(o.sdf, o.res, cq, g8);
            return grid_interp(g8, ch);
        };
        float nrm[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            float xp[3] = {lp[0], lp[1], lp[2]}, xm[3] = {lp[0], lp[1], lp[2]};
            xp[a] = lp[a] + eps;
            xm[a] = lp[a] - eps;
            const float v1 = sdf_at(xp);
            const float v0 = sdf_at(xm);
            nrm[a] = (v1 - v0) * (1.0f / (2.0f * eps));
        }
        const float modn = sqrtf(dot3(nrm[0], nrm[1], nrm[2], nrm[0], nrm[1], nrm[2]));
        if (!(modn > 1e-9f)) continue;
        float rn[3];
        mat_vec(o.Rm, nrm, rn);
        const float inv_modn = 1.0f / modn;
#pragma unroll
        for (int a = 0; a < 3; a++) rn[a] = rn[a] * inv_modn;
        const float d = fmaxf(dist + 0.125f * c.h, 0.5f * c.h);
        float xij[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float xb = pos[a] - d * rn[a];
            xij[a] = pos[a] - xb;
        }
        const float xm = sqrtf(dot3(xij[0], xij[1], xij[2], xij[0], xij[1], xij[2]));
        psi += volume * kernel_w(xm, c);
        const float vg = volume * kernel_dw(xm, c);
        const float inv_xm = xm > 0.0f ? 1.0f / xm : 0.0f;
#pragma unroll
        for (int a = 0; a < 3; a++) grad[a] += vg * (xm > 0.0f ? xij[a] * inv_xm : 0.0f);
    }
}

// ------------------------------------------------------------------ kernels

// The hardware deals workgroups to the 8 XCDs round robin; each XCD has its own
// L2.  XCD x is given chunks x, x + 8, x + 16, ... of kXcdChunk consecutive
// blocks (compact slabs of sorted particles, so a chunk's neighbour cells are
// mostly in the same L2), and the blocks past the last whole round of chunks
// contiguously.  Chunks rather than one contiguous eighth per XCD: where the
// fluid is uneven an eighth of the sorted order can hold 7% more neighbour work
// than the mean, and the slowest XCD sets the kernel's end.  Measured on 1M
// particles in contact: 16 blocks per chunk 0.452 ms per substep, 8 0.456,
// 64 0.455, 256 0.458, one eighth per XCD 0.459, no remap (1) 0.472.
// nblocks is a multiple of 8.
constexpr uint32_t kXcdChunk = 16;
static_assert(kXcdChunk >= 1u && kXcdChunk <= (1u << 28), "8 chunks must fit the 32-bit block count");
__device__ __forceinline__ uint32_t xcd_block(uint32_t nblocks) {
    const uint32_t b = blockIdx.x, x = b & 7u, k = b >> 3;
    const uint32_t round = 8u * kXcdChunk;
    const uint32_t full = nblocks / round;  // whole rounds of eight chunks
    if (k < full * kXcdChunk) return (k / kXcdChunk) * round + x * kXcdChunk + k % kXcdChunk;
    const uint32_t rest = (nblocks - full * round) >> 3;
    return full * round + x * rest + (k - full * kXcdChunk);
}

// Raw buffer resources over the particle arrays: 32-bit offsets and hardware
// range checks (a load past num_records bytes returns zeros), so loads by
// neighbour index need no 64-bit address arithmetic, and the prefetch of a
// stale queue slot cannot fault.
constexpr int kBufferDword3 = 0x00020000;  // raw (untyped) buffer access on gfx9

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, kBufferDword3);
}

__device__ __forceinline__ float4 load_f4(__amdgpu_buffer_rsrc_t r, uint32_t index) {
    const auto raw = __builtin_amdgcn_raw_buffer_load_b128(r, (int)(index * 16u), 0, 0);
    float4 v;
    __builtin_memcpy(&v, &raw, sizeof(v));
    return v;
}

__device__ __forceinline__ float2 load_f2(__amdgpu_buffer_rsrc_t r, uint32_t index) {
    const auto raw = __builtin_amdgcn_raw_buffer_load_b64(r, (int)(index * 8u), 0, 0);
    float2 v;
    __builtin_memcpy(&v, &raw, sizeof(v));
    return v;
}


}  // namespace

}  // namespace vfs
