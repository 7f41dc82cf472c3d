// Reverse-mode (adjoint) kernels of the WCSPH substep, and the density-grid
// loss.  The forward substep is the one of wcsph.hip; its recorded quantities
// (permutation, sorted x v with the EOS terms in .w, densities, boundary
// terms, neighbour lists) are read back per substep in sorted order.
//
// The neighbour lists are frozen: the adjoint differentiates the substep with
// the pair set of the forward pass (the smoothing kernel and its gradient
// vanish with their first derivatives at r = h, so pairs entering or leaving
// the support add nothing).  Every per-particle adjoint is gathered over the
// particle's own list, no atomics: a pair term of particle j that depends on
// x_i is found through i's list entry for j.  A pair appears in both lists
// equally often unless the 27 stencil keys of one of the two particles repeat
// (hash collisions); then the neighbour's term is weighted by the ratio of the
// two multiplicities, counted from the stencil keys.
#include "wcsph_math.h"

namespace vfs {

namespace {

// dG/dr of GradCubicSpline (G = kernel_dw, which carries the extra 1/h).
__device__ __forceinline__ float kernel_ddw(float r, const SimConsts& c) {
    const float q = r * c.inv_h;
    if ((r > 1e-9f) && (q <= 1.0f)) {
        if (q > 0.5f) return ((2.0f * c.grad_sigma) * (1.0f - q)) * c.inv_h;
        return (c.grad_sigma * (6.0f * q - 2.0f)) * c.inv_h;
    }
    return 0.0f;
}

// dp/drho of PressureFromDensityEOS: B gamma (rho/rho0)^(gamma-1) / rho0 above rho0.
__device__ __forceinline__ float eos_dp(float rho, const SimConsts& c) {
    if (!(rho > c.rho0)) return 0.0f;
    const double x = (double)(rho / c.rho0);
    double xp;
    if (c.gamma_int >= 1) {
        unsigned e = (unsigned)(c.gamma_int - 1);
        double base = x;
        xp = 1.0;
        while (e) {
            if (e & 1u) xp *= base;
            base *= base;
            e >>= 1;
        }
    } else {
        xp = pow(x, (double)c.gamma - 1.0);
    }
    return (float)((double)c.B * (double)c.gamma * xp / (double)c.rho0);
}

__device__ __forceinline__ uint32_t own_key(float x, float y, float z, const SimConsts& c) {
    return cell_key(floor_to_cell(x / c.cell), floor_to_cell(y / c.cell), floor_to_cell(z / c.cell), 0, 0, 0, c);
}

// The 27 stencil keys of a position, in the walk order.
__device__ __forceinline__ void stencil_keys(float x, float y, float z, const SimConsts& c, uint32_t* k) {
    const int32_t cx = floor_to_cell(x / c.cell), cy = floor_to_cell(y / c.cell), cz = floor_to_cell(z / c.cell);
    int t = 0;
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) k[t++] = cell_key(cx, cy, cz, dx, dy, dz, c);
}

__device__ __forceinline__ uint32_t stencil_count(float x, float y, float z, uint32_t key, const SimConsts& c) {
    uint32_t k[27];
    stencil_keys(x, y, z, c, k);
    uint32_t m = 0;
    for (int t = 0; t < 27; t++) m += k[t] == key ? 1u : 0u;
    return m;
}

// Weight of a neighbour's pair term found through i's list entry for j:
// multiplicity of i in j's list / multiplicity of j in i's list.
__device__ __forceinline__ float pair_weight(const float4& xi, const float4& xj, bool dup_i, bool dup_j,
                                             const SimConsts& c) {
    if (!dup_i && !dup_j) return 1.0f;
    const uint32_t ki = own_key(xi.x, xi.y, xi.z, c), kj = own_key(xj.x, xj.y, xj.z, c);
    const uint32_t m_ij = stencil_count(xi.x, xi.y, xi.z, kj, c);  // times j is in i's list
    const uint32_t m_ji = stencil_count(xj.x, xj.y, xj.z, ki, c);  // times i is in j's list
    return m_ij ? (float)m_ji / (float)m_ij : 0.0f;
}

// d(trilinear value)/d(uvw) for corners g and cell coordinates chi.
__device__ __forceinline__ void grid_grad(const float* g, const float* chi, const uint32_t* res, float* out) {
    const float f0[3] = {1.0f - chi[0], 1.0f - chi[1], 1.0f - chi[2]};
    const float f1[3] = {1.0f + chi[0], 1.0f + chi[1], 1.0f + chi[2]};
    // corner k has offsets (ox, oy, oz) = (k >> 2, k & 1, (k >> 1) & 1)
    float dx = 0.0f, dy = 0.0f, dz = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int ox = k >> 2, oy = k & 1, oz = (k >> 1) & 1;
        const float fx = ox ? f1[0] : f0[0], fy = oy ? f1[1] : f0[1], fz = oz ? f1[2] : f0[2];
        const float sx = ox ? 1.0f : -1.0f, sy = oy ? 1.0f : -1.0f, sz = oz ? 1.0f : -1.0f;
        dx += g[k] * (sx * fy * fz);
        dy += g[k] * (fx * sy * fz);
        dz += g[k] * (fx * fy * sz);
    }
    out[0] = dx * (2.0f * ((float)res[0] - 1.0f)) / 8.0f;
    out[1] = dy * (2.0f * ((float)res[1] - 1.0f)) / 8.0f;
    out[2] = dz * (2.0f * ((float)res[2] - 1.0f)) / 8.0f;
}

// Adjoint of CalculateBoundaryVolume (boundary_volume) at `pos`: adds to g
// the derivative of psi_bar * psi + grad_bar . grad psi with respect to pos,
// on the branch the forward took (grid cells held fixed).
__device__ void boundary_volume_adjoint(const DevObject* __restrict__ objs, uint32_t n_obj, const float* pos,
                                        const SimConsts& c, float psi_bar, const float* grad_bar, float* gpos) {
    for (uint32_t oi = 0; oi < n_obj; oi++) {
        const DevObject& o = objs[oi];
        float lp[3];
        mat_vec(o.T, pos, lp);
        bool outside = false;
#pragma unroll
        for (int a = 0; a < 3; a++)
            outside |= (lp[a] < o.box_pos[a]) || (lp[a] > o.box_pos[a] + o.box_size[a]);
        if (outside) continue;
        const float inv_bs[3] = {1.0f / o.box_size[0], 1.0f / o.box_size[1], 1.0f / o.box_size[2]};
        float uvw[3], chi[3], gs[8], gv[8];
        to_uvw(o, lp, uvw);
        const uint32_t cell = grid_cell(o.res, uvw, chi);
        grid_corners(o.sdf, o.res, cell, gs);
        const float dist = grid_interp(gs, chi);
        if (!(dist >= 0.0f && dist < c.h)) continue;
        grid_corners(o.vm, o.res, cell, gv);
        const float volume = grid_interp(gv, chi);
        if (!(volume > 0.0f)) continue;
        float gS[3], gV[3];
        grid_grad(gs, chi, o.res, gS);
        grid_grad(gv, chi, o.res, gV);
#pragma unroll
        for (int a = 0; a < 3; a++) {
            gS[a] *= inv_bs[a];
            gV[a] *= inv_bs[a];
        }
        // central-difference normal and its Jacobian Jn[a][b] = d nrm_a / d lp_b
        const float eps = 0.1f * c.h;
        float nrm[3], Jn[3][3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            float v[2], gr[2][3];
#pragma unroll
            for (int sgn = 0; sgn < 2; sgn++) {
                float q[3] = {lp[0], lp[1], lp[2]};
                q[a] = sgn == 0 ? lp[a] + eps : lp[a] - eps;
                float u[3], ch[3], g8[8];
                to_uvw(o, q, u);
                const uint32_t cq = grid_cell(o.res, u, ch);
                grid_corners(o.sdf, o.res, cq, g8);
                v[sgn] = grid_interp(g8, ch);
                grid_grad(g8, ch, o.res, gr[sgn]);
            }
            nrm[a] = (v[0] - v[1]) * (1.0f / (2.0f * eps));
#pragma unroll
            for (int b2 = 0; b2 < 3; b2++) Jn[a][b2] = (gr[0][b2] - gr[1][b2]) * inv_bs[b2] / (2.0f * eps);
        }
        const float modn = sqrtf(dot3(nrm[0], nrm[1], nrm[2], nrm[0], nrm[1], nrm[2]));
        if (!(modn > 1e-9f)) continue;
        // rn = (R nrm + t) / modn
        float Rn[3], RJ[3][3];
#pragma unroll
        for (int r = 0; r < 3; r++) {
            Rn[r] = o.Rm[r] * nrm[0] + o.Rm[4 + r] * nrm[1] + o.Rm[8 + r] * nrm[2] + o.Rm[12 + r];
#pragma unroll
            for (int b2 = 0; b2 < 3; b2++)
                RJ[r][b2] = o.Rm[r] * Jn[0][b2] + o.Rm[4 + r] * Jn[1][b2] + o.Rm[8 + r] * Jn[2][b2];
        }
        const float inv_m = 1.0f / modn;
        float nJ[3];  // nrm^T Jn
#pragma unroll
        for (int b2 = 0; b2 < 3; b2++) nJ[b2] = nrm[0] * Jn[0][b2] + nrm[1] * Jn[1][b2] + nrm[2] * Jn[2][b2];
        float rn[3], Jr[3][3];
#pragma unroll
        for (int r = 0; r < 3; r++) {
            rn[r] = Rn[r] * inv_m;
#pragma unroll
            for (int b2 = 0; b2 < 3; b2++) Jr[r][b2] = RJ[r][b2] * inv_m - Rn[r] * nJ[b2] * (inv_m * inv_m * inv_m);
        }
        const float dd = dist + 0.125f * c.h;
        const float d = fmaxf(dd, 0.5f * c.h);
        const bool d_follows = dd > 0.5f * c.h;
        // xij = d rn, Jx = rn gd^T + d Jr
        float xij[3], Jx[3][3];
#pragma unroll
        for (int r = 0; r < 3; r++) {
            xij[r] = d * rn[r];
#pragma unroll
            for (int b2 = 0; b2 < 3; b2++) Jx[r][b2] = (d_follows ? rn[r] * gS[b2] : 0.0f) + d * Jr[r][b2];
        }
        const float xm = sqrtf(dot3(xij[0], xij[1], xij[2], xij[0], xij[1], xij[2]));
        if (!(xm > 0.0f)) continue;
        const float xh[3] = {xij[0] / xm, xij[1] / xm, xij[2] / xm};
        float dxm[3];  // x^T Jx
#pragma unroll
        for (int b2 = 0; b2 < 3; b2++) dxm[b2] = xh[0] * Jx[0][b2] + xh[1] * Jx[1][b2] + xh[2] * Jx[2][b2];
        const float W = kernel_w(xm, c), G = kernel_dw(xm, c), Gp = kernel_ddw(xm, c);
        const float Wp = c.h * G;
        const float xg = xh[0] * grad_bar[0] + xh[1] * grad_bar[1] + xh[2] * grad_bar[2];
        float perp[3];  // (I - x x^T) grad_bar / xm
#pragma unroll
        for (int r = 0; r < 3; r++) perp[r] = (grad_bar[r] - xg * xh[r]) / xm;
        float glp[3];
#pragma unroll
        for (int b2 = 0; b2 < 3; b2++) {
            const float jp = Jx[0][b2] * perp[0] + Jx[1][b2] * perp[1] + Jx[2][b2] * perp[2];
            glp[b2] = psi_bar * (W * gV[b2] + volume * Wp * dxm[b2]) + xg * (G * gV[b2] + volume * Gp * dxm[b2]) +
                      volume * G * jp;
        }
        // lp = T pos + t
#pragma unroll
        for (int b2 = 0; b2 < 3; b2++)
            gpos[b2] += o.T[b2 * 4 + 0] * glp[0] + o.T[b2 * 4 + 1] * glp[1] + o.T[b2 * 4 + 2] * glp[2];
    }
}

struct TapeView {
    const uint32_t* perm;
    const float4* xs;  // w: p / rho^2
    const float4* vs;  // w: m / max(rho, 1e-6)
    const float* rho;
    const float4* bnd;
    const uint32_t* nbr;
    const uint32_t* cnt;
};

struct AdjBuffers {
    const float4* gx_in;  // adjoint of the substep's output positions (sorted order)
    const float4* gv_in;
    float4* abar;         // adjoint of the acceleration
    float4* gx;           // adjoint of the substep's sorted input positions
    float4* gv;
    float* rbar;          // adjoint of the densities
    float4* gpsi;         // adjoint of grad psi
    uint32_t* dup;        // 1: the particle's stencil keys repeat
    float4* gx_out;       // gx, gv in the order of the substep's input arrays
    float4* gv_out;
    double* nu_part;      // per-block viscosity adjoint of this substep
    uint32_t n;
};

// Symplectic Euler and wall penalty backwards; stencil-key repeats.
__global__ __launch_bounds__(256) void k_adj_integrate(TapeView t, AdjBuffers a, SimConsts c) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.n) return;
    const float4 gxo = a.gx_in[i], gvo = a.gv_in[i];
    const float vt[3] = {gvo.x + c.dt * gxo.x, gvo.y + c.dt * gxo.y, gvo.z + c.dt * gxo.z};
    const float ab[3] = {c.dt * vt[0], c.dt * vt[1], c.dt * vt[2]};
    const float4 x4 = t.xs[i];
    const float p[3] = {x4.x, x4.y, x4.z};
    float gx[3] = {gxo.x, gxo.y, gxo.z};
    float diff[3];
#pragma unroll
    for (int k = 0; k < 3; k++) diff[k] = fminf(p[k] - c.bb_lo[k], 0.0f) - fminf(c.bb_hi[k] - p[k], 0.0f);
    const bool wall = fabsf(diff[0]) > 1e-9f || fabsf(diff[1]) > 1e-9f || fabsf(diff[2]) > 1e-9f;
    if (wall) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const float D = (p[k] < c.bb_lo[k] ? 1.0f : 0.0f) + (p[k] > c.bb_hi[k] ? 1.0f : 0.0f);
            gx[k] += c.kwall_neg * D * ab[k];
        }
    }
    a.abar[i] = make_float4(ab[0], ab[1], ab[2], 0.0f);
    a.gx[i] = make_float4(gx[0], gx[1], gx[2], 0.0f);
    a.gv[i] = make_float4(vt[0], vt[1], vt[2], 0.0f);
    uint32_t k[27];
    stencil_keys(p[0], p[1], p[2], c, k);
    uint32_t dup = 0;
    for (int u = 1; u < 27 && !dup; u++)
        for (int w = 0; w < u; w++) dup |= k[u] == k[w] ? 1u : 0u;
    a.dup[i] = dup;
}

// Pressure, viscosity and boundary pressure backwards: adjoints of x, v (pair
// terms), of the EOS terms and so of rho, of grad psi, and of the viscosity.
__global__ __launch_bounds__(256) void k_adj_forces(TapeView t, AdjBuffers a, SimConsts c) {
    __shared__ double s_nu[256 / 64];
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    double nu_local = 0.0;
    if (i < a.n) {
        const float4 xi = t.xs[i], vi = t.vs[i];
        const float Pi = xi.w, Mi = vi.w;
        const float4 abi = a.abar[i];
        const bool dup_i = a.dup[i] != 0u;
        const float ai[3] = {abi.x, abi.y, abi.z};
        // per-particle sums of large, cancelling pair terms: double accumulators
        double Pbar = 0.0, Mbar = 0.0, gx[3] = {0.0, 0.0, 0.0}, gv[3] = {0.0, 0.0, 0.0};
        double fv[3] = {0.0, 0.0, 0.0};
        const uint32_t count = t.cnt[i];
        const uint32_t* list = t.nbr;
        for (uint32_t e = 0; e < count; e++) {
            const uint32_t j = list[(size_t)(e >> 2) * 4u * a.n + 4u * i + (e & 3u)];
            if (j == i) continue;
            const float4 xj = t.xs[j], vj = t.vs[j];
            const float xij[3] = {xi.x - xj.x, xi.y - xj.y, xi.z - xj.z};
            const float r2 = dot3(xij[0], xij[1], xij[2], xij[0], xij[1], xij[2]);
            const float r = sqrtf(r2);
            if (!(r > 1e-9f)) continue;
            const float inv_r = 1.0f / r;
            const float nh[3] = {xij[0] * inv_r, xij[1] * inv_r, xij[2] * inv_r};
            const float G = kernel_dw(r, c), Gp = kernel_ddw(r, c);
            const float F[3] = {nh[0] * G, nh[1] * G, nh[2] * G};
            const float vij[3] = {vi.x - vj.x, vi.y - vj.y, vi.z - vj.z};
            const float sig = dot3(vij[0], vij[1], vij[2], xij[0], xij[1], xij[2]);
            const float inv_den = 1.0f / (r2 + c.visc_eps_h2);
            const float Pj = xj.w, Mj = vj.w;
            const float4 abj4 = a.abar[j];
            const float aj[3] = {abj4.x, abj4.y, abj4.z};
            const float w = pair_weight(xi, xj, dup_i, a.dup[j] != 0u, c);
            // forward viscous sum (for the viscosity adjoint)
#pragma unroll
            for (int k = 0; k < 3; k++) fv[k] += Mj * sig * inv_den * F[k];
            // own term of a_i: -m (Pi + Pj) F + nu10 Mj sig inv_den F
            const float aF = ai[0] * F[0] + ai[1] * F[1] + ai[2] * F[2];
            Pbar += -c.mass * aF;
            float Fb[3];
#pragma unroll
            for (int k = 0; k < 3; k++) Fb[k] = -c.mass * (Pi + Pj) * ai[k] + c.nu10 * Mj * sig * inv_den * ai[k];
            float sb = c.nu10 * Mj * inv_den * aF;
            float db = -c.nu10 * Mj * sig * inv_den * inv_den * aF;
            // neighbour term of a_j (its pair with i): F_ji = -F, x_ji = -x_ij, v_ji = -v_ij
            const float bF = -(aj[0] * F[0] + aj[1] * F[1] + aj[2] * F[2]);  // a_j . F_ji
            Pbar += w * (-c.mass * bF);
            Mbar += w * (c.nu10 * sig * inv_den * bF);
            // d(a_j term)/d x_i = -d/d x_ji; fold it into the own-term adjoint by
            // expressing it in x_ij: F_ji(x_ji) = -F(x_ij), sig and den even
            float Fb2[3];
#pragma unroll
            for (int k = 0; k < 3; k++)
                Fb2[k] = -(-c.mass * (Pj + Pi) * aj[k] + c.nu10 * Mi * sig * inv_den * aj[k]);
            const float sb2 = c.nu10 * Mi * inv_den * bF;
            const float db2 = -c.nu10 * Mi * sig * inv_den * inv_den * bF;
#pragma unroll
            for (int k = 0; k < 3; k++) Fb[k] += w * Fb2[k];
            sb += w * sb2;
            db += w * db2;
            // F = x_ij G / r:  dF/dx_ij^T Fb = (G/r) Fb + (Fb . n)(G' - G/r) n
            const float Fn = Fb[0] * nh[0] + Fb[1] * nh[1] + Fb[2] * nh[2];
            const float gr = G * inv_r;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                gx[k] += gr * Fb[k] + Fn * (Gp - gr) * nh[k] + sb * vij[k] + 2.0f * db * xij[k];
                gv[k] += sb * xij[k];
            }
        }
        // viscosity: a_i = ... + nu10 * fv, nu10 = 10 nu
        nu_local = 10.0 * ((double)ai[0] * fv[0] + (double)ai[1] * fv[1] + (double)ai[2] * fv[2]);
        // boundary pressure: a_i = ... - (Pi + pb) rho0 grad psi
        const float rhoi = t.rho[i];
        const float4 bnd = t.bnd[i];
        const float p = eos(rhoi, c);
        const float pb = rhoi > 0.0f ? p / c.rho0_sq : 0.0f;
        const double ag = (double)ai[0] * bnd.x + (double)ai[1] * bnd.y + (double)ai[2] * bnd.z;
        Pbar += -c.rho0 * ag;
        const double pbbar = -c.rho0 * ag;
        const float sbnd = -(Pi + pb) * c.rho0;
        a.gpsi[i] = make_float4(sbnd * ai[0], sbnd * ai[1], sbnd * ai[2], 0.0f);
        // EOS terms to rho
        const double dp = eos_dp(rhoi, c);
        const double rd = rhoi;
        double rbar = rhoi > 0.0f ? pbbar * dp / c.rho0_sq : 0.0;
        if (rhoi > 0.0f) rbar += Pbar * (dp / (rd * rd) - 2.0 * (double)p / (rd * rd * rd));
        if (rhoi > 1e-6f) rbar += Mbar * (-(double)c.mass / (rd * rd));
        a.rbar[i] = (float)rbar;
        const float4 g0 = a.gx[i], v0 = a.gv[i];
        a.gx[i] = make_float4((float)(g0.x + gx[0]), (float)(g0.y + gx[1]), (float)(g0.z + gx[2]), 0.0f);
        a.gv[i] = make_float4((float)(v0.x + gv[0]), (float)(v0.y + gv[1]), (float)(v0.z + gv[2]), 0.0f);
    }
    // block sum of the viscosity adjoint
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) nu_local += __shfl_xor(nu_local, o, 64);
    if ((threadIdx.x & 63u) == 0u) s_nu[threadIdx.x >> 6] = nu_local;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < 256 / 64; w++) s += s_nu[w];
        a.nu_part[blockIdx.x] = s;
    }
}

// nu_bar += sum of the block partials, in a fixed order (run-to-run identical).
__global__ __launch_bounds__(256) void k_add_partials(const double* __restrict__ part, uint32_t n,
                                                      double* __restrict__ total) {
    __shared__ double s[256];
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += 256u) acc += part[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t o = 128; o >= 1; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total += s[0];
}

// Density and boundary volume backwards, then the adjoints of the substep's
// input in the order of its input arrays (inverse of the reorder).
__global__ __launch_bounds__(256) void k_adj_density(TapeView t, AdjBuffers a, SimConsts c,
                                                     const DevObject* __restrict__ objs) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.n) return;
    const float4 xi = t.xs[i];
    const float rbi = a.rbar[i];
    const bool dup_i = a.dup[i] != 0u;
    const float4 g0 = a.gx[i];
    double gx[3] = {g0.x, g0.y, g0.z};
    const uint32_t count = t.cnt[i];
    for (uint32_t e = 0; e < count; e++) {
        const uint32_t j = t.nbr[(size_t)(e >> 2) * 4u * a.n + 4u * i + (e & 3u)];
        if (j == i) continue;
        const float4 xj = t.xs[j];
        const float xij[3] = {xi.x - xj.x, xi.y - xj.y, xi.z - xj.z};
        const float r = sqrtf(dot3(xij[0], xij[1], xij[2], xij[0], xij[1], xij[2]));
        if (!(r > 1e-9f)) continue;
        const float w = pair_weight(xi, xj, dup_i, a.dup[j] != 0u, c);
        const double s = (double)(c.mass * (c.h * kernel_dw(r, c))) * ((double)rbi + w * (double)a.rbar[j]) / r;
#pragma unroll
        for (int k = 0; k < 3; k++) gx[k] += s * xij[k];
    }
    if (c.obj_count) {
        const float4 gp = a.gpsi[i];
        const float gb[3] = {gp.x, gp.y, gp.z};
        const float pos[3] = {xi.x, xi.y, xi.z};
        float gbnd[3] = {0.0f, 0.0f, 0.0f};
        boundary_volume_adjoint(objs + c.obj_begin, c.obj_count, pos, c, c.rho0 * rbi, gb, gbnd);
#pragma unroll
        for (int k = 0; k < 3; k++) gx[k] += gbnd[k];
    }
    const uint32_t o = t.perm[i];
    a.gx_out[o] = make_float4((float)gx[0], (float)gx[1], (float)gx[2], 0.0f);
    a.gv_out[o] = a.gv[i];
}

// Density field on grid points: rho(g) = sum over particles within h of
// m W(|g - x_j|), each particle once (repeated stencil keys are skipped).
struct GridDesc {
    float o[3], d[3];
    uint32_t dims[3];
};

__device__ __forceinline__ void grid_point(const GridDesc& gd, uint32_t idx, float* g) {
    const uint32_t ix = idx % gd.dims[0], iy = (idx / gd.dims[0]) % gd.dims[1], iz = idx / (gd.dims[0] * gd.dims[1]);
    g[0] = gd.o[0] + (float)ix * gd.d[0];
    g[1] = gd.o[1] + (float)iy * gd.d[1];
    g[2] = gd.o[2] + (float)iz * gd.d[2];
}

__global__ __launch_bounds__(256) void k_grid_density(GridDesc gd, const float4* __restrict__ xs,
                                                      const uint32_t* __restrict__ cs, SimConsts c,
                                                      double* __restrict__ out) {
    const uint32_t idx = blockIdx.x * 256u + threadIdx.x;
    const uint32_t total = gd.dims[0] * gd.dims[1] * gd.dims[2];
    if (idx >= total) return;
    float g[3];
    grid_point(gd, idx, g);
    uint32_t k[27];
    stencil_keys(g[0], g[1], g[2], c, k);
    // the loss is a small difference of densities: sum in double
    double sum = 0.0;
    for (int u = 0; u < 27; u++) {
        bool seen = false;
        for (int w = 0; w < u; w++) seen |= k[w] == k[u];
        if (seen) continue;
        for (uint32_t j = cs[k[u]]; j < cs[k[u] + 1u]; j++) {
            const float4 q = xs[j];
            const float dx = q.x - g[0], dy = q.y - g[1], dz = q.z - g[2];
            const float r2 = dot3(dx, dy, dz, dx, dy, dz);
            if (!(r2 > c.h2)) sum += (double)(c.mass * kernel_w(sqrtf(r2), c));
        }
    }
    out[idx] = sum;
}

// d loss / d x_i for loss = mean((rho(g) - target(g))^2), per particle over the
// grid points within h (the same pairs as k_grid_density).
__global__ __launch_bounds__(256) void k_grid_loss_grad(GridDesc gd, const float4* __restrict__ x, uint32_t n,
                                                        const double* __restrict__ rho_g,
                                                        const float* __restrict__ target, SimConsts c,
                                                        float4* __restrict__ gx) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const float4 p = x[i];
    const float pp[3] = {p.x, p.y, p.z};
    const double total = (double)gd.dims[0] * gd.dims[1] * gd.dims[2];
    int lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        // clamped in float before the conversion (NaN -> whole range; such
        // particles then match no grid point)
        lo[a] = (int)fmaxf(floorf((pp[a] - c.h - gd.o[a]) / gd.d[a]) - 1.0f, 0.0f);
        hi[a] = (int)fminf(ceilf((pp[a] + c.h - gd.o[a]) / gd.d[a]) + 1.0f, (float)gd.dims[a] - 1.0f);
    }
    double acc[3] = {0.0, 0.0, 0.0};
    for (int iz = lo[2]; iz <= hi[2]; iz++)
        for (int iy = lo[1]; iy <= hi[1]; iy++)
            for (int ix = lo[0]; ix <= hi[0]; ix++) {
                const uint32_t idx = (uint32_t)ix + gd.dims[0] * ((uint32_t)iy + gd.dims[1] * (uint32_t)iz);
                float g[3];
                grid_point(gd, idx, g);
                const float dx = pp[0] - g[0], dy = pp[1] - g[1], dz = pp[2] - g[2];
                const float r2 = dot3(dx, dy, dz, dx, dy, dz);
                if (r2 > c.h2) continue;
                const float r = sqrtf(r2);
                if (!(r > 1e-9f)) continue;
                const double cg = 2.0 * (rho_g[idx] - (target ? (double)target[idx] : 0.0)) / total;
                const double s = cg * (double)(c.mass * (c.h * kernel_dw(r, c)) / r);
                acc[0] += s * dx;
                acc[1] += s * dy;
                acc[2] += s * dz;
            }
    gx[i] = make_float4((float)acc[0], (float)acc[1], (float)acc[2], 0.0f);
}

__global__ void k_max_u32(const uint32_t* __restrict__ a, uint32_t n, uint32_t* __restrict__ out) {
    uint32_t m = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) m = max(m, a[i]);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = max(m, (uint32_t)__shfl_xor(m, o, 64));
    if ((threadIdx.x & 63u) == 0u) atomicMax(out, m);
}

__global__ void k_f3_to_f4(const float* __restrict__ src, float4* __restrict__ dst, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = make_float4(src[3 * (size_t)i], src[3 * (size_t)i + 1], src[3 * (size_t)i + 2], 0.0f);
}

__global__ void k_f4_to_f3(const float4* __restrict__ src, float* __restrict__ dst, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float4 v = src[i];
        dst[3 * (size_t)i] = v.x;
        dst[3 * (size_t)i + 1] = v.y;
        dst[3 * (size_t)i + 2] = v.z;
    }
}

__global__ void k_d_to_f(const double* __restrict__ src, float* __restrict__ dst, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (float)src[i];
}

__global__ void k_zero_f4(float4* __restrict__ a, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

inline uint32_t blocks_for(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

GridDesc to_desc(const float* origin, const float* spacing, const uint32_t* dims) {
    GridDesc g;
    for (int a = 0; a < 3; a++) {
        g.o[a] = origin[a];
        g.d[a] = spacing[a];
        g.dims[a] = dims[a];
    }
    return g;
}

}  // namespace

// ---- launchers (declared in adjoint_internal.h)
int launch_adjoint_substep(const uint32_t* perm, const float4* xs, const float4* vs, const float* rho,
                           const float4* bnd, const uint32_t* nbr, const uint32_t* cnt, const float4* gx_in,
                           const float4* gv_in, float4* abar, float4* gx, float4* gv, float* rbar, float4* gpsi,
                           uint32_t* dup, float4* gx_out, float4* gv_out, double* nu_part, double* nu_bar,
                           uint32_t n, const SimConsts& c, const DevObject* objs, hipStream_t s) {
    if (n == 0) return 0;
    const TapeView t{perm, xs, vs, rho, bnd, nbr, cnt};
    const AdjBuffers a{gx_in, gv_in, abar, gx, gv, rbar, gpsi, dup, gx_out, gv_out, nu_part, n};
    const uint32_t nb = blocks_for(n, 256);
    hipLaunchKernelGGL(k_adj_integrate, dim3(nb), dim3(256), 0, s, t, a, c);
    VFS_TRY(check_launch("adj_integrate"));
    hipLaunchKernelGGL(k_adj_forces, dim3(nb), dim3(256), 0, s, t, a, c);
    VFS_TRY(check_launch("adj_forces"));
    hipLaunchKernelGGL(k_add_partials, dim3(1), dim3(256), 0, s, nu_part, nb, nu_bar);
    VFS_TRY(check_launch("add_partials"));
    hipLaunchKernelGGL(k_adj_density, dim3(nb), dim3(256), 0, s, t, a, c, objs);
    return check_launch("adj_density");
}

int launch_grid_density(const float* origin, const float* spacing, const uint32_t* dims, const float4* xs,
                        const uint32_t* cell_start, const SimConsts& c, double* out, hipStream_t s) {
    const GridDesc g = to_desc(origin, spacing, dims);
    const uint64_t total = (uint64_t)dims[0] * dims[1] * dims[2];
    if (total == 0) return 0;
    hipLaunchKernelGGL(k_grid_density, dim3(blocks_for(total, 256)), dim3(256), 0, s, g, xs, cell_start, c, out);
    return check_launch("grid_density");
}

int launch_grid_loss_grad(const float* origin, const float* spacing, const uint32_t* dims, const float4* x,
                          uint32_t n, const double* rho_g, const float* target, const SimConsts& c, float4* gx,
                          hipStream_t s) {
    if (n == 0) return 0;
    const GridDesc g = to_desc(origin, spacing, dims);
    hipLaunchKernelGGL(k_grid_loss_grad, dim3(blocks_for(n, 256)), dim3(256), 0, s, g, x, n, rho_g, target, c, gx);
    return check_launch("grid_loss_grad");
}

int launch_max_u32(const uint32_t* a, uint32_t n, uint32_t* out, hipStream_t s) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_max_u32, dim3(std::min<uint32_t>(blocks_for(n, 256), 1024u)), dim3(256), 0, s, a, n, out);
    return check_launch("max_u32");
}

int launch_f3_to_f4(const float* src, float4* dst, uint32_t n, hipStream_t s) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_f3_to_f4, dim3(blocks_for(n, 256)), dim3(256), 0, s, src, dst, n);
    return check_launch("f3_to_f4");
}

int launch_f4_to_f3(const float4* src, float* dst, uint32_t n, hipStream_t s) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_f4_to_f3, dim3(blocks_for(n, 256)), dim3(256), 0, s, src, dst, n);
    return check_launch("f4_to_f3");
}

int launch_d_to_f(const double* src, float* dst, uint32_t n, hipStream_t s) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_d_to_f, dim3(blocks_for(n, 256)), dim3(256), 0, s, src, dst, n);
    return check_launch("d_to_f");
}

int launch_zero_f4(float4* a, uint32_t n, hipStream_t s) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_zero_f4, dim3(blocks_for(n, 256)), dim3(256), 0, s, a, n);
    return check_launch("zero_f4");
}

}  // namespace vfs
