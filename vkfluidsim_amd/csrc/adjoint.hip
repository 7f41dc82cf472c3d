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
#include <cmath>

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

// Constants of the reverse pair terms in double, from the model's float
// parameters (the pair sums cancel strongly, so their terms are evaluated in
// double from the float tape; see AdjConsts::make).
struct AdjConsts {
    double h, inv_h, h2, sigma, grad_sigma, mass, nu10, visc_eps_h2, rho0, B, gamma, dt, kwall_neg;
    int32_t gamma_int;
};

__device__ __forceinline__ double w_d(double r, const AdjConsts& k) {
    const double q = r * k.inv_h;
    if (q <= 1.0) {
        if (q > 0.5) {
            const double t = 1.0 - q;
            return (k.sigma * 2.0) * ((t * t) * t);
        }
        const double q2 = q * q;
        const double q3 = q2 * q;
        return k.sigma * ((6.0 * q3 - 6.0 * q2) + 1.0);
    }
    return 0.0;
}

__device__ __forceinline__ double dw_d(double r, const AdjConsts& k) {
    const double q = r * k.inv_h;
    if ((r > (double)1e-9f) && (q <= 1.0)) {
        if (q > 0.5) {
            const double t = 1.0 - q;
            return (-k.grad_sigma) * (t * t);
        }
        return (k.grad_sigma * q) * (3.0 * q - 2.0);
    }
    return 0.0;
}

__device__ __forceinline__ double ddw_d(double r, const AdjConsts& k) {
    const double q = r * k.inv_h;
    if ((r > (double)1e-9f) && (q <= 1.0)) {
        if (q > 0.5) return ((2.0 * k.grad_sigma) * (1.0 - q)) * k.inv_h;
        return (k.grad_sigma * (6.0 * q - 2.0)) * k.inv_h;
    }
    return 0.0;
}

// (rho / rho0)^e in double: binary powering for an integral exponent.
__device__ __forceinline__ double pow_ratio_d(double rho, int32_t e_int, double e, const AdjConsts& k) {
    const double x = rho / k.rho0;
    if (e_int >= 0) {
        unsigned n = (unsigned)e_int;
        double base = x, r = 1.0;
        while (n) {
            if (n & 1u) r *= base;
            base *= base;
            n >>= 1;
        }
        return r;
    }
    return pow(x, e);
}

__device__ __forceinline__ double eos_d(double rho, const AdjConsts& k) {
    return k.B * (pow_ratio_d(fmax(rho, k.rho0), k.gamma_int, k.gamma, k) - 1.0);
}

__device__ __forceinline__ double eos_dp_d(double rho, const AdjConsts& k) {
    if (!(rho > k.rho0)) return 0.0;
    return k.B * k.gamma * pow_ratio_d(rho, k.gamma_int >= 1 ? k.gamma_int - 1 : -1, k.gamma - 1.0, k) / k.rho0;
}

// p / rho^2 and m / max(rho, 1e-6) of a particle, in double from its float density.
__device__ __forceinline__ double pd_d(float rho, const AdjConsts& k) {
    const double r = rho;
    return rho > 0.0f ? eos_d(r, k) / (r * r) : 0.0;
}
__device__ __forceinline__ double md_d(float rho, const AdjConsts& k) {
    return k.mass / fmax((double)rho, (double)1e-6f);
}

__device__ __forceinline__ double dot3d(double ax, double ay, double az, double bx, double by, double bz) {
    return fma(az, bz, fma(ay, by, ax * bx));
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

// SampleGrid's cell, value and d(value)/d(uvw) in double (base cell clamped
// to [0, n-2] as in the forward), on float grids.
__device__ __forceinline__ uint32_t grid_cell_d(const uint32_t* res, const double* uvw, double* chi) {
    uint32_t ci[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const double x = ((double)res[a] - 1.0) * uvw[a];
        const double f = floor(x);
        const uint32_t hi = res[a] >= 2u ? res[a] - 2u : 0u;
        const uint32_t cc = f >= 0.0 ? (f <= (double)hi ? (uint32_t)f : hi) : 0u;
        ci[a] = cc;
        chi[a] = 2.0 * (x - (double)cc) - 1.0;
    }
    return ci[0] + res[0] * ci[1] + (res[0] * res[1]) * ci[2];
}

__device__ __forceinline__ void grid_corners_d(const float* __restrict__ grid, const uint32_t* res, uint32_t b0,
                                               double* g) {
    float f[8];
    grid_corners(grid, res, b0, f);
#pragma unroll
    for (int q = 0; q < 8; q++) g[q] = f[q];
}

// value and gradient with respect to uvw; corner q has offsets
// (ox, oy, oz) = (q >> 2, q & 1, (q >> 1) & 1)
__device__ __forceinline__ double grid_value_grad_d(const double* g, const double* chi, const uint32_t* res,
                                                    double* out) {
    const double f0[3] = {1.0 - chi[0], 1.0 - chi[1], 1.0 - chi[2]};
    const double f1[3] = {1.0 + chi[0], 1.0 + chi[1], 1.0 + chi[2]};
    double v = 0.0, dx = 0.0, dy = 0.0, dz = 0.0;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int ox = q >> 2, oy = q & 1, oz = (q >> 1) & 1;
        const double fx = ox ? f1[0] : f0[0], fy = oy ? f1[1] : f0[1], fz = oz ? f1[2] : f0[2];
        const double sx = ox ? 1.0 : -1.0, sy = oy ? 1.0 : -1.0, sz = oz ? 1.0 : -1.0;
        v += g[q] * ((fx * fy) * fz);
        dx += g[q] * (sx * fy * fz);
        dy += g[q] * (fx * sy * fz);
        dz += g[q] * (fx * fy * sz);
    }
    out[0] = dx * (2.0 * ((double)res[0] - 1.0)) / 8.0;
    out[1] = dy * (2.0 * ((double)res[1] - 1.0)) / 8.0;
    out[2] = dz * (2.0 * ((double)res[2] - 1.0)) / 8.0;
    return v / 8.0;
}

// The SDF value and gradient (d/d object-space position) at q, in double.
__device__ __forceinline__ double sdf_value_grad_d(const DevObject& o, const double* q, double* grad) {
    double u[3], ch[3], g8[8];
#pragma unroll
    for (int a = 0; a < 3; a++) u[a] = (q[a] - (double)o.box_pos[a]) / (double)o.box_size[a];
    const uint32_t cq = grid_cell_d(o.res, u, ch);
    grid_corners_d(o.sdf, o.res, cq, g8);
    const double v = grid_value_grad_d(g8, ch, o.res, grad);
#pragma unroll
    for (int a = 0; a < 3; a++) grad[a] /= (double)o.box_size[a];
    return v;
}

// Adjoint of CalculateBoundaryVolume (boundary_volume) at `pos`: adds to gpos
// the derivative of psi_bar * psi + grad_bar . grad psi with respect to pos,
// on the branch the forward took (grid cells held fixed), in double (the
// normal's Jacobian is a difference of nearby grid gradients).
__device__ void boundary_volume_adjoint(const DevObject* __restrict__ objs, uint32_t n_obj, const float* pos_f,
                                        const AdjConsts& k, double psi_bar, const double* grad_bar, double* gpos) {
    const double pos[3] = {pos_f[0], pos_f[1], pos_f[2]};
    for (uint32_t oi = 0; oi < n_obj; oi++) {
        const DevObject& o = objs[oi];
        double lp[3];
#pragma unroll
        for (int r = 0; r < 3; r++)
            lp[r] = (((double)o.T[r] * pos[0] + (double)o.T[4 + r] * pos[1]) + (double)o.T[8 + r] * pos[2]) +
                    (double)o.T[12 + r];
        bool outside = false;
#pragma unroll
        for (int a = 0; a < 3; a++)
            outside |= (lp[a] < (double)o.box_pos[a]) || (lp[a] > (double)o.box_pos[a] + (double)o.box_size[a]);
        if (outside) continue;
        double uvw[3], chi[3], gs[8], gv[8], gS[3], gV[3];
#pragma unroll
        for (int a = 0; a < 3; a++) uvw[a] = (lp[a] - (double)o.box_pos[a]) / (double)o.box_size[a];
        const uint32_t cell = grid_cell_d(o.res, uvw, chi);
        grid_corners_d(o.sdf, o.res, cell, gs);
        const double dist = grid_value_grad_d(gs, chi, o.res, gS);
        if (!(dist >= 0.0 && dist < k.h)) continue;
        grid_corners_d(o.vm, o.res, cell, gv);
        const double volume = grid_value_grad_d(gv, chi, o.res, gV);
        if (!(volume > 0.0)) continue;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            gS[a] /= (double)o.box_size[a];
            gV[a] /= (double)o.box_size[a];
        }
        // central-difference normal and its Jacobian Jn[a][b] = d nrm_a / d lp_b
        const double eps = (double)0.1f * k.h;
        double nrm[3], Jn[3][3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            double qp[3] = {lp[0], lp[1], lp[2]}, qm[3] = {lp[0], lp[1], lp[2]};
            qp[a] = lp[a] + eps;
            qm[a] = lp[a] - eps;
            double gp[3], gm[3];
            const double vp = sdf_value_grad_d(o, qp, gp), vm = sdf_value_grad_d(o, qm, gm);
            nrm[a] = (vp - vm) / (2.0 * eps);
#pragma unroll
            for (int b2 = 0; b2 < 3; b2++) Jn[a][b2] = (gp[b2] - gm[b2]) / (2.0 * eps);
        }
        const double modn = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
        if (!(modn > (double)1e-9f)) continue;
        // rn = (R nrm + t) / modn
        double Rn[3], RJ[3][3], nJ[3];
#pragma unroll
        for (int r = 0; r < 3; r++) {
            Rn[r] = (double)o.Rm[r] * nrm[0] + (double)o.Rm[4 + r] * nrm[1] + (double)o.Rm[8 + r] * nrm[2] +
                    (double)o.Rm[12 + r];
#pragma unroll
            for (int b2 = 0; b2 < 3; b2++)
                RJ[r][b2] = (double)o.Rm[r] * Jn[0][b2] + (double)o.Rm[4 + r] * Jn[1][b2] + (double)o.Rm[8 + r] * Jn[2][b2];
        }
#pragma unroll
        for (int b2 = 0; b2 < 3; b2++) nJ[b2] = nrm[0] * Jn[0][b2] + nrm[1] * Jn[1][b2] + nrm[2] * Jn[2][b2];
        double rn[3], Jr[3][3];
#pragma unroll
        for (int r = 0; r < 3; r++) {
            rn[r] = Rn[r] / modn;
#pragma unroll
            for (int b2 = 0; b2 < 3; b2++) Jr[r][b2] = RJ[r][b2] / modn - Rn[r] * nJ[b2] / (modn * modn * modn);
        }
        const double dd = dist + (double)0.125f * k.h;
        const double d = fmax(dd, (double)0.5f * k.h);
        const bool d_follows = dd > (double)0.5f * k.h;
        // xij = d rn, Jx = rn gd^T + d Jr
        double xij[3], Jx[3][3];
#pragma unroll
        for (int r = 0; r < 3; r++) {
            xij[r] = d * rn[r];
#pragma unroll
            for (int b2 = 0; b2 < 3; b2++) Jx[r][b2] = (d_follows ? rn[r] * gS[b2] : 0.0) + d * Jr[r][b2];
        }
        const double xm = sqrt(xij[0] * xij[0] + xij[1] * xij[1] + xij[2] * xij[2]);
        if (!(xm > 0.0)) continue;
        const double xh[3] = {xij[0] / xm, xij[1] / xm, xij[2] / xm};
        double dxm[3];  // x^T Jx
#pragma unroll
        for (int b2 = 0; b2 < 3; b2++) dxm[b2] = xh[0] * Jx[0][b2] + xh[1] * Jx[1][b2] + xh[2] * Jx[2][b2];
        const double W = w_d(xm, k), G = dw_d(xm, k), Gp = ddw_d(xm, k);
        const double Wp = k.h * G;
        const double xg = xh[0] * grad_bar[0] + xh[1] * grad_bar[1] + xh[2] * grad_bar[2];
        double perp[3];  // (I - x x^T) grad_bar / xm
#pragma unroll
        for (int r = 0; r < 3; r++) perp[r] = (grad_bar[r] - xg * xh[r]) / xm;
        double glp[3];
#pragma unroll
        for (int b2 = 0; b2 < 3; b2++) {
            const double jp = Jx[0][b2] * perp[0] + Jx[1][b2] * perp[1] + Jx[2][b2] * perp[2];
            glp[b2] = psi_bar * (W * gV[b2] + volume * Wp * dxm[b2]) + xg * (G * gV[b2] + volume * Gp * dxm[b2]) +
                      volume * G * jp;
        }
        // lp = T pos + t
#pragma unroll
        for (int b2 = 0; b2 < 3; b2++)
            gpos[b2] += (double)o.T[b2 * 4 + 0] * glp[0] + (double)o.T[b2 * 4 + 1] * glp[1] +
                        (double)o.T[b2 * 4 + 2] * glp[2];
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
    const double4* gx_in;  // adjoint of the substep's output positions (sorted order)
    const double4* gv_in;
    double4* abar;         // adjoint of the acceleration
    double4* gx;           // adjoint of the substep's sorted input positions
    double4* gv;
    double* rbar;          // adjoint of the densities
    double4* gpsi;         // adjoint of grad psi
    uint32_t* dup;         // 1: the particle's stencil keys repeat
    double4* gx_out;       // gx, gv in the order of the substep's input arrays
    double4* gv_out;
    double* nu_part;       // per-block viscosity adjoint of this substep
    uint32_t n;
};

// Symplectic Euler and wall penalty backwards; stencil-key repeats.
__global__ __launch_bounds__(256) void k_adj_integrate(TapeView t, AdjBuffers a, SimConsts c, AdjConsts k) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.n) return;
    const double4 gxo = a.gx_in[i], gvo = a.gv_in[i];
    const double vt[3] = {gvo.x + k.dt * gxo.x, gvo.y + k.dt * gxo.y, gvo.z + k.dt * gxo.z};
    const double ab[3] = {k.dt * vt[0], k.dt * vt[1], k.dt * vt[2]};
    const float4 x4 = t.xs[i];
    const float p[3] = {x4.x, x4.y, x4.z};
    double gx[3] = {gxo.x, gxo.y, gxo.z};
    float diff[3];
#pragma unroll
    for (int q = 0; q < 3; q++) diff[q] = fminf(p[q] - c.bb_lo[q], 0.0f) - fminf(c.bb_hi[q] - p[q], 0.0f);
    const bool wall = fabsf(diff[0]) > 1e-9f || fabsf(diff[1]) > 1e-9f || fabsf(diff[2]) > 1e-9f;
    if (wall) {
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const double D = (p[q] < c.bb_lo[q] ? 1.0 : 0.0) + (p[q] > c.bb_hi[q] ? 1.0 : 0.0);
            gx[q] += k.kwall_neg * D * ab[q];
        }
    }
    a.abar[i] = make_double4(ab[0], ab[1], ab[2], 0.0);
    a.gx[i] = make_double4(gx[0], gx[1], gx[2], 0.0);
    a.gv[i] = make_double4(vt[0], vt[1], vt[2], 0.0);
    uint32_t sk[27];
    stencil_keys(p[0], p[1], p[2], c, sk);
    uint32_t dup = 0;
    for (int u = 1; u < 27 && !dup; u++)
        for (int w = 0; w < u; w++) dup |= sk[u] == sk[w] ? 1u : 0u;
    a.dup[i] = dup;
}

// Pressure, viscosity and boundary pressure backwards: adjoints of x, v (pair
// terms), of the EOS terms and so of rho, of grad psi, and of the viscosity.
// The pair terms are evaluated in double from the float tape (positions,
// velocities, densities): the per-particle sums cancel strongly.
// Four lanes per particle in the two neighbour kernels: lane `sub` of a
// particle takes entries sub, sub + 4, ... of its list (entry sub of every
// chunk, so the four lanes read a chunk as one 16-byte run), and the partial
// sums are added over the four lanes (xor 1, then xor 2: every lane ends with
// the same total).  A substep of 65,536 particles is otherwise one wave per
// SIMD, each walking its lists alone.
constexpr uint32_t kAdjLanes = 4;

__device__ __forceinline__ double quad_sum(double v) {
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    return v;
}

__global__ __launch_bounds__(256) void k_adj_forces(TapeView t, AdjBuffers a, SimConsts c, AdjConsts k) {
    __shared__ double s_nu[256 / 64];
    const uint32_t gtid = blockIdx.x * 256u + threadIdx.x;
    const uint32_t i = gtid / kAdjLanes, sub = gtid % kAdjLanes;
    double nu_local = 0.0;
    double Pbar = 0.0, Mbar = 0.0, gx[3] = {0.0, 0.0, 0.0}, gv[3] = {0.0, 0.0, 0.0};
    double fv[3] = {0.0, 0.0, 0.0};
    if (i < a.n) {  // the four lanes of a particle are active together
        const float4 xi = t.xs[i], vi = t.vs[i];
        const float rhoi = t.rho[i];
        const double Pi = pd_d(rhoi, k), Mi = md_d(rhoi, k);
        const double4 abi = a.abar[i];
        const bool dup_i = a.dup[i] != 0u;
        const double ai[3] = {abi.x, abi.y, abi.z};
        const uint32_t count = t.cnt[i];
        const uint32_t* list = t.nbr;
        for (uint32_t e = sub; e < count; e += kAdjLanes) {
            const uint32_t j = list[(size_t)(e >> 2) * 4u * a.n + 4u * i + (e & 3u)];
            if (j == i) continue;
            const float4 xj = t.xs[j], vj = t.vs[j];
            const double xij[3] = {(double)xi.x - xj.x, (double)xi.y - xj.y, (double)xi.z - xj.z};
            const double r2 = dot3d(xij[0], xij[1], xij[2], xij[0], xij[1], xij[2]);
            const double r = sqrt(r2);
            if (!(r > (double)1e-9f)) continue;
            const double inv_r = 1.0 / r;
            const double nh[3] = {xij[0] * inv_r, xij[1] * inv_r, xij[2] * inv_r};
            const double G = dw_d(r, k), Gp = ddw_d(r, k);
            const double F[3] = {nh[0] * G, nh[1] * G, nh[2] * G};
            const double vij[3] = {(double)vi.x - vj.x, (double)vi.y - vj.y, (double)vi.z - vj.z};
            const double sig = dot3d(vij[0], vij[1], vij[2], xij[0], xij[1], xij[2]);
            const double inv_den = 1.0 / (r2 + k.visc_eps_h2);
            const float rhoj = t.rho[j];
            const double Pj = pd_d(rhoj, k), Mj = md_d(rhoj, k);
            const double4 abj4 = a.abar[j];
            const double aj[3] = {abj4.x, abj4.y, abj4.z};
            const double w = pair_weight(xi, xj, dup_i, a.dup[j] != 0u, c);
            // forward viscous sum (for the viscosity adjoint)
#pragma unroll
            for (int q = 0; q < 3; q++) fv[q] += Mj * sig * inv_den * F[q];
            // own term of a_i: -m (Pi + Pj) F + nu10 Mj sig inv_den F
            const double aF = ai[0] * F[0] + ai[1] * F[1] + ai[2] * F[2];
            // neighbour term of a_j (its pair with i): F_ji = -F, x_ji = -x_ij, v_ji = -v_ij
            const double bF = -(aj[0] * F[0] + aj[1] * F[1] + aj[2] * F[2]);  // a_j . F_ji
            Pbar += -k.mass * aF + w * (-k.mass * bF);
            Mbar += w * (k.nu10 * sig * inv_den * bF);
            // d(a_j term)/d x_i = -d/d x_ji, folded into the own-term adjoint in
            // x_ij: F_ji(x_ji) = -F(x_ij), sig and den even
            double Fb[3];
#pragma unroll
            for (int q = 0; q < 3; q++)
                Fb[q] = (-k.mass * (Pi + Pj) + k.nu10 * Mj * sig * inv_den) * ai[q] -
                        w * (-k.mass * (Pj + Pi) + k.nu10 * Mi * sig * inv_den) * aj[q];
            const double sb = k.nu10 * Mj * inv_den * aF + w * k.nu10 * Mi * inv_den * bF;
            const double db = -k.nu10 * Mj * sig * inv_den * inv_den * aF - w * k.nu10 * Mi * sig * inv_den * inv_den * bF;
            // F = x_ij G / r:  dF/dx_ij^T Fb = (G/r) Fb + (Fb . n)(G' - G/r) n
            const double Fn = Fb[0] * nh[0] + Fb[1] * nh[1] + Fb[2] * nh[2];
            const double gr = G * inv_r;
#pragma unroll
            for (int q = 0; q < 3; q++) {
                gx[q] += gr * Fb[q] + Fn * (Gp - gr) * nh[q] + sb * vij[q] + 2.0 * db * xij[q];
                gv[q] += sb * xij[q];
            }
        }
        Pbar = quad_sum(Pbar);
        Mbar = quad_sum(Mbar);
#pragma unroll
        for (int q = 0; q < 3; q++) {
            gx[q] = quad_sum(gx[q]);
            gv[q] = quad_sum(gv[q]);
            fv[q] = quad_sum(fv[q]);
        }
    }
    if (i < a.n && sub == 0u) {
        const float rhoi = t.rho[i];
        const double Pi = pd_d(rhoi, k);
        const double4 abi = a.abar[i];
        const double ai[3] = {abi.x, abi.y, abi.z};
        // viscosity: a_i = ... + nu10 * fv, nu10 = 10 nu
        nu_local = 10.0 * (ai[0] * fv[0] + ai[1] * fv[1] + ai[2] * fv[2]);
        // boundary pressure: a_i = ... - (Pi + pb) rho0 grad psi
        const double rd = rhoi;
        const float4 bnd = t.bnd[i];
        const double p = eos_d(rd, k);
        const double pb = rhoi > 0.0f ? p / (k.rho0 * k.rho0) : 0.0;
        const double ag = ai[0] * bnd.x + ai[1] * bnd.y + ai[2] * bnd.z;
        Pbar += -k.rho0 * ag;
        const double sbnd = -(Pi + pb) * k.rho0;
        a.gpsi[i] = make_double4(sbnd * ai[0], sbnd * ai[1], sbnd * ai[2], 0.0);
        // EOS terms to rho
        const double dp = eos_dp_d(rd, k);
        double rbar = rhoi > 0.0f ? -k.rho0 * ag * dp / (k.rho0 * k.rho0) : 0.0;
        if (rhoi > 0.0f) rbar += Pbar * (dp / (rd * rd) - 2.0 * p / (rd * rd * rd));
        if (rhoi > 1e-6f) rbar += Mbar * (-k.mass / (rd * rd));
        a.rbar[i] = rbar;
        const double4 g0 = a.gx[i], v0 = a.gv[i];
        a.gx[i] = make_double4(g0.x + gx[0], g0.y + gx[1], g0.z + gx[2], 0.0);
        a.gv[i] = make_double4(v0.x + gv[0], v0.y + gv[1], v0.z + gv[2], 0.0);
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

// Boundary volume backwards (particles with boundary objects only): adds the
// adjoint of psi and grad psi through the grids' trilinear gradients into the
// position adjoint, before the density pass reads it.  A kernel of its own:
// its double grid arithmetic needs far more registers than the neighbour sums.
__global__ __launch_bounds__(256) void k_adj_boundary(TapeView t, AdjBuffers a, SimConsts c, AdjConsts k,
                                                      const DevObject* __restrict__ objs) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.n) return;
    const float4 xi = t.xs[i];
    const double4 g0 = a.gx[i], gp = a.gpsi[i];
    double gx[3] = {g0.x, g0.y, g0.z};
    const double gb[3] = {gp.x, gp.y, gp.z};
    const float pos[3] = {xi.x, xi.y, xi.z};
    boundary_volume_adjoint(objs + c.obj_begin, c.obj_count, pos, k, k.rho0 * a.rbar[i], gb, gx);
    a.gx[i] = make_double4(gx[0], gx[1], gx[2], 0.0);
}

// Density backwards (four lanes per particle, as k_adj_forces), then the
// adjoints of the substep's input in the order of its input arrays (inverse
// of the reorder).
__global__ __launch_bounds__(256) void k_adj_density(TapeView t, AdjBuffers a, SimConsts c, AdjConsts k) {
    const uint32_t gtid = blockIdx.x * 256u + threadIdx.x;
    const uint32_t i = gtid / kAdjLanes, sub = gtid % kAdjLanes;
    if (i >= a.n) return;  // the four lanes of a particle leave together
    const float4 xi = t.xs[i];
    const double rbi = a.rbar[i];
    const bool dup_i = a.dup[i] != 0u;
    double gx[3] = {0.0, 0.0, 0.0};
    const uint32_t count = t.cnt[i];
    for (uint32_t e = sub; e < count; e += kAdjLanes) {
        const uint32_t j = t.nbr[(size_t)(e >> 2) * 4u * a.n + 4u * i + (e & 3u)];
        if (j == i) continue;
        const float4 xj = t.xs[j];
        const double xij[3] = {(double)xi.x - xj.x, (double)xi.y - xj.y, (double)xi.z - xj.z};
        const double r = sqrt(dot3d(xij[0], xij[1], xij[2], xij[0], xij[1], xij[2]));
        if (!(r > (double)1e-9f)) continue;
        const double w = pair_weight(xi, xj, dup_i, a.dup[j] != 0u, c);
        const double s = k.mass * (k.h * dw_d(r, k)) * (rbi + w * a.rbar[j]) / r;
#pragma unroll
        for (int q = 0; q < 3; q++) gx[q] += s * xij[q];
    }
#pragma unroll
    for (int q = 0; q < 3; q++) gx[q] = quad_sum(gx[q]);
    if (sub != 0u) return;
    const double4 g0 = a.gx[i];
    const uint32_t o = t.perm[i];
    a.gx_out[o] = make_double4(g0.x + gx[0], g0.y + gx[1], g0.z + gx[2], 0.0);
    a.gv_out[o] = a.gv[i];
}

// Density field on grid points: rho(g) = sum over particles within h of
// m W(|g - x_j|), each particle once (repeated stencil keys are skipped).
struct GridDesc {
    float o[3], d[3];
    uint32_t dims[3];
};

// The loss evaluates the cubic spline in double from the float positions and
// grid points: it is a small difference of two density fields, which float
// kernel values would blur by ~1e-3 of its size.  Constants from the float
// smoothing radius and target density, as the model's FP64 reading.
struct LossKernel {
    double h2, inv_h, h, sigma, grad_sigma, mass;
};

__device__ __forceinline__ double loss_w(double r, const LossKernel& k) {
    const double q = r * k.inv_h;
    if (q <= 1.0) {
        if (q > 0.5) {
            const double t = 1.0 - q;
            return (k.sigma * 2.0) * ((t * t) * t);
        }
        const double q2 = q * q;
        const double q3 = q2 * q;
        return k.sigma * ((6.0 * q3 - 6.0 * q2) + 1.0);
    }
    return 0.0;
}

__device__ __forceinline__ double loss_dw(double r, const LossKernel& k) {
    const double q = r * k.inv_h;
    if ((r > (double)1e-9f) && (q <= 1.0)) {
        if (q > 0.5) {
            const double t = 1.0 - q;
            return (-k.grad_sigma) * (t * t);
        }
        return (k.grad_sigma * q) * (3.0 * q - 2.0);
    }
    return 0.0;
}

__device__ __forceinline__ double dot3d(double ax, double ay, double az) {
    return fma(az, az, fma(ay, ay, ax * ax));
}

__device__ __forceinline__ void grid_point(const GridDesc& gd, uint32_t idx, float* g) {
    const uint32_t ix = idx % gd.dims[0], iy = (idx / gd.dims[0]) % gd.dims[1], iz = idx / (gd.dims[0] * gd.dims[1]);
    g[0] = gd.o[0] + (float)ix * gd.d[0];
    g[1] = gd.o[1] + (float)iy * gd.d[1];
    g[2] = gd.o[2] + (float)iz * gd.d[2];
}

__global__ __launch_bounds__(256) void k_grid_density(GridDesc gd, const float4* __restrict__ xs,
                                                      const uint32_t* __restrict__ cs, SimConsts c, LossKernel lk,
                                                      double* __restrict__ out) {
    const uint32_t idx = blockIdx.x * 256u + threadIdx.x;
    const uint32_t total = gd.dims[0] * gd.dims[1] * gd.dims[2];
    if (idx >= total) return;
    float g[3];
    grid_point(gd, idx, g);
    uint32_t k[27];
    stencil_keys(g[0], g[1], g[2], c, k);
    double sum = 0.0;
    for (int u = 0; u < 27; u++) {
        bool seen = false;
        for (int w = 0; w < u; w++) seen |= k[w] == k[u];
        if (seen) continue;
        for (uint32_t j = cs[k[u]]; j < cs[k[u] + 1u]; j++) {
            const float4 q = xs[j];
            const double dx = (double)q.x - g[0], dy = (double)q.y - g[1], dz = (double)q.z - g[2];
            const double r2 = dot3d(dx, dy, dz);
            if (!(r2 > lk.h2)) sum += lk.mass * loss_w(sqrt(r2), lk);
        }
    }
    out[idx] = sum;
}

// d loss / d x_i for loss = mean((rho(g) - target(g))^2), per particle over the
// grid points within h (the same pairs as k_grid_density).
__global__ __launch_bounds__(256) void k_grid_loss_grad(GridDesc gd, const float4* __restrict__ x, uint32_t n,
                                                        const double* __restrict__ rho_g,
                                                        const float* __restrict__ target, SimConsts c,
                                                        LossKernel lk, double4* __restrict__ gx) {
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
                const double dx = (double)pp[0] - g[0], dy = (double)pp[1] - g[1], dz = (double)pp[2] - g[2];
                const double r2 = dot3d(dx, dy, dz);
                if (r2 > lk.h2) continue;
                const double r = sqrt(r2);
                if (!(r > 1e-9)) continue;
                const double cg = 2.0 * (rho_g[idx] - (target ? (double)target[idx] : 0.0)) / total;
                const double s = cg * lk.mass * (lk.h * loss_dw(r, lk)) / r;
                acc[0] += s * dx;
                acc[1] += s * dy;
                acc[2] += s * dz;
            }
    gx[i] = make_double4(acc[0], acc[1], acc[2], 0.0);
}

// out = max(out, max of a[0..n)): sixteen elements per thread (loads issued
// together), a block reduction, and one atomic per block (atomics on one
// address serialise, so a thousand of them per launch cost ~10 us).
constexpr uint32_t kMaxPerThread = 16;
__global__ __launch_bounds__(256) void k_max_u32(const uint32_t* __restrict__ a, uint32_t n,
                                                 uint32_t* __restrict__ out) {
    __shared__ uint32_t s_max[256 / 64];
    const uint32_t base = blockIdx.x * 256u * kMaxPerThread + threadIdx.x;
    uint32_t v[kMaxPerThread];
#pragma unroll
    for (uint32_t u = 0; u < kMaxPerThread; u++) {
        const uint32_t i = base + u * 256u;
        v[u] = i < n ? a[i] : 0u;
    }
    uint32_t m = 0;
#pragma unroll
    for (uint32_t u = 0; u < kMaxPerThread; u++) m = max(m, v[u]);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = max(m, (uint32_t)__shfl_xor(m, o, 64));
    if ((threadIdx.x & 63u) == 0u) s_max[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 256 / 64; w++) m = max(m, s_max[w]);
        atomicMax(out, m);
    }
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

__global__ void k_zero_d4(double4* __restrict__ a, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = make_double4(0.0, 0.0, 0.0, 0.0);
}

__global__ void k_d4_to_f3(const double4* __restrict__ src, float* __restrict__ dst, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double4 v = src[i];
        dst[3 * (size_t)i] = (float)v.x;
        dst[3 * (size_t)i + 1] = (float)v.y;
        dst[3 * (size_t)i + 2] = (float)v.z;
    }
}

inline uint32_t blocks_for(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

// The model's constants in double from its float parameters (the FP64
// reading: 48 / (pi h^4) / h, rho0 * 0.8 h^3 / 8, 2 (d + 2) nu, 0.01 h^2).
// The reverse pass is checked against the FP64 adjoint of the FP32
// trajectory, which uses exactly these.
AdjConsts adj_consts(const SimConsts& c) {
    AdjConsts k;
    k.h = (double)c.h;
    k.inv_h = 1.0 / k.h;
    k.h2 = k.h * k.h;
    k.sigma = 8.0 / (M_PI * std::pow(k.h, 3));
    k.grad_sigma = 48.0 / (M_PI * std::pow(k.h, 4)) / k.h;
    k.mass = (double)c.rho0 * ((double)0.8f * k.h * k.h * k.h / 8.0);
    k.dt = (double)c.dt;
    k.kwall_neg = (double)c.kwall_neg;
    k.nu10 = 10.0 * (double)c.nu;
    k.visc_eps_h2 = (double)0.01f * k.h2;
    k.rho0 = (double)c.rho0;
    k.B = (double)c.B;
    k.gamma = (double)c.gamma;
    k.gamma_int = c.gamma_int;
    return k;
}

LossKernel loss_kernel(const SimConsts& c) {
    LossKernel k;
    k.h = (double)c.h;
    k.h2 = k.h * k.h;
    k.inv_h = 1.0 / k.h;
    k.sigma = 8.0 / (M_PI * std::pow(k.h, 3));
    k.grad_sigma = 48.0 / (M_PI * std::pow(k.h, 4)) / k.h;
    k.mass = (double)c.rho0 * ((double)0.8f * k.h * k.h * k.h / 8.0);
    return k;
}

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
                           const float4* bnd, const uint32_t* nbr, const uint32_t* cnt, const double4* gx_in,
                           const double4* gv_in, double4* abar, double4* gx, double4* gv, double* rbar,
                           double4* gpsi, uint32_t* dup, double4* gx_out, double4* gv_out, double* nu_part,
                           double* nu_bar, uint32_t n, const SimConsts& c, const DevObject* objs, hipStream_t s,
                           StageTimer* timer) {
    if (n == 0) return 0;
    const TapeView t{perm, xs, vs, rho, bnd, nbr, cnt};
    const AdjBuffers a{gx_in, gv_in, abar, gx, gv, rbar, gpsi, dup, gx_out, gv_out, nu_part, n};
    const uint32_t nb = blocks_for(n, 256);
    const uint32_t nb_lanes = blocks_for((uint64_t)n * kAdjLanes, 256);  // four lanes per particle
    const AdjConsts k = adj_consts(c);
    if (timer) timer->begin(s, "adj_integrate");
    hipLaunchKernelGGL(k_adj_integrate, dim3(nb), dim3(256), 0, s, t, a, c, k);
    VFS_TRY(check_launch("adj_integrate"));
    if (timer) timer->end(s);
    if (timer) timer->begin(s, "adj_forces");
    hipLaunchKernelGGL(k_adj_forces, dim3(nb_lanes), dim3(256), 0, s, t, a, c, k);
    VFS_TRY(check_launch("adj_forces"));
    if (timer) timer->end(s);
    if (timer) timer->begin(s, "adj_viscosity_sum");
    hipLaunchKernelGGL(k_add_partials, dim3(1), dim3(256), 0, s, nu_part, nb_lanes, nu_bar);
    VFS_TRY(check_launch("add_partials"));
    if (timer) timer->end(s);
    if (c.obj_count) {
        if (timer) timer->begin(s, "adj_boundary");
        hipLaunchKernelGGL(k_adj_boundary, dim3(nb), dim3(256), 0, s, t, a, c, k, objs);
        VFS_TRY(check_launch("adj_boundary"));
        if (timer) timer->end(s);
    }
    if (timer) timer->begin(s, "adj_density");
    hipLaunchKernelGGL(k_adj_density, dim3(nb_lanes), dim3(256), 0, s, t, a, c, k);
    VFS_TRY(check_launch("adj_density"));
    if (timer) timer->end(s);
    return 0;
}

int launch_grid_density(const float* origin, const float* spacing, const uint32_t* dims, const float4* xs,
                        const uint32_t* cell_start, const SimConsts& c, double* out, hipStream_t s) {
    const GridDesc g = to_desc(origin, spacing, dims);
    const uint64_t total = (uint64_t)dims[0] * dims[1] * dims[2];
    if (total == 0) return 0;
    hipLaunchKernelGGL(k_grid_density, dim3(blocks_for(total, 256)), dim3(256), 0, s, g, xs, cell_start, c,
                       loss_kernel(c), out);
    return check_launch("grid_density");
}

int launch_grid_loss_grad(const float* origin, const float* spacing, const uint32_t* dims, const float4* x,
                          uint32_t n, const double* rho_g, const float* target, const SimConsts& c, double4* gx,
                          hipStream_t s) {
    if (n == 0) return 0;
    const GridDesc g = to_desc(origin, spacing, dims);
    hipLaunchKernelGGL(k_grid_loss_grad, dim3(blocks_for(n, 256)), dim3(256), 0, s, g, x, n, rho_g, target, c,
                       loss_kernel(c), gx);
    return check_launch("grid_loss_grad");
}

int launch_max_u32(const uint32_t* a, uint32_t n, uint32_t* out, hipStream_t s) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_max_u32, dim3(blocks_for(n, 256u * kMaxPerThread)), dim3(256), 0, s, a, n, out);
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

int launch_zero_d4(double4* a, uint32_t n, hipStream_t s) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_zero_d4, dim3(blocks_for(n, 256)), dim3(256), 0, s, a, n);
    return check_launch("zero_d4");
}

int launch_d4_to_f3(const double4* src, float* dst, uint32_t n, hipStream_t s) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_d4_to_f3, dim3(blocks_for(n, 256)), dim3(256), 0, s, src, dst, n);
    return check_launch("d4_to_f3");
}

}  // namespace vfs
