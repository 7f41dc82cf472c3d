// TEST INFRASTRUCTURE ONLY -- the checker, never the product.
//
// CPU restatement of the reference WCSPH substep (luihabl/VkFluidSim), written
// operation by operation from the shader text so that an FP32 instantiation
// (built with -ffp-contract=off) is the parity oracle for the HIP kernels and an
// FP64 instantiation measures FP32 drift.  Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg load this library.
//
// Canonical choices where the reference leaves bits open (documented in the
// design notes as "parity unpinned" against the shaders themselves):
//   * pow(t,3) = (t*t)*t, pow(t,2) = t*t, pow(h,3) = (h*h)*h;
//   * EOS power: integral exponent in [0,64] -> binary powering in double,
//     rounded once to the working type; otherwise std::pow in double;
//   * dot(a,b) = fma(az, bz, fma(ay, by, ax*bx)) (fused, as GPU shader
//     compilers evaluate dot); M*v sums columns left to right, unfused;
//   * a vector divided by a scalar (xij / |xij|, the viscous term over
//     r^2 + 0.01h^2, the boundary normal, the SDF gradient) is multiplied by
//     the IEEE reciprocal of the scalar, and so is r / h in the smoothing
//     kernels (h is uniform); other scalar quotients are IEEE divisions
//     (SPIR-V OpFDiv itself is only specified to 2.5 ulp under Vulkan);
//   * the 1e-9 / 0.01 / 0.125 / 0.8 literals are taken as float literals;
//   * float->int cell conversion saturates and maps NaN to 0;
//   * SampleGrid clamps the base cell to [0, n-2] (the reference reads out of
//     bounds at uvw == 1 and outside the grid);
//   * the counting sort is stable (ascending current index inside a key); the
//     reference scatter order inside a bucket is nondeterministic.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ------------------------------------------------------------------ hashing
// spatial_hash_3d.slang:3-10
static const int kOffsets3D[27][3] = {
    {-1, -1, -1}, {0, -1, -1}, {1, -1, -1}, {-1, 0, -1}, {0, 0, -1}, {1, 0, -1}, {-1, 1, -1},
    {0, 1, -1},   {1, 1, -1},  {-1, -1, 0}, {0, -1, 0},  {1, -1, 0}, {-1, 0, 0}, {0, 0, 0},
    {1, 0, 0},    {-1, 1, 0},  {0, 1, 0},   {1, 1, 0},   {-1, -1, 1}, {0, -1, 1}, {1, -1, 1},
    {-1, 0, 1},   {0, 0, 1},   {1, 0, 1},   {-1, 1, 1},  {0, 1, 1},  {1, 1, 1}};

// HashCell3D, spatial_hash_3d.slang:20-28.  (uint3)(cell + 25) wraps mod 2^32.
static inline uint32_t hash_cell(int32_t cx, int32_t cy, int32_t cz) {
    const uint32_t B = 50;
    const uint32_t ux = (uint32_t)cx + 25u, uy = (uint32_t)cy + 25u, uz = (uint32_t)cz + 25u;
    const uint32_t lx = ux % B, ly = uy % B, lz = uz % B;
    const uint32_t bx = ux / B, by = uy / B, bz = uz / B;
    const uint32_t block_hash = bx * 15823u + by * 9737333u + bz * 440817757u;
    return lx + B * (ly + B * lz) + block_hash;
}

// (int)floor(v): GetCell3D, spatial_hash_3d.slang:16-18, with a saturating,
// NaN->0 conversion in place of the shader's undefined out-of-range cast.
template <class R>
static inline int32_t floor_to_cell(R v) {
    R f = std::floor(v);
    if (!(f == f)) return 0;
    if (f <= R(-2147483648.0)) return std::numeric_limits<int32_t>::min();
    if (f >= R(2147483647.0)) return std::numeric_limits<int32_t>::max();
    return (int32_t)f;
}

template <class R>
static inline void get_cell(const R* p, R cell_size, int32_t c[3]) {
    c[0] = floor_to_cell<R>(p[0] / cell_size);
    c[1] = floor_to_cell<R>(p[1] / cell_size);
    c[2] = floor_to_cell<R>(p[2] / cell_size);
}

// ------------------------------------------------------------- parameters
struct ParamsC {  // identical layout to vfs_params (include/vfs_c.h)
    float gravity[3];
    float smooth_radius;
    float time_scale;
    int32_t iterations;
    int32_t n_particles;
    float fixed_dt;
    float target_density;
    float bbox_size[3];
    float bbox_pos[3];
    float wall_stiffness;
    float stiffness;
    float expoent;
    float viscosity_strenght;
    float cell_size;
    uint32_t flags;
};
static_assert(sizeof(ParamsC) == 84, "vfs_params layout");

struct ObjectC {  // identical layout to vfs_boundary_object
    float transform[16];
    float rotation[16];
    float box_size[3];
    float box_pos[3];
    uint32_t resolution[3];
    uint32_t reserved;
    const float* sdf_grid;
    const float* volume_map_grid;
};
static_assert(sizeof(ObjectC) == 184, "vfs_boundary_object layout");

// Canonical power used by the EOS (see header comment).
template <class R>
static inline R eos_pow(R x, R e);
template <>
inline float eos_pow<float>(float x, float e) {
    if (e >= 0.0f && e <= 64.0f && std::floor(e) == e) {
        unsigned n = (unsigned)e;
        double base = (double)x, result = 1.0;
        while (n) {
            if (n & 1u) result *= base;
            base *= base;
            n >>= 1;
        }
        return (float)result;
    }
    return (float)std::pow((double)x, (double)e);
}
template <>
inline double eos_pow<double>(double x, double e) {
    return std::pow(x, e);
}

template <class R>
struct Model {
    // derived uniforms
    R g[3], h, h2, inv_h, cell, rho0, B, gamma, nu, kwall, dt, mass;
    // the adjoint of an FP32 trajectory takes the forward's cells: float
    // division of the (float) positions by the cell size
    bool float_cells = false;
    R sigma;       // cubic_spline_scale (model.cpp:231)
    R grad_sigma;  // grad_cubic_spline_scale / smooth_radius (kernels_3d.slang:80-81)
    R bb_lo[3], bb_hi[3];
    uint32_t n;

    struct Obj {
        R T[16], Rm[16], box_pos[3], box_size[3];
        uint32_t res[3];
        std::vector<float> sdf, vm;
    };
    std::vector<Obj> objs;

    void set_params(const ParamsC& p);

    // kernels_3d.slang:62-76
    R W(R r) const {
        const R q = r * inv_h;
        if (q <= R(1)) {
            if (q > R(0.5)) {
                const R t = R(1) - q;
                return (sigma * R(2)) * ((t * t) * t);
            }
            const R q2 = q * q;
            const R q3 = q2 * q;
            return sigma * ((R(6) * q3 - R(6) * q2) + R(1));
        }
        return R(0);
    }
    // kernels_3d.slang:78-91 (note the extra 1/h in grad_sigma)
    R dW(R r) const {
        const R q = r * inv_h;
        if ((r > R(1e-9f)) && (q <= R(1))) {
            if (q > R(0.5)) {
                const R t = R(1) - q;
                return (-grad_sigma) * (t * t);
            }
            return (grad_sigma * q) * (R(3) * q - R(2));
        }
        return R(0);
    }
    // wcsph_with_boundary_model.slang:199-203
    R eos(R rho) const {
        rho = std::fmax(rho, rho0);
        return B * (eos_pow<R>(rho / rho0, gamma) - R(1));
    }
};

template <class R>
static inline R kernel_scale(double num, float h) {
    return R(num) / (R(3.14159265358979323846) * R(h) * R(h) * R(h));
}

template <>
void Model<float>::set_params(const ParamsC& p) {
    for (int a = 0; a < 3; a++) g[a] = p.gravity[a];
    h = p.smooth_radius;
    h2 = h * h;
    inv_h = 1.0f / h;
    cell = p.cell_size > 0.0f ? p.cell_size : p.smooth_radius;
    rho0 = p.target_density;
    B = p.stiffness;
    gamma = p.expoent;
    nu = p.viscosity_strenght;
    kwall = p.wall_stiffness;
    dt = p.time_scale * p.fixed_dt / (float)p.iterations;  // wcsph_with_boundary_model.cpp:84-89
    // ParticleVolume (slang:39-41) times target_density (slang:164)
    const float h3 = (h * h) * h;
    mass = rho0 * ((0.8f * h3) / 8.0f);
    // CalcKernelCoefficients, model.cpp:225-233: glm::pi<float>() and (float)std::pow(double, int)
    const float pi_f = 3.14159265358979323846f;
    sigma = 8.0f / (pi_f * (float)std::pow((double)h, 3));
    const float grad_scale = 48.0f / (pi_f * (float)std::pow((double)h, 4));
    grad_sigma = grad_scale / h;
    for (int a = 0; a < 3; a++) {
        bb_lo[a] = p.bbox_pos[a];
        bb_hi[a] = p.bbox_pos[a] + p.bbox_size[a];
    }
    n = (uint32_t)p.n_particles;
}

template <>
void Model<double>::set_params(const ParamsC& p) {
    for (int a = 0; a < 3; a++) g[a] = p.gravity[a];
    h = p.smooth_radius;
    h2 = h * h;
    inv_h = 1.0 / h;
    cell = p.cell_size > 0.0f ? p.cell_size : p.smooth_radius;
    rho0 = p.target_density;
    B = p.stiffness;
    gamma = p.expoent;
    nu = p.viscosity_strenght;
    kwall = p.wall_stiffness;
    dt = (double)p.time_scale * (double)p.fixed_dt / (double)p.iterations;
    mass = rho0 * ((double)0.8f * h * h * h / 8.0);
    sigma = 8.0 / (M_PI * std::pow(h, 3));
    grad_sigma = 48.0 / (M_PI * std::pow(h, 4)) / h;
    for (int a = 0; a < 3; a++) {
        bb_lo[a] = p.bbox_pos[a];
        bb_hi[a] = (double)p.bbox_pos[a] + (double)p.bbox_size[a];
    }
    n = (uint32_t)p.n_particles;
}

// ------------------------------------------------------------ grid sampling
template <class R>
static inline R dot3(const R* a, const R* b) {
    return std::fma(a[2], b[2], std::fma(a[1], b[1], a[0] * b[0]));
}

// SampleGrid, wcsph_with_boundary_model.slang:51-66 (base cell clamped to [0, n-2])
template <class R>
static R sample_grid(const float* grid, const uint32_t* n, const R* uvw) {
    uint32_t c[3];
    R chi[3];
    for (int a = 0; a < 3; a++) {
        const R x = (R(n[a]) - R(1)) * uvw[a];
        const R f = std::floor(x);
        const uint32_t hi = n[a] >= 2 ? n[a] - 2 : 0;
        uint32_t ci;
        if (f >= R(0))
            ci = f <= R(hi) ? (uint32_t)f : hi;
        else
            ci = 0;
        c[a] = ci;
        chi[a] = R(2) * (x - R(ci)) - R(1);
    }
    static const uint32_t off[8][3] = {{0, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1},
                                       {1, 0, 0}, {1, 1, 0}, {1, 0, 1}, {1, 1, 1}};
    R val = R(0);
    for (int i = 0; i < 8; i++) {
        R f[3];
        for (int a = 0; a < 3; a++) f[a] = R(1) + (off[i][a] ? R(1) : R(-1)) * chi[a];
        const uint32_t ix = c[0] + off[i][0], iy = c[1] + off[i][1], iz = c[2] + off[i][2];
        const uint32_t idx = ix + n[0] * iy + n[0] * n[1] * iz;  // GetIndex3D, slang:47-49
        val += R(grid[idx]) * ((f[0] * f[1]) * f[2]);
    }
    return val / R(8);
}

template <class R>
static inline void to_uvw(const R* box_pos, const R* box_size, const R* p, R* uvw) {
    for (int a = 0; a < 3; a++) uvw[a] = (p[a] - box_pos[a]) / box_size[a];  // slang:43-45
}

// GradientVector, slang:68-85
template <class R>
static void gradient_vector(const typename Model<R>::Obj& o, R h, const R* x, R* out) {
    const R eps = R(0.1f) * h;
    R v1[3], v0[3];
    for (int a = 0; a < 3; a++) {
        R xp[3] = {x[0], x[1], x[2]}, xm[3] = {x[0], x[1], x[2]};
        xp[a] = x[a] + eps;
        xm[a] = x[a] - eps;
        // the shader adds the zero components too: x + 0 == x
        R uvw[3];
        to_uvw<R>(o.box_pos, o.box_size, xp, uvw);
        v1[a] = sample_grid<R>(o.sdf.data(), o.res, uvw);
        to_uvw<R>(o.box_pos, o.box_size, xm, uvw);
        v0[a] = sample_grid<R>(o.sdf.data(), o.res, uvw);
    }
    const R inv = R(1) / (R(2) * eps);
    for (int a = 0; a < 3; a++) out[a] = (v1[a] - v0[a]) * inv;
}

// mul(M, float4(v, w)).xyz for a column-major 4x4 matrix.
template <class R>
static inline void mat_vec(const R* M, const R* v, R w, R* out) {
    for (int r = 0; r < 3; r++)
        out[r] = ((M[0 * 4 + r] * v[0] + M[1 * 4 + r] * v[1]) + M[2 * 4 + r] * v[2]) + M[3 * 4 + r] * w;
}

// CalculateBoundaryVolume for one particle, slang:102-151
template <class R>
static void boundary_volume(const Model<R>& m, const R* pos, R& psi, R* grad) {
    psi = R(0);
    grad[0] = grad[1] = grad[2] = R(0);
    for (const auto& o : m.objs) {
        R lp[3];
        mat_vec<R>(o.T, pos, R(1), lp);
        bool outside = false;
        for (int a = 0; a < 3; a++)
            if (lp[a] < o.box_pos[a] || lp[a] > o.box_pos[a] + o.box_size[a]) outside = true;
        if (outside) continue;
        R uvw[3];
        to_uvw<R>(o.box_pos, o.box_size, lp, uvw);
        const R dist = sample_grid<R>(o.sdf.data(), o.res, uvw);
        if (!(dist >= R(0) && dist < m.h)) continue;
        const R volume = sample_grid<R>(o.vm.data(), o.res, uvw);
        if (!(volume > R(0))) continue;
        R nrm[3];
        gradient_vector<R>(o, m.h, lp, nrm);
        const R modn = std::sqrt(dot3<R>(nrm, nrm));
        if (!(modn > R(1e-9f))) continue;
        R rn[3];
        mat_vec<R>(o.Rm, nrm, R(1), rn);
        const R inv_modn = R(1) / modn;
        for (int a = 0; a < 3; a++) rn[a] = rn[a] * inv_modn;
        const R d = std::fmax(dist + R(0.125f) * m.h, R(0.5f) * m.h);
        R xij[3];
        for (int a = 0; a < 3; a++) {
            const R xb = pos[a] - d * rn[a];
            xij[a] = pos[a] - xb;
        }
        const R xm = std::sqrt(dot3<R>(xij, xij));
        R xn[3];
        const R inv_xm = xm > R(0) ? R(1) / xm : R(0);
        for (int a = 0; a < 3; a++) xn[a] = xm > R(0) ? xij[a] * inv_xm : R(0);
        psi += volume * m.W(xm);
        const R vg = volume * m.dW(xm);
        for (int a = 0; a < 3; a++) grad[a] += vg * xn[a];
    }
}

// ---------------------------------------------------------------- simulation
template <class R>
struct Sim {
    Model<R> m;
    uint32_t n = 0;
    std::vector<R> x, v, a, rho, psi, bgrad;
    std::vector<uint32_t> pid, keys, sorted_keys, perm, offsets, counts;
    std::vector<R> tx, tv;
    std::vector<uint32_t> tpid;
    int threads = 1;

    // Per-substep record for the host adjoint: the sorted input state, the
    // density-loop neighbours of every particle (self and repeats included, in
    // loop order), densities, boundary terms and the permutation.
    struct TapeStep {
        std::vector<uint32_t> perm;
        std::vector<R> x, v, rho, psi, bgrad;
        std::vector<std::vector<uint32_t>> nbr;
    };
    std::vector<TapeStep>* tape = nullptr;
    std::vector<std::vector<uint32_t>> lists;

    void hash_and_sort() {
        // UpdateSpatialHash (update_spatial_hash.slang:18-28)
#pragma omp parallel for schedule(static) num_threads(threads)
        for (int64_t i = 0; i < (int64_t)n; i++) {
            int32_t c[3];
            get_cell<R>(&x[3 * i], m.cell, c);
            keys[i] = hash_cell(c[0], c[1], c[2]) % n;
        }
        // GPUCountSort (sort.slang:13-55) with the stable scatter order
        std::fill(counts.begin(), counts.end(), 0u);
        for (uint32_t i = 0; i < n; i++) counts[keys[i]]++;
        uint32_t run = 0;
        for (uint32_t k = 0; k < n; k++) {
            const uint32_t c = counts[k];
            counts[k] = run;
            run += c;
        }
        for (uint32_t i = 0; i < n; i++) {
            const uint32_t dst = counts[keys[i]]++;
            sorted_keys[dst] = keys[i];
            perm[dst] = i;
        }
        // InitOffsets / CalculateOffsets (spatial_offsets.slang:10-31)
        std::fill(offsets.begin(), offsets.end(), n);
        for (uint32_t i = 0; i < n; i++) {
            const uint32_t key_prev = i == 0 ? n : sorted_keys[i - 1];
            if (sorted_keys[i] != key_prev) offsets[sorted_keys[i]] = i;
        }
        // Reorder + Copyback of x, v (reorder.slang:11-19, sort.cpp:206-227)
        for (uint32_t i = 0; i < n; i++) {
            const uint32_t s = perm[i];
            for (int c = 0; c < 3; c++) {
                tx[3 * i + c] = x[3 * s + c];
                tv[3 * i + c] = v[3 * s + c];
            }
            tpid[i] = pid[s];
        }
        x.swap(tx);
        v.swap(tv);
        pid.swap(tpid);
    }

    void substep() {
        if (n == 0) return;
        // ExternalAccel (slang:92-100)
        for (uint32_t i = 0; i < n; i++)
            for (int c = 0; c < 3; c++) a[3 * i + c] = m.g[c];
        hash_and_sort();
        if (tape && lists.size() != n) lists.assign(n, {});
        // CalculateBoundaryVolume
#pragma omp parallel for schedule(static) num_threads(threads)
        for (int64_t i = 0; i < (int64_t)n; i++) boundary_volume<R>(m, &x[3 * i], psi[i], &bgrad[3 * i]);
        // CalculateDensities (slang:153-197)
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads)
        for (int64_t ii = 0; ii < (int64_t)n; ii++) {
            const uint32_t i = (uint32_t)ii;
            const R* pos = &x[3 * i];
            int32_t oc[3];
            get_cell<R>(pos, m.cell, oc);
            R density = R(0);
            for (int o = 0; o < 27; o++) {
                const uint32_t key =
                    hash_cell(oc[0] + kOffsets3D[o][0], oc[1] + kOffsets3D[o][1], oc[2] + kOffsets3D[o][2]) % n;
                uint32_t cur = offsets[key];
                while (cur < n) {
                    const uint32_t j = cur++;
                    if (sorted_keys[j] != key) break;
                    R d[3];
                    for (int c = 0; c < 3; c++) d[c] = x[3 * j + c] - pos[c];
                    const R r2 = dot3<R>(d, d);
                    if (r2 > m.h2) continue;
                    const R r = std::sqrt(r2);
                    density += m.mass * m.W(r);
                    if (tape) lists[i].push_back(j);
                }
            }
            density += m.rho0 * psi[i];
            rho[i] = density;
        }
        if (tape) {
            TapeStep t;
            t.perm = perm;
            t.x = x;
            t.v = v;
            t.rho = rho;
            t.psi = psi;
            t.bgrad = bgrad;
            t.nbr.swap(lists);
            tape->push_back(std::move(t));
            lists.assign(n, {});
        }
        // CalculatePressureAccel (slang:205-271)
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads)
        for (int64_t ii = 0; ii < (int64_t)n; ii++) {
            const uint32_t i = (uint32_t)ii;
            const R di = rho[i];
            const R pi = m.eos(di);
            const R pdi = di > R(0) ? pi / (di * di) : R(0);
            const R* xi = &x[3 * i];
            int32_t oc[3];
            get_cell<R>(xi, m.cell, oc);
            R pf[3] = {R(0), R(0), R(0)};
            for (int o = 0; o < 27; o++) {
                const uint32_t key =
                    hash_cell(oc[0] + kOffsets3D[o][0], oc[1] + kOffsets3D[o][1], oc[2] + kOffsets3D[o][2]) % n;
                uint32_t cur = offsets[key];
                while (cur < n) {
                    const uint32_t j = cur++;
                    if (j == i) continue;
                    if (sorted_keys[j] != key) break;
                    R xij[3];
                    for (int c = 0; c < 3; c++) xij[c] = xi[c] - x[3 * j + c];
                    const R r2 = dot3<R>(xij, xij);
                    if (r2 > m.h2) continue;
                    const R dj = rho[j];
                    const R pj = m.eos(dj);
                    const R r = std::sqrt(r2);
                    R xn[3];
                    if (r > R(0)) {
                        const R inv_r = R(1) / r;
                        for (int c = 0; c < 3; c++) xn[c] = xij[c] * inv_r;
                    } else {
                        xn[0] = R(0);
                        xn[1] = R(1);
                        xn[2] = R(0);
                    }
                    const R pdj = dj > R(0) ? pj / (dj * dj) : R(0);
                    const R s = (m.mass * (pdi + pdj)) * m.dW(r);
                    for (int c = 0; c < 3; c++) pf[c] += s * xn[c];
                }
            }
            const R pdb = di > R(0) ? pi / (m.rho0 * m.rho0) : R(0);
            const R sb = (pdi + pdb) * m.rho0;
            for (int c = 0; c < 3; c++) {
                pf[c] += sb * bgrad[3 * i + c];
                a[3 * i + c] += -pf[c];
            }
        }
        // CalculateViscousAccel (slang:273-334)
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads)
        for (int64_t ii = 0; ii < (int64_t)n; ii++) {
            const uint32_t i = (uint32_t)ii;
            const R* xi = &x[3 * i];
            const R* vi = &v[3 * i];
            int32_t oc[3];
            get_cell<R>(xi, m.cell, oc);
            R vf[3] = {R(0), R(0), R(0)};
            for (int o = 0; o < 27; o++) {
                const uint32_t key =
                    hash_cell(oc[0] + kOffsets3D[o][0], oc[1] + kOffsets3D[o][1], oc[2] + kOffsets3D[o][2]) % n;
                uint32_t cur = offsets[key];
                while (cur < n) {
                    const uint32_t j = cur++;
                    if (j == i) continue;
                    if (sorted_keys[j] != key) break;
                    R xij[3];
                    for (int c = 0; c < 3; c++) xij[c] = xi[c] - x[3 * j + c];
                    const R r2 = dot3<R>(xij, xij);
                    if (r2 > m.h2) continue;
                    const R r = std::sqrt(r2);
                    R xn[3];
                    if (r > R(0)) {
                        const R inv_r = R(1) / r;
                        for (int c = 0; c < 3; c++) xn[c] = xij[c] * inv_r;
                    } else {
                        xn[0] = R(0);
                        xn[1] = R(1);
                        xn[2] = R(0);
                    }
                    R vij[3];
                    for (int c = 0; c < 3; c++) vij[c] = vi[c] - v[3 * j + c];
                    const R rhoj = std::fmax(rho[j], R(1e-6f));
                    const R G = m.dW(r);
                    const R s = (m.mass / rhoj) * dot3<R>(vij, xij);
                    const R inv_den = R(1) / (r2 + R(0.01f) * m.h2);
                    for (int c = 0; c < 3; c++) vf[c] += (s * (xn[c] * G)) * inv_den;
                }
            }
            const R f = (R(2) * (R(3) + R(2))) * m.nu;
            for (int c = 0; c < 3; c++) a[3 * i + c] += f * vf[c];
        }
        // WallAcceleration + UpdatePositions (slang:336-360)
#pragma omp parallel for schedule(static) num_threads(threads)
        for (int64_t ii = 0; ii < (int64_t)n; ii++) {
            const uint32_t i = (uint32_t)ii;
            R diff[3];
            bool any = false;
            for (int c = 0; c < 3; c++) {
                const R p = x[3 * i + c];
                diff[c] = std::fmin(p - m.bb_lo[c], R(0)) - std::fmin(m.bb_hi[c] - p, R(0));
                if (std::fabs(diff[c]) > R(1e-9f)) any = true;
            }
            for (int c = 0; c < 3; c++) {
                const R w = any ? (-m.kwall) * diff[c] : R(0);
                a[3 * i + c] += w;
                v[3 * i + c] += a[3 * i + c] * m.dt;
                x[3 * i + c] += v[3 * i + c] * m.dt;
            }
        }
    }

    void init(const ParamsC& p, const ObjectC* objs, uint32_t n_obj) {
        m.set_params(p);
        n = m.n;
        x.assign(3 * (size_t)n, R(0));
        v.assign(3 * (size_t)n, R(0));
        a.assign(3 * (size_t)n, R(0));
        tx.assign(3 * (size_t)n, R(0));
        tv.assign(3 * (size_t)n, R(0));
        bgrad.assign(3 * (size_t)n, R(0));
        rho.assign(n, R(0));
        psi.assign(n, R(0));
        pid.resize(n);
        tpid.resize(n);
        for (uint32_t i = 0; i < n; i++) pid[i] = i;
        keys.assign(n, 0);
        sorted_keys.assign(n, 0);
        perm.assign(n, 0);
        offsets.assign(n, n);
        counts.assign(n, 0);
        m.objs.clear();
        for (uint32_t k = 0; k < n_obj; k++) {
            typename Model<R>::Obj o;
            for (int e = 0; e < 16; e++) {
                o.T[e] = objs[k].transform[e];
                o.Rm[e] = objs[k].rotation[e];
            }
            for (int a2 = 0; a2 < 3; a2++) {
                o.box_pos[a2] = objs[k].box_pos[a2];
                o.box_size[a2] = objs[k].box_size[a2];
                o.res[a2] = objs[k].resolution[a2];
            }
            const size_t cnt = (size_t)o.res[0] * o.res[1] * o.res[2];
            o.sdf.assign(objs[k].sdf_grid, objs[k].sdf_grid + cnt);
            o.vm.assign(objs[k].volume_map_grid, objs[k].volume_map_grid + cnt);
            m.objs.push_back(std::move(o));
        }
    }
};

// ------------------------------------------------------ host adjoint (FP64)
// Reverse pass of the substep over the recorded neighbour lists, and the
// density-grid loss: the restatement of the GPU adjoint, in double, used to
// check it (and checked itself against finite differences of the FP64 forward).
struct GridC {
    float origin[3], spacing[3];
    uint32_t dims[3];
};

template <class R>
static R kernel_ddw(const Model<R>& m, R r) {
    const R q = r * m.inv_h;
    if ((r > R(1e-9f)) && (q <= R(1))) {
        if (q > R(0.5)) return ((R(2) * m.grad_sigma) * (R(1) - q)) * m.inv_h;
        return (m.grad_sigma * (R(6) * q - R(2))) * m.inv_h;
    }
    return R(0);
}

template <class R>
static R eos_dp(const Model<R>& m, R rho) {
    if (!(rho > m.rho0)) return R(0);
    return m.B * m.gamma * std::pow(rho / m.rho0, m.gamma - R(1)) / m.rho0;
}

template <class R>
static void grid_point(const GridC& g, uint32_t ix, uint32_t iy, uint32_t iz, R* p) {
    // the same float products as the device
    p[0] = (R)(g.origin[0] + (float)ix * g.spacing[0]);
    p[1] = (R)(g.origin[1] + (float)iy * g.spacing[1]);
    p[2] = (R)(g.origin[2] + (float)iz * g.spacing[2]);
}

template <class R>
static void grid_range(const GridC& g, const R* p, R h, int* lo, int* hi) {
    for (int a = 0; a < 3; a++) {
        lo[a] = (int)std::max(std::floor((double)(p[a] - h - (R)g.origin[a]) / g.spacing[a]) - 1.0, 0.0);
        hi[a] = (int)std::min(std::ceil((double)(p[a] + h - (R)g.origin[a]) / g.spacing[a]) + 1.0,
                              (double)g.dims[a] - 1.0);
    }
}

// rho(g) = sum_j m W(|g - x_j|) over particles within h.
template <class R>
static void grid_density(const Model<R>& m, const std::vector<R>& x, uint32_t n, const GridC& g,
                         std::vector<R>& out) {
    out.assign((size_t)g.dims[0] * g.dims[1] * g.dims[2], R(0));
    for (uint32_t j = 0; j < n; j++) {
        int lo[3], hi[3];
        grid_range<R>(g, &x[3 * j], m.h, lo, hi);
        for (int iz = lo[2]; iz <= hi[2]; iz++)
            for (int iy = lo[1]; iy <= hi[1]; iy++)
                for (int ix = lo[0]; ix <= hi[0]; ix++) {
                    R gp[3], d[3];
                    grid_point<R>(g, ix, iy, iz, gp);
                    for (int a = 0; a < 3; a++) d[a] = x[3 * j + a] - gp[a];
                    const R r2 = dot3<R>(d, d);
                    if (r2 > m.h2) continue;
                    out[(size_t)ix + g.dims[0] * ((size_t)iy + (size_t)g.dims[1] * iz)] += m.mass * m.W(std::sqrt(r2));
                }
    }
}

template <class R>
static void stencil_keys(const Model<R>& m, const R* p, uint32_t n, uint32_t* k) {
    int32_t c[3];
    if (m.float_cells) {
        const float pf[3] = {(float)p[0], (float)p[1], (float)p[2]};
        get_cell<float>(pf, (float)m.cell, c);
    } else {
        get_cell<R>(p, m.cell, c);
    }
    for (int o = 0; o < 27; o++)
        k[o] = hash_cell(c[0] + kOffsets3D[o][0], c[1] + kOffsets3D[o][1], c[2] + kOffsets3D[o][2]) % n;
}

template <class R>
static R pair_weight(const Model<R>& m, const R* xi, const R* xj, uint32_t n) {
    uint32_t ki[27], kj[27];
    stencil_keys<R>(m, xi, n, ki);
    stencil_keys<R>(m, xj, n, kj);
    // own keys are the stencil's centre (offset 13)
    uint32_t mij = 0, mji = 0;
    for (int o = 0; o < 27; o++) {
        mij += ki[o] == kj[13] ? 1u : 0u;
        mji += kj[o] == ki[13] ? 1u : 0u;
    }
    return mij ? (R)mji / (R)mij : R(0);
}

// value and d/duvw of SampleGrid with its base cell held fixed
template <class R>
static R sample_grid_grad(const float* grid, const uint32_t* n, const R* uvw, R* grad) {
    uint32_t c[3];
    R chi[3];
    for (int a = 0; a < 3; a++) {
        const R x = (R(n[a]) - R(1)) * uvw[a];
        const R f = std::floor(x);
        const uint32_t hi = n[a] >= 2 ? n[a] - 2 : 0;
        c[a] = f >= R(0) ? (f <= R(hi) ? (uint32_t)f : hi) : 0;
        chi[a] = R(2) * (x - R(c[a])) - R(1);
    }
    static const uint32_t off[8][3] = {{0, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1},
                                       {1, 0, 0}, {1, 1, 0}, {1, 0, 1}, {1, 1, 1}};
    R val = R(0);
    grad[0] = grad[1] = grad[2] = R(0);
    for (int i = 0; i < 8; i++) {
        R f[3], s[3];
        for (int a = 0; a < 3; a++) {
            s[a] = off[i][a] ? R(1) : R(-1);
            f[a] = R(1) + s[a] * chi[a];
        }
        const uint32_t idx = (c[0] + off[i][0]) + n[0] * (c[1] + off[i][1]) + n[0] * n[1] * (c[2] + off[i][2]);
        const R gv = R(grid[idx]);
        val += gv * ((f[0] * f[1]) * f[2]);
        grad[0] += gv * s[0] * f[1] * f[2];
        grad[1] += gv * f[0] * s[1] * f[2];
        grad[2] += gv * f[0] * f[1] * s[2];
    }
    for (int a = 0; a < 3; a++) grad[a] *= R(2) * (R(n[a]) - R(1)) / R(8);
    return val / R(8);
}

template <class R>
static void boundary_adjoint(const Model<R>& m, const R* pos, R psi_bar, const R* gb, R* gpos) {
    for (const auto& o : m.objs) {
        R lp[3];
        mat_vec<R>(o.T, pos, R(1), lp);
        bool outside = false;
        for (int a = 0; a < 3; a++)
            if (lp[a] < o.box_pos[a] || lp[a] > o.box_pos[a] + o.box_size[a]) outside = true;
        if (outside) continue;
        R uvw[3], gS[3], gV[3];
        to_uvw<R>(o.box_pos, o.box_size, lp, uvw);
        const R dist = sample_grid_grad<R>(o.sdf.data(), o.res, uvw, gS);
        if (!(dist >= R(0) && dist < m.h)) continue;
        const R vol = sample_grid_grad<R>(o.vm.data(), o.res, uvw, gV);
        if (!(vol > R(0))) continue;
        for (int a = 0; a < 3; a++) {
            gS[a] /= o.box_size[a];
            gV[a] /= o.box_size[a];
        }
        const R eps = R(0.1f) * m.h;
        R nrm[3], Jn[3][3];
        for (int a = 0; a < 3; a++) {
            R v[2], g[2][3];
            for (int sg = 0; sg < 2; sg++) {
                R q[3] = {lp[0], lp[1], lp[2]};
                q[a] = sg == 0 ? lp[a] + eps : lp[a] - eps;
                R u[3];
                to_uvw<R>(o.box_pos, o.box_size, q, u);
                v[sg] = sample_grid_grad<R>(o.sdf.data(), o.res, u, g[sg]);
            }
            nrm[a] = (v[0] - v[1]) / (R(2) * eps);
            for (int b = 0; b < 3; b++) Jn[a][b] = (g[0][b] - g[1][b]) / o.box_size[b] / (R(2) * eps);
        }
        const R modn = std::sqrt(dot3<R>(nrm, nrm));
        if (!(modn > R(1e-9f))) continue;
        R Rn[3], RJ[3][3], nJ[3];
        for (int r = 0; r < 3; r++) {
            Rn[r] = o.Rm[r] * nrm[0] + o.Rm[4 + r] * nrm[1] + o.Rm[8 + r] * nrm[2] + o.Rm[12 + r];
            for (int b = 0; b < 3; b++) RJ[r][b] = o.Rm[r] * Jn[0][b] + o.Rm[4 + r] * Jn[1][b] + o.Rm[8 + r] * Jn[2][b];
        }
        for (int b = 0; b < 3; b++) nJ[b] = nrm[0] * Jn[0][b] + nrm[1] * Jn[1][b] + nrm[2] * Jn[2][b];
        R rn[3], Jr[3][3];
        for (int r = 0; r < 3; r++) {
            rn[r] = Rn[r] / modn;
            for (int b = 0; b < 3; b++) Jr[r][b] = RJ[r][b] / modn - Rn[r] * nJ[b] / (modn * modn * modn);
        }
        const R dd = dist + R(0.125f) * m.h;
        const R d = std::fmax(dd, R(0.5f) * m.h);
        R xij[3], Jx[3][3];
        for (int r = 0; r < 3; r++) {
            xij[r] = d * rn[r];
            for (int b = 0; b < 3; b++) Jx[r][b] = (dd > R(0.5f) * m.h ? rn[r] * gS[b] : R(0)) + d * Jr[r][b];
        }
        const R xm = std::sqrt(dot3<R>(xij, xij));
        if (!(xm > R(0))) continue;
        R xh[3], dxm[3], perp[3];
        for (int r = 0; r < 3; r++) xh[r] = xij[r] / xm;
        for (int b = 0; b < 3; b++) dxm[b] = xh[0] * Jx[0][b] + xh[1] * Jx[1][b] + xh[2] * Jx[2][b];
        const R W = m.W(xm), G = m.dW(xm), Gp = kernel_ddw<R>(m, xm), Wp = m.h * G;
        const R xg = xh[0] * gb[0] + xh[1] * gb[1] + xh[2] * gb[2];
        for (int r = 0; r < 3; r++) perp[r] = (gb[r] - xg * xh[r]) / xm;
        R glp[3];
        for (int b = 0; b < 3; b++) {
            const R jp = Jx[0][b] * perp[0] + Jx[1][b] * perp[1] + Jx[2][b] * perp[2];
            glp[b] = psi_bar * (W * gV[b] + vol * Wp * dxm[b]) + xg * (G * gV[b] + vol * Gp * dxm[b]) + vol * G * jp;
        }
        for (int b = 0; b < 3; b++) gpos[b] += o.T[b * 4 + 0] * glp[0] + o.T[b * 4 + 1] * glp[1] + o.T[b * 4 + 2] * glp[2];
    }
}

// One substep backwards: gx/gv hold the adjoints of the substep's outputs
// (sorted order) on entry and of its inputs (input order) on return.
template <class R>
static void adjoint_substep(const Model<R>& m, const typename Sim<R>::TapeStep& t, uint32_t n, std::vector<R>& gx,
                            std::vector<R>& gv, R& nu_bar) {
    std::vector<R> ab(3 * (size_t)n), ax(3 * (size_t)n), av(3 * (size_t)n), rbar(n), gpsi(3 * (size_t)n);
    std::vector<uint8_t> dup(n);
    for (uint32_t i = 0; i < n; i++) {
        const R* p = &t.x[3 * i];
        R diff[3];
        bool wall = false;
        for (int k = 0; k < 3; k++) {
            diff[k] = std::fmin(p[k] - m.bb_lo[k], R(0)) - std::fmin(m.bb_hi[k] - p[k], R(0));
            if (std::fabs(diff[k]) > R(1e-9f)) wall = true;
        }
        for (int k = 0; k < 3; k++) {
            const R vt = gv[3 * i + k] + m.dt * gx[3 * i + k];
            ab[3 * i + k] = m.dt * vt;
            av[3 * i + k] = vt;
            ax[3 * i + k] = gx[3 * i + k];
            if (wall) {
                const R D = (p[k] < m.bb_lo[k] ? R(1) : R(0)) + (p[k] > m.bb_hi[k] ? R(1) : R(0));
                ax[3 * i + k] += -m.kwall * D * ab[3 * i + k];
            }
        }
        uint32_t k27[27];
        stencil_keys<R>(m, p, n, k27);
        bool d = false;
        for (int u = 1; u < 27 && !d; u++)
            for (int w = 0; w < u; w++) d |= k27[u] == k27[w];
        dup[i] = d ? 1 : 0;
    }
    const R nu10 = (R(2) * (R(3) + R(2))) * m.nu;
    auto P = [&](uint32_t j) {
        const R rj = t.rho[j];
        return rj > R(0) ? m.eos(rj) / (rj * rj) : R(0);
    };
    auto M = [&](uint32_t j) { return m.mass / std::fmax(t.rho[j], R(1e-6f)); };
    for (uint32_t i = 0; i < n; i++) {
        const R* xi = &t.x[3 * i];
        const R* vi = &t.v[3 * i];
        const R Pi = P(i), Mi = M(i);
        const R* ai = &ab[3 * i];
        R Pbar = 0, Mbar = 0, gxl[3] = {0, 0, 0}, gvl[3] = {0, 0, 0}, fv[3] = {0, 0, 0};
        for (uint32_t j : t.nbr[i]) {
            if (j == i) continue;
            const R* xj = &t.x[3 * j];
            const R* vj = &t.v[3 * j];
            R xij[3], vij[3];
            for (int k = 0; k < 3; k++) {
                xij[k] = xi[k] - xj[k];
                vij[k] = vi[k] - vj[k];
            }
            const R r2 = dot3<R>(xij, xij), r = std::sqrt(r2);
            if (!(r > R(1e-9f))) continue;
            R nh[3], F[3];
            const R G = m.dW(r), Gp = kernel_ddw<R>(m, r);
            for (int k = 0; k < 3; k++) {
                nh[k] = xij[k] / r;
                F[k] = nh[k] * G;
            }
            const R sig = dot3<R>(vij, xij);
            const R inv_den = R(1) / (r2 + R(0.01f) * m.h2);
            const R Pj = P(j), Mj = M(j);
            const R* aj = &ab[3 * j];
            const R w = (dup[i] || dup[j]) ? pair_weight<R>(m, xi, xj, n) : R(1);
            for (int k = 0; k < 3; k++) fv[k] += Mj * sig * inv_den * F[k];
            const R aF = ai[0] * F[0] + ai[1] * F[1] + ai[2] * F[2];
            const R bF = -(aj[0] * F[0] + aj[1] * F[1] + aj[2] * F[2]);
            Pbar += -m.mass * aF + w * (-m.mass * bF);
            Mbar += w * (nu10 * sig * inv_den * bF);
            R Fb[3];
            for (int k = 0; k < 3; k++)
                Fb[k] = (-m.mass * (Pi + Pj) + nu10 * Mj * sig * inv_den) * ai[k] -
                        w * (-m.mass * (Pj + Pi) + nu10 * Mi * sig * inv_den) * aj[k];
            const R sb = nu10 * Mj * inv_den * aF + w * nu10 * Mi * inv_den * bF;
            const R db = -nu10 * Mj * sig * inv_den * inv_den * aF - w * nu10 * Mi * sig * inv_den * inv_den * bF;
            const R Fn = Fb[0] * nh[0] + Fb[1] * nh[1] + Fb[2] * nh[2];
            for (int k = 0; k < 3; k++) {
                gxl[k] += (G / r) * Fb[k] + Fn * (Gp - G / r) * nh[k] + sb * vij[k] + R(2) * db * xij[k];
                gvl[k] += sb * xij[k];
            }
        }
        nu_bar += R(10) * (ai[0] * fv[0] + ai[1] * fv[1] + ai[2] * fv[2]);
        const R rhoi = t.rho[i];
        const R p = m.eos(rhoi);
        const R pb = rhoi > R(0) ? p / (m.rho0 * m.rho0) : R(0);
        const R ag = ai[0] * t.bgrad[3 * i] + ai[1] * t.bgrad[3 * i + 1] + ai[2] * t.bgrad[3 * i + 2];
        Pbar += -m.rho0 * ag;
        for (int k = 0; k < 3; k++) gpsi[3 * i + k] = -(Pi + pb) * m.rho0 * ai[k];
        const R dp = eos_dp<R>(m, rhoi);
        R rb = rhoi > R(0) ? -m.rho0 * ag * dp / (m.rho0 * m.rho0) : R(0);
        if (rhoi > R(0)) rb += Pbar * (dp / (rhoi * rhoi) - R(2) * p / (rhoi * rhoi * rhoi));
        if (rhoi > R(1e-6f)) rb += Mbar * (-m.mass / (rhoi * rhoi));
        rbar[i] = rb;
        for (int k = 0; k < 3; k++) {
            ax[3 * i + k] += gxl[k];
            av[3 * i + k] += gvl[k];
        }
    }
    for (uint32_t i = 0; i < n; i++) {
        const R* xi = &t.x[3 * i];
        R g[3] = {ax[3 * i], ax[3 * i + 1], ax[3 * i + 2]};
        for (uint32_t j : t.nbr[i]) {
            if (j == i) continue;
            const R* xj = &t.x[3 * j];
            R xij[3];
            for (int k = 0; k < 3; k++) xij[k] = xi[k] - xj[k];
            const R r = std::sqrt(dot3<R>(xij, xij));
            if (!(r > R(1e-9f))) continue;
            const R w = (dup[i] || dup[j]) ? pair_weight<R>(m, xi, xj, n) : R(1);
            const R s = m.mass * (m.h * m.dW(r)) * (rbar[i] + w * rbar[j]) / r;
            for (int k = 0; k < 3; k++) g[k] += s * xij[k];
        }
        if (!m.objs.empty()) boundary_adjoint<R>(m, xi, m.rho0 * rbar[i], &gpsi[3 * i], g);
        const uint32_t o = t.perm[i];
        for (int k = 0; k < 3; k++) {
            gx[3 * o + k] = g[k];
            gv[3 * o + k] = av[3 * i + k];
        }
    }
}

// ------------------------------------------------------- volume map (FP64)
// Gauss-Legendre nodes/weights on [-1,1] by Newton iteration on P_n.
static void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 0; i < (n + 1) / 2; i++) {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
        double pp = 0.0;
        for (int it = 0; it < 100; it++) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; j++) {
                const double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            const double z1 = z;
            z = z1 - p1 / pp;
            if (std::fabs(z - z1) < 1e-16) break;
        }
        x[i] = -z;
        x[n - 1 - i] = z;
        w[i] = w[n - 1 - i] = 2.0 / ((1.0 - z * z) * pp * pp);
    }
}

struct Shape {
    int kind;
    double p[6];
    double sdf(const float* q) const {
        const double px = q[0], py = q[1], pz = q[2];
        if (kind == 1) {
            const double dx = px - p[0], dy = py - p[1], dz = pz - p[2];
            return std::sqrt(dx * dx + dy * dy + dz * dz) - p[3];
        }
        const double qx = std::fabs(px - p[0]) - p[3], qy = std::fabs(py - p[1]) - p[4],
                     qz = std::fabs(pz - p[2]) - p[5];
        const double ox = std::fmax(qx, 0.0), oy = std::fmax(qy, 0.0), oz = std::fmax(qz, 0.0);
        const double outside = std::sqrt(ox * ox + oy * oy + oz * oz);
        const double inside = std::fmin(std::fmax(qx, std::fmax(qy, qz)), 0.0);
        const double d = outside + inside;
        return kind == 2 ? -d : d;
    }
    void aabb(float lo[3], float hi[3]) const {
        const double e[3] = {p[3], kind == 1 ? p[3] : p[4], kind == 1 ? p[3] : p[5]};
        for (int a = 0; a < 3; a++) {
            lo[a] = (float)(p[a] - e[a]);
            hi[a] = (float)(p[a] + e[a]);
        }
    }
};

// CubicSplineKernel (src/util/kernel.cpp:10-34), FP64
struct KernelD {
    double radius, factor, w_zero;
    explicit KernelD(double r) : radius(r), factor(8.0 / (M_PI * r * r * r)) { w_zero = W(0.0); }
    double W(double r) const {
        const double q = r / radius;
        if (q <= 1.0) {
            if (q <= 0.5) {
                const double q2 = q * q, q3 = q2 * q;
                return factor * (6.0 * q3 - 6.0 * q2 + 1.0);
            }
            return factor * (2.0 * std::pow(1.0 - q, 3.0));
        }
        return 0.0;
    }
};

// Volume-map value of one grid node (volume_map.cpp:23-52) for any signed
// distance sdf(const float q[3]) -> double; loops in the reference's order.
template <class SDF>
double volume_map_node(const SDF& sdf, const float* x, float support_radius, const std::vector<double>& gx,
                       const std::vector<double>& gw, const KernelD& kernel) {
    const double dist = sdf(x);
    if (dist > 2.0 * support_radius) return 0.0;
    const double c0 = 0.5 * (double)(support_radius - (-support_radius));
    const double c1 = 0.5 * (double)(support_radius + (-support_radius));
    const float sr2 = support_radius * support_radius;
    double integral = 0.0;
    for (int i = 0; i < 16; i++) {
        const double wi = gw[i];
        for (int j = 0; j < 16; j++) {
            const double wij = wi * gw[j];
            for (int k = 0; k < 16; k++) {
                const double wijk = wij * gw[k];
                const float xi[3] = {(float)(c0 * gx[i] + c1), (float)(c0 * gx[j] + c1), (float)(c0 * gx[k] + c1)};
                double val;
                if ((xi[0] * xi[0] + xi[1] * xi[1]) + xi[2] * xi[2] > sr2) {
                    val = 0.0;
                } else {
                    const float q[3] = {x[0] + xi[0], x[1] + xi[1], x[2] + xi[2]};
                    const double di = sdf(q);
                    if (di <= 0.0)
                        val = 1.0;
                    else if (di < support_radius)
                        val = kernel.W(di) / kernel.w_zero;
                    else
                        val = 0.0;
                }
                integral += wijk * val;
            }
        }
    }
    integral *= c0 * c0 * c0;
    return 0.8 * integral;
}

// ------------------------------------------------------------ Lague model
// LagueModel (src/models/lague_model.cpp:87-125, lague_model.slang), the
// Clavet double-density model: per substep ExternalForces (predicted
// positions), the spatial hash of the predicted positions with x, v and the
// predictions reordered, CalculateDensities, CalculatePressureForces and
// UpdatePositions.  Canonical readings as for the WCSPH restatement (fused
// dot products, vector / scalar as a multiplication by the reciprocal,
// integral pow by binary powering in double), and one more: the reference's
// pressure kernel reads neighbour velocities that other invocations of the
// same dispatch are updating (a race); here they are read as they were
// before the dispatch.
struct LagueParamsC {  // identical layout to vfs_lague_params
    float wall_damping_factor;
    float pressure_multiplier;
    float near_pressure_multiplier;
    float viscosity_strenght;
};

template <class R>
struct LagueSim {
    R g[3], h, cell, dt, pf, rho0, pm, npm, visc, damp;
    R sp2, sp3, dsp2, dsp3, poly6;
    R bb_lo[3], bb_hi[3];
    uint32_t n = 0;
    int threads = 1;
    std::vector<R> x, v, pred, rho, near_rho, nv;
    std::vector<uint32_t> pid, keys, sorted_keys, perm, offsets, counts;

    static R ipow(R x, int e) {  // canonical pow with an integral exponent
        double base = (double)x, result = 1.0;
        unsigned k = (unsigned)e;
        while (k) {
            if (k & 1u) result *= base;
            base *= base;
            k >>= 1;
        }
        return (R)result;
    }

    void init(const ParamsC& p, const LagueParamsC& lp) {
        for (int c = 0; c < 3; c++) g[c] = p.gravity[c];
        h = p.smooth_radius;
        cell = p.cell_size > 0.0f ? p.cell_size : p.smooth_radius;
        dt = (R)(p.time_scale * p.fixed_dt / (float)p.iterations);
        pf = (R)(1.0f / 120.0f);  // prediction_factor, lague_model.slang:218
        rho0 = p.target_density;
        pm = lp.pressure_multiplier;
        npm = lp.near_pressure_multiplier;
        visc = lp.viscosity_strenght;
        damp = lp.wall_damping_factor;
        // CalcKernelCoefficients (model.cpp:224-233): glm::pi<float>() and std::pow(float, int) in double
        const float pi_f = 3.14159265358979323846f;
        const float r = p.smooth_radius;
        sp3 = 15.0f / (pi_f * (float)std::pow((double)r, 6));
        sp2 = 15.0f / (2.0f * pi_f * (float)std::pow((double)r, 5));
        dsp3 = 45.0f / (pi_f * (float)std::pow((double)r, 6));
        dsp2 = 15.0f / (pi_f * (float)std::pow((double)r, 5));
        // Poly6 (kernels_3d.slang:16-24): 315 / (64 PI pow(|r|, 9)) with the shader's PI
        const float pi_shader = 3.1415926f;
        poly6 = (R)(315.0f / ((64.0f * pi_shader) * (float)ipow((float)std::fabs(r), 9)));
        for (int c = 0; c < 3; c++) {
            bb_lo[c] = p.bbox_pos[c];
            bb_hi[c] = (R)(p.bbox_pos[c] + p.bbox_size[c]);
        }
        n = (uint32_t)p.n_particles;
        x.assign(3 * (size_t)n, R(0));
        v.assign(3 * (size_t)n, R(0));
        pred.assign(3 * (size_t)n, R(0));
        nv.assign(3 * (size_t)n, R(0));
        rho.assign(n, R(0));
        near_rho.assign(n, R(0));
        pid.resize(n);
        for (uint32_t i = 0; i < n; i++) pid[i] = i;
        keys.assign(n, 0u);
        sorted_keys.assign(n, 0u);
        perm.assign(n, 0u);
        offsets.assign(n, 0u);
        counts.assign(n, 0u);
    }

    // kernels_3d.slang:26-60
    R spiky3(R d) const { return d < h ? (((h - d) * (h - d)) * (h - d)) * sp3 : R(0); }
    R spiky2(R d) const { return d < h ? ((h - d) * (h - d)) * sp2 : R(0); }
    R grad_spiky3(R d) const { return d <= h ? ((-(h - d)) * (h - d)) * dsp3 : R(0); }
    R grad_spiky2(R d) const { return d <= h ? (-(h - d)) * dsp2 : R(0); }
    R poly6_w(R d) const {
        if (d < h) {
            const R w = h * h - d * d;
            return ((w * w) * w) * poly6;
        }
        return R(0);
    }

    template <class F>
    void for_neighbours(const R* pos, F&& f) const {
        int32_t oc[3];
        get_cell<R>(pos, cell, oc);
        for (int o = 0; o < 27; o++) {
            const uint32_t key =
                hash_cell(oc[0] + kOffsets3D[o][0], oc[1] + kOffsets3D[o][1], oc[2] + kOffsets3D[o][2]) % n;
            uint32_t cur = offsets[key];
            while (cur < n) {
                const uint32_t j = cur++;
                if (!f(j, key)) break;
            }
        }
    }

    void substep() {
        if (n == 0) return;
        // ExternalForces (slang:211-221)
        for (uint32_t i = 0; i < n; i++)
            for (int c = 0; c < 3; c++) {
                v[3 * i + c] = v[3 * i + c] + g[c] * dt;
                pred[3 * i + c] = x[3 * i + c] + v[3 * i + c] * pf;
            }
        // spatial hash of the predicted positions; reorder of x, v and the predictions
        for (uint32_t i = 0; i < n; i++) {
            int32_t c[3];
            get_cell<R>(&pred[3 * i], cell, c);
            keys[i] = hash_cell(c[0], c[1], c[2]) % n;
        }
        std::fill(counts.begin(), counts.end(), 0u);
        for (uint32_t i = 0; i < n; i++) counts[keys[i]]++;
        uint32_t run = 0;
        for (uint32_t k = 0; k < n; k++) {
            const uint32_t c = counts[k];
            counts[k] = run;
            run += c;
        }
        for (uint32_t i = 0; i < n; i++) {
            const uint32_t d = counts[keys[i]]++;
            sorted_keys[d] = keys[i];
            perm[d] = i;
        }
        std::fill(offsets.begin(), offsets.end(), n);
        for (uint32_t i = 0; i < n; i++)
            if (i == 0 || sorted_keys[i] != sorted_keys[i - 1]) offsets[sorted_keys[i]] = i;
        {
            std::vector<R> tx(x.size()), tv(v.size()), tp(pred.size());
            std::vector<uint32_t> tid(n);
            for (uint32_t i = 0; i < n; i++) {
                const uint32_t s = perm[i];
                for (int c = 0; c < 3; c++) {
                    tx[3 * i + c] = x[3 * s + c];
                    tv[3 * i + c] = v[3 * s + c];
                    tp[3 * i + c] = pred[3 * s + c];
                }
                tid[i] = pid[s];
            }
            x.swap(tx);
            v.swap(tv);
            pred.swap(tp);
            pid.swap(tid);
        }
        const R h2 = h * h;
        // CalculateDensities (slang:60-97): self included
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads)
        for (int64_t ii = 0; ii < (int64_t)n; ii++) {
            const uint32_t i = (uint32_t)ii;
            const R* pos = &pred[3 * i];
            R density = R(0), near_density = R(0);
            for_neighbours(pos, [&](uint32_t j, uint32_t key) {
                if (sorted_keys[j] != key) return false;
                R d[3];
                for (int c = 0; c < 3; c++) d[c] = pred[3 * j + c] - pos[c];
                const R r2 = dot3<R>(d, d);
                if (r2 > h2) return true;
                const R dst = std::sqrt(r2);
                density += spiky2(dst);
                near_density += spiky3(dst);
                return true;
            });
            rho[i] = density;
            near_rho[i] = near_density;
        }
        // CalculatePressureForces (slang:99-160), velocities from before the dispatch
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads)
        for (int64_t ii = 0; ii < (int64_t)n; ii++) {
            const uint32_t i = (uint32_t)ii;
            const R density = rho[i], density_near = near_rho[i];
            const R pressure = (density - rho0) * pm;
            const R near_pressure = npm * density_near;
            const R* vel = &v[3 * i];
            const R* pos = &pred[3 * i];
            R pf3[3] = {0, 0, 0}, vf3[3] = {0, 0, 0};
            for_neighbours(pos, [&](uint32_t j, uint32_t key) {
                if (j == i) return true;
                if (sorted_keys[j] != key) return false;
                R off[3];
                for (int c = 0; c < 3; c++) off[c] = pred[3 * j + c] - pos[c];
                const R r2 = dot3<R>(off, off);
                if (r2 > h2) return true;
                const R nd = rho[j], nnd = near_rho[j];
                const R np = (nd - rho0) * pm;
                const R nnp = npm * nnd;
                const R shared = (pressure + np) * R(0.5);
                const R shared_near = (near_pressure + nnp) * R(0.5);
                const R dst = std::sqrt(r2);
                R dir[3];
                if (dst > R(0)) {
                    const R inv = R(1) / dst;
                    for (int c = 0; c < 3; c++) dir[c] = off[c] * inv;
                } else {
                    dir[0] = R(0);
                    dir[1] = R(1);
                    dir[2] = R(0);
                }
                const R dd = grad_spiky2(dst), ndd = grad_spiky3(dst);
                const R inv_nd = R(1) / nd, inv_nnd = R(1) / nnd;
                for (int c = 0; c < 3; c++) pf3[c] += ((dir[c] * dd) * shared) * inv_nd;
                for (int c = 0; c < 3; c++) pf3[c] += ((dir[c] * ndd) * shared_near) * inv_nnd;
                const R w = poly6_w(dst);
                for (int c = 0; c < 3; c++) vf3[c] += (v[3 * j + c] - vel[c]) * w;
                return true;
            });
            const R inv_density = R(1) / density;
            for (int c = 0; c < 3; c++) {
                const R acc = pf3[c] * inv_density + vf3[c] * visc;
                nv[3 * i + c] = vel[c] + acc * dt;
            }
        }
        v.swap(nv);
        // UpdatePositions + ResolveCollisions (slang:232-250, 162-200)
        for (uint32_t i = 0; i < n; i++) {
            R* pos = &x[3 * i];
            R* vel = &v[3 * i];
            for (int c = 0; c < 3; c++) pos[c] = pos[c] + vel[c] * dt;
            for (int c = 0; c < 3; c++)
                if (pos[c] < bb_lo[c]) {
                    pos[c] = bb_lo[c];
                    vel[c] = (-vel[c]) * damp;
                }
            for (int c = 0; c < 3; c++)
                if (pos[c] > bb_hi[c]) {
                    pos[c] = bb_hi[c];
                    vel[c] = (-vel[c]) * damp;
                }
        }
    }
};

// ------------------------------------------------------------------ meshes
// Signed distance to a triangle mesh (geometry.cpp:37-286,
// mesh_pseudonormals.cpp:11-103), by exhaustive search over the triangles in
// index order (the reference's BVH only prunes this search; ties keep the
// first triangle found).

enum class Feature { V0, V1, V2, E01, E12, E02, F };

struct PointTri {
    double d2;
    double point[3];
    Feature feature;
};

// DistancePointToTriangle (geometry.cpp:37-217): minimise
// |v0 + s e0 + t e1 - p|^2 over the triangle by Eberly's region classification.
PointTri point_triangle(const double p[3], const double v0[3], const double v1[3], const double v2[3]) {
    double diff[3], e0[3], e1[3];
    for (int a = 0; a < 3; a++) {
        diff[a] = v0[a] - p[a];
        e0[a] = v1[a] - v0[a];
        e1[a] = v2[a] - v0[a];
    }
    auto dot = [](const double* u, const double* w) { return (u[0] * w[0] + u[1] * w[1]) + u[2] * w[2]; };
    const double a00 = dot(e0, e0), a01 = dot(e0, e1), a11 = dot(e1, e1);
    const double b0 = dot(diff, e0), b1 = dot(diff, e1), c = dot(diff, diff);
    const double det = std::fabs(a00 * a11 - a01 * a01);
    double s = a01 * b1 - a11 * b0;
    double t = a01 * b0 - a00 * b1;
    double d2 = -1.0;
    Feature f = Feature::F;
    const double d2_v1 = a00 + 2.0 * b0 + c, d2_v2 = a11 + 2.0 * b1 + c;
    auto interior = [&]() { return s * (a00 * s + a01 * t + 2.0 * b0) + t * (a01 * s + a11 * t + 2.0 * b1) + c; };
    if (s + t <= det) {
        if (s < 0.0 && t < 0.0 && b0 < 0.0) {  // region 4, toward v1
            t = 0.0;
            if (-b0 >= a00) { f = Feature::V1; s = 1.0; d2 = d2_v1; }
            else { f = Feature::E01; s = -b0 / a00; d2 = b0 * s + c; }
        } else if (s < 0.0) {  // regions 4 (rest) and 3: edge s = 0
            s = 0.0;
            if (b1 >= 0.0) { f = Feature::V0; t = 0.0; d2 = c; }
            else if (-b1 >= a11) { f = Feature::V2; t = 1.0; d2 = d2_v2; }
            else { f = Feature::E02; t = -b1 / a11; d2 = b1 * t + c; }
        } else if (t < 0.0) {  // region 5: edge t = 0
            t = 0.0;
            if (b0 >= 0.0) { f = Feature::V0; s = 0.0; d2 = c; }
            else if (-b0 >= a00) { f = Feature::V1; s = 1.0; d2 = d2_v1; }
            else { f = Feature::E01; s = -b0 / a00; d2 = b0 * s + c; }
        } else {  // region 0: interior
            const double inv_det = 1.0 / det;
            s *= inv_det;
            t *= inv_det;
            f = Feature::F;
            d2 = interior();
        }
    } else if (s < 0.0) {  // region 2
        const double tmp0 = a01 + b0, tmp1 = a11 + b1;
        if (tmp1 > tmp0) {
            const double numer = tmp1 - tmp0, denom = a00 - 2.0 * a01 + a11;
            if (numer >= denom) { f = Feature::V1; s = 1.0; t = 0.0; d2 = d2_v1; }
            else { f = Feature::E12; s = numer / denom; t = 1.0 - s; d2 = interior(); }
        } else {
            s = 0.0;
            if (tmp1 <= 0.0) { f = Feature::V2; t = 1.0; d2 = d2_v2; }
            else if (b1 >= 0.0) { f = Feature::V0; t = 0.0; d2 = c; }
            else { f = Feature::E02; t = -b1 / a11; d2 = b1 * t + c; }
        }
    } else if (t < 0.0) {  // region 6
        const double tmp0 = a01 + b1, tmp1 = a00 + b0;
        if (tmp1 > tmp0) {
            const double numer = tmp1 - tmp0, denom = a00 - 2.0 * a01 + a11;
            if (numer >= denom) { f = Feature::V2; t = 1.0; s = 0.0; d2 = d2_v2; }
            else { f = Feature::E12; t = numer / denom; s = 1.0 - t; d2 = interior(); }
        } else {
            t = 0.0;
            if (tmp1 <= 0.0) { f = Feature::V1; s = 1.0; d2 = d2_v1; }
            else if (b0 >= 0.0) { f = Feature::V0; s = 0.0; d2 = c; }
            else { f = Feature::E01; s = -b0 / a00; d2 = b0 * s + c; }
        }
    } else {  // region 1
        const double numer = a11 + b1 - a01 - b0;
        if (numer <= 0.0) { f = Feature::V2; s = 0.0; t = 1.0; d2 = d2_v2; }
        else {
            const double denom = a00 - 2.0 * a01 + a11;
            if (numer >= denom) { f = Feature::V1; s = 1.0; t = 0.0; d2 = d2_v1; }
            else { f = Feature::E12; s = numer / denom; t = 1.0 - s; d2 = interior(); }
        }
    }
    if (d2 < 0.0) d2 = 0.0;
    PointTri r;
    r.d2 = d2;
    r.feature = f;
    for (int a = 0; a < 3; a++) r.point[a] = v0[a] + s * e0[a] + t * e1[a];
    return r;
}

struct MeshO {
    std::vector<float> v;       // 3 per vertex
    std::vector<uint32_t> t;    // 3 per triangle
    std::vector<float> face;    // 3 per triangle
    std::vector<float> edge;    // 9 per triangle: edges (0,1), (1,2), (0,2)
    std::vector<float> vert;    // 3 per vertex
    float lo[3], hi[3];

    // float vector helpers with glm's evaluation order
    static void fsub(const float* a, const float* b, float* o) {
        for (int k = 0; k < 3; k++) o[k] = a[k] - b[k];
    }
    static float fdot(const float* a, const float* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
    static void fnormalize(const float* a, float* o) {
        const float inv = 1.0f / std::sqrt(fdot(a, a));
        for (int k = 0; k < 3; k++) o[k] = a[k] * inv;
    }

    MeshO(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt)
        : v(verts, verts + 3 * (size_t)nv), t(tris, tris + 3 * (size_t)nt) {
        face.assign(3 * (size_t)nt, 0.0f);
        edge.assign(9 * (size_t)nt, 0.0f);
        vert.assign(3 * (size_t)nv, 0.0f);
        std::map<uint64_t, std::array<float, 3>> sums;  // edge key -> normal sum (stable references)
        auto slot = [&](uint32_t a, uint32_t b) -> std::array<float, 3>& {
            const uint64_t key = ((uint64_t)std::min(a, b) << 32) | std::max(a, b);
            return sums.emplace(key, std::array<float, 3>{0.0f, 0.0f, 0.0f}).first->second;
        };
        for (int k = 0; k < 3; k++) {
            lo[k] = std::numeric_limits<float>::max();
            hi[k] = std::numeric_limits<float>::lowest();
        }
        for (uint32_t i = 0; i < nt; i++) {
            const uint32_t ia = t[3 * i], ib = t[3 * i + 1], ic = t[3 * i + 2];
            const float *a = &v[3 * (size_t)ia], *b = &v[3 * (size_t)ib], *c = &v[3 * (size_t)ic];
            for (const float* p : {a, b, c})
                for (int k = 0; k < 3; k++) {
                    lo[k] = std::min(lo[k], p[k]);
                    hi[k] = std::max(hi[k], p[k]);
                }
            float ab[3], ac[3], ba[3], bc[3], cb[3], ca[3], n[3];
            fsub(b, a, ab);
            fsub(c, a, ac);
            fsub(a, b, ba);
            fsub(c, b, bc);
            fsub(b, c, cb);
            fsub(a, c, ca);
            const float cr[3] = {ab[1] * ac[2] - ac[1] * ab[2], ab[2] * ac[0] - ac[2] * ab[0],
                                 ab[0] * ac[1] - ac[0] * ab[1]};
            fnormalize(cr, n);
            for (int k = 0; k < 3; k++) face[3 * (size_t)i + k] = n[k];
            float u1[3], u2[3];
            fnormalize(ab, u1);
            fnormalize(ac, u2);
            const float al0 = std::acos(std::fabs(fdot(u1, u2)));
            fnormalize(ba, u1);
            fnormalize(bc, u2);
            const float al1 = std::acos(std::fabs(fdot(u1, u2)));
            fnormalize(cb, u1);
            fnormalize(ca, u2);
            const float al2 = std::acos(std::fabs(fdot(u1, u2)));
            for (int k = 0; k < 3; k++) {
                vert[3 * (size_t)ia + k] += al0 * n[k];
                vert[3 * (size_t)ib + k] += al1 * n[k];
                vert[3 * (size_t)ic + k] += al2 * n[k];
            }
            for (auto* e : {&slot(ia, ib), &slot(ib, ic), &slot(ia, ic)})
                for (int k = 0; k < 3; k++) (*e)[k] += n[k];
        }
        for (uint32_t j = 0; j < nv; j++) {
            float o[3];
            fnormalize(&vert[3 * (size_t)j], o);
            for (int k = 0; k < 3; k++) vert[3 * (size_t)j + k] = o[k];
        }
        for (uint32_t i = 0; i < nt; i++) {
            const uint32_t ia = t[3 * i], ib = t[3 * i + 1], ic = t[3 * i + 2];
            const std::array<float, 3>* es[3] = {&slot(ia, ib), &slot(ib, ic), &slot(ia, ic)};
            for (int e = 0; e < 3; e++) fnormalize(es[e]->data(), &edge[9 * (size_t)i + 3 * e]);
        }
    }

    // Exactly equal distances resolve to the lowest triangle index (the first
    // found in index order); the reference's choice depends on its BVH
    // traversal order.
    double signed_distance(const float* q) const {
        const double p[3] = {q[0], q[1], q[2]};
        double best = (double)std::numeric_limits<float>::max();
        uint32_t best_t = 0;
        PointTri best_r{};
        const uint32_t nt = (uint32_t)(t.size() / 3);
        for (uint32_t i = 0; i < nt; i++) {
            double tv[3][3];
            for (int c = 0; c < 3; c++)
                for (int k = 0; k < 3; k++) tv[c][k] = v[3 * (size_t)t[3 * i + c] + k];
            const PointTri r = point_triangle(p, tv[0], tv[1], tv[2]);
            if (r.d2 < best) {
                best = r.d2;
                best_t = i;
                best_r = r;
            }
        }
        const float* n;
        switch (best_r.feature) {
            case Feature::V0: n = &vert[3 * (size_t)t[3 * best_t]]; break;
            case Feature::V1: n = &vert[3 * (size_t)t[3 * best_t + 1]]; break;
            case Feature::V2: n = &vert[3 * (size_t)t[3 * best_t + 2]]; break;
            case Feature::E01: n = &edge[9 * (size_t)best_t]; break;
            case Feature::E12: n = &edge[9 * (size_t)best_t + 3]; break;
            case Feature::E02: n = &edge[9 * (size_t)best_t + 6]; break;
            default: n = &face[3 * (size_t)best_t]; break;
        }
        const float cp[3] = {(float)best_r.point[0], (float)best_r.point[1], (float)best_r.point[2]};
        const float u[3] = {q[0] - cp[0], q[1] - cp[1], q[2] - cp[2]};
        const double sign = fdot(u, n) >= 0.0f ? 1.0 : -1.0;
        return sign * std::sqrt(best);
    }
};

}  // namespace orc

using namespace orc;

extern "C" {

uint32_t orc_hash_cell(int32_t cx, int32_t cy, int32_t cz) { return hash_cell(cx, cy, cz); }

void orc_hash_keys(const float* x, uint32_t n, float cell_size, uint32_t table_size, uint32_t* keys) {
    for (uint32_t i = 0; i < n; i++) {
        int32_t c[3];
        get_cell<float>(&x[3 * (size_t)i], cell_size, c);
        keys[i] = hash_cell(c[0], c[1], c[2]) % table_size;
    }
}

// Stable counting sort of (iota, keys): sorted_keys and perm (perm[dst] = src).
void orc_stable_sort(const uint32_t* keys, uint32_t n, uint32_t max_value, uint32_t* sorted_keys,
                     uint32_t* perm) {
    std::vector<uint32_t> counts((size_t)max_value + 1, 0u);
    for (uint32_t i = 0; i < n; i++) counts[keys[i]]++;
    uint32_t run = 0;
    for (size_t k = 0; k < counts.size(); k++) {
        const uint32_t c = counts[k];
        counts[k] = run;
        run += c;
    }
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t dst = counts[keys[i]]++;
        sorted_keys[dst] = keys[i];
        perm[dst] = i;
    }
}

void orc_offsets(const uint32_t* sorted_keys, uint32_t n, uint32_t* offsets) {
    for (uint32_t k = 0; k < n; k++) offsets[k] = n;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t prev = i == 0 ? n : sorted_keys[i - 1];
        if (sorted_keys[i] != prev) offsets[sorted_keys[i]] = i;
    }
}

void orc_exclusive_scan(const uint32_t* in, uint32_t n, uint32_t* out) {
    uint32_t run = 0;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t c = in[i];
        out[i] = run;
        run += c;
    }
}

float orc_W_f32(const ParamsC* p, float r) {
    Model<float> m;
    m.set_params(*p);
    return m.W(r);
}
float orc_dW_f32(const ParamsC* p, float r) {
    Model<float> m;
    m.set_params(*p);
    return m.dW(r);
}
double orc_W_f64(const ParamsC* p, double r) {
    Model<double> m;
    m.set_params(*p);
    return m.W(r);
}
double orc_dW_f64(const ParamsC* p, double r) {
    Model<double> m;
    m.set_params(*p);
    return m.dW(r);
}
float orc_eos_f32(const ParamsC* p, float rho) {
    Model<float> m;
    m.set_params(*p);
    return m.eos(rho);
}
// Derived uniforms: {mass, sigma, grad_sigma, dt}
void orc_uniforms_f32(const ParamsC* p, float* out4) {
    Model<float> m;
    m.set_params(*p);
    out4[0] = m.mass;
    out4[1] = m.sigma;
    out4[2] = m.grad_sigma;
    out4[3] = m.dt;
}

float orc_sample_grid_f32(const float* grid, const uint32_t* n, const float* uvw) {
    return sample_grid<float>(grid, n, uvw);
}
double orc_sample_grid_f64(const float* grid, const uint32_t* n, const double* uvw) {
    return sample_grid<double>(grid, n, uvw);
}

struct OrcSim {
    int fp64;
    Sim<float> s32;
    Sim<double> s64;
};

void* orc_sim_create(const ParamsC* p, const ObjectC* objs, uint32_t n_obj, int fp64) {
    if (!p || p->n_particles < 0) return nullptr;
    auto* s = new OrcSim();
    s->fp64 = fp64;
    if (fp64)
        s->s64.init(*p, objs, n_obj);
    else
        s->s32.init(*p, objs, n_obj);
    return s;
}
void orc_sim_destroy(void* h) { delete (OrcSim*)h; }

// GradientVector (slang:68-85) of boundary object k's SDF grid at the
// object-space position x, in the simulation's working type.
int orc_sim_gradient_vector(void* h, uint32_t k, const double* x, double* out) {
    auto* s = (OrcSim*)h;
    if (s->fp64) {
        if (k >= s->s64.m.objs.size()) return -1;
        gradient_vector<double>(s->s64.m.objs[k], s->s64.m.h, x, out);
        return 0;
    }
    if (k >= s->s32.m.objs.size()) return -1;
    const float xf[3] = {(float)x[0], (float)x[1], (float)x[2]};
    float g[3];
    gradient_vector<float>(s->s32.m.objs[k], s->s32.m.h, xf, g);
    for (int a = 0; a < 3; a++) out[a] = g[a];
    return 0;
}

void orc_sim_set_threads(void* h, int threads) {
    auto* s = (OrcSim*)h;
    s->s32.threads = s->s64.threads = threads < 1 ? 1 : threads;
}

// positions/velocities AoS float3 (offset/count as SetParticleState)
void orc_sim_set_state(void* h, const float* x, const float* v, uint32_t offset, uint32_t count) {
    auto* s = (OrcSim*)h;
    auto set = [&](auto& sim) {
        for (size_t i = 0; i < 3 * (size_t)count; i++) {
            sim.x[3 * (size_t)offset + i] = x[i];
            sim.v[3 * (size_t)offset + i] = v[i];
        }
        for (uint32_t i = 0; i < count; i++) sim.rho[offset + i] = 0;
        for (uint32_t i = 0; i < sim.n; i++) sim.pid[i] = i;  // ids restart at the current indices
    };
    if (s->fp64)
        set(s->s64);
    else
        set(s->s32);
}

// State in double (rounded to float for the FP32 instance): finite
// differences of the FP64 instance need perturbations below float resolution.
void orc_sim_set_state_f64(void* h, const double* x, const double* v, uint32_t offset, uint32_t count) {
    auto* s = (OrcSim*)h;
    auto set = [&](auto& sim) {
        using R = typename std::decay<decltype(sim.x[0])>::type;
        for (size_t i = 0; i < 3 * (size_t)count; i++) {
            sim.x[3 * (size_t)offset + i] = (R)x[i];
            sim.v[3 * (size_t)offset + i] = (R)v[i];
        }
        for (uint32_t i = 0; i < count; i++) sim.rho[offset + i] = 0;
        for (uint32_t i = 0; i < sim.n; i++) sim.pid[i] = i;
    };
    if (s->fp64)
        set(s->s64);
    else
        set(s->s32);
}

// FP64 host adjoint: from the simulation's current state, k substeps with the
// density-grid loss against target; returns the loss and writes the gradients
// with respect to the state at the start (positions, velocities, AoS) and to
// the viscosity.  The simulation is left at the final state.  FP64 instance only.
int orc_adjoint_run(void* h, uint32_t k, const orc::GridC* grid, const double* target, double* loss,
                    double* grad_x, double* grad_v, double* grad_nu, double* density) {
    auto* o = (OrcSim*)h;
    if (!o->fp64 || !grid) return -1;
    auto& sim = o->s64;
    const uint32_t n = sim.n;
    std::vector<orc::Sim<double>::TapeStep> tape;
    sim.tape = &tape;
    for (uint32_t s = 0; s < k; s++) sim.substep();
    sim.tape = nullptr;
    sim.lists.clear();
    std::vector<double> rg;
    orc::grid_density<double>(sim.m, sim.x, n, *grid, rg);
    const size_t G = rg.size();
    double L = 0.0;
    for (size_t g = 0; g < G; g++) {
        const double d = rg[g] - (target ? target[g] : 0.0);
        L += d * d;
    }
    if (loss) *loss = L / (double)G;
    if (density)
        for (size_t g = 0; g < G; g++) density[g] = rg[g];
    std::vector<double> gx(3 * (size_t)n, 0.0), gv(3 * (size_t)n, 0.0);
    for (uint32_t j = 0; j < n; j++) {
        int lo[3], hi[3];
        orc::grid_range<double>(*grid, &sim.x[3 * j], sim.m.h, lo, hi);
        for (int iz = lo[2]; iz <= hi[2]; iz++)
            for (int iy = lo[1]; iy <= hi[1]; iy++)
                for (int ix = lo[0]; ix <= hi[0]; ix++) {
                    double gp[3], d[3];
                    orc::grid_point<double>(*grid, ix, iy, iz, gp);
                    for (int a = 0; a < 3; a++) d[a] = sim.x[3 * j + a] - gp[a];
                    const double r2 = orc::dot3<double>(d, d);
                    if (r2 > sim.m.h2) continue;
                    const double r = std::sqrt(r2);
                    if (!(r > 1e-9)) continue;
                    const size_t idx = (size_t)ix + grid->dims[0] * ((size_t)iy + (size_t)grid->dims[1] * iz);
                    const double cg = 2.0 * (rg[idx] - (target ? target[idx] : 0.0)) / (double)G;
                    const double s2 = cg * sim.m.mass * (sim.m.h * sim.m.dW(r)) / r;
                    for (int a = 0; a < 3; a++) gx[3 * j + a] += s2 * d[a];
                }
    }
    double nu_bar = 0.0;
    for (int s = (int)k - 1; s >= 0; s--) orc::adjoint_substep<double>(sim.m, tape[s], n, gx, gv, nu_bar);
    for (size_t i = 0; i < 3 * (size_t)n; i++) {
        if (grad_x) grad_x[i] = gx[i];
        if (grad_v) grad_v[i] = gv[i];
    }
    if (grad_nu) *grad_nu = nu_bar;
    return 0;
}

// The FP64 adjoint of the FP32 trajectory: the forward runs on the FP32
// handle h32 (the trajectory the GPU reproduces bit for bit), its tape is
// widened to double, and the loss gradient and the adjoint use the model of
// the FP64 handle h64 (same parameters and objects; its state is unused).
// Separates the adjoint's arithmetic from the divergence of FP32 and FP64
// trajectories.  The loss is evaluated in double (as the GPU evaluates it:
// float positions and grid points, kernel values in double).
int orc_adjoint_run_f32_trajectory(void* h32, void* h64, uint32_t k, const orc::GridC* grid, const double* target,
                                   double* loss, double* grad_x, double* grad_v, double* grad_nu) {
    auto* o32 = (OrcSim*)h32;
    auto* o64 = (OrcSim*)h64;
    if (o32->fp64 || !o64->fp64 || !grid) return -1;
    auto& sim = o32->s32;
    orc::Model<double> m = o64->s64.m;
    m.float_cells = true;
    const uint32_t n = sim.n;
    std::vector<orc::Sim<float>::TapeStep> tape32;
    sim.tape = &tape32;
    for (uint32_t s = 0; s < k; s++) sim.substep();
    sim.tape = nullptr;
    sim.lists.clear();
    auto widen = [](const std::vector<float>& a) { return std::vector<double>(a.begin(), a.end()); };
    std::vector<orc::Sim<double>::TapeStep> tape(tape32.size());
    for (size_t s = 0; s < tape32.size(); s++) {
        tape[s].perm = tape32[s].perm;
        tape[s].x = widen(tape32[s].x);
        tape[s].v = widen(tape32[s].v);
        tape[s].rho = widen(tape32[s].rho);
        tape[s].psi = widen(tape32[s].psi);
        tape[s].bgrad = widen(tape32[s].bgrad);
        tape[s].nbr.swap(tape32[s].nbr);
    }
    tape32.clear();
    const std::vector<double> x = widen(sim.x);
    std::vector<double> rg;
    orc::grid_density<double>(m, x, n, *grid, rg);
    const size_t G = rg.size();
    double L = 0.0;
    for (size_t g = 0; g < G; g++) {
        const double d = rg[g] - (target ? target[g] : 0.0);
        L += d * d;
    }
    if (loss) *loss = L / (double)G;
    std::vector<double> gx(3 * (size_t)n, 0.0), gv(3 * (size_t)n, 0.0);
    for (uint32_t j = 0; j < n; j++) {
        int lo[3], hi[3];
        orc::grid_range<double>(*grid, &x[3 * j], m.h, lo, hi);
        for (int iz = lo[2]; iz <= hi[2]; iz++)
            for (int iy = lo[1]; iy <= hi[1]; iy++)
                for (int ix = lo[0]; ix <= hi[0]; ix++) {
                    double gp[3], d[3];
                    orc::grid_point<double>(*grid, ix, iy, iz, gp);
                    for (int a = 0; a < 3; a++) d[a] = x[3 * j + a] - gp[a];
                    const double r2 = orc::dot3<double>(d, d);
                    if (r2 > m.h2) continue;
                    const double r = std::sqrt(r2);
                    if (!(r > 1e-9)) continue;
                    const size_t idx = (size_t)ix + grid->dims[0] * ((size_t)iy + (size_t)grid->dims[1] * iz);
                    const double cg = 2.0 * (rg[idx] - (target ? target[idx] : 0.0)) / (double)G;
                    const double s2 = cg * m.mass * (m.h * m.dW(r)) / r;
                    for (int a = 0; a < 3; a++) gx[3 * j + a] += s2 * d[a];
                }
    }
    double nu_bar = 0.0;
    for (int s = (int)k - 1; s >= 0; s--) orc::adjoint_substep<double>(m, tape[s], n, gx, gv, nu_bar);
    for (size_t i = 0; i < 3 * (size_t)n; i++) {
        if (grad_x) grad_x[i] = gx[i];
        if (grad_v) grad_v[i] = gv[i];
    }
    if (grad_nu) *grad_nu = nu_bar;
    return 0;
}

void orc_sim_substeps(void* h, uint32_t k) {
    auto* s = (OrcSim*)h;
    for (uint32_t i = 0; i < k; i++) {
        if (s->fp64)
            s->s64.substep();
        else
            s->s32.substep();
    }
}

// State as doubles (exact for the FP32 instance); any pointer may be null.
void orc_sim_get_state(void* h, double* x, double* v, double* rho, uint32_t* pid, double* accel) {
    auto* s = (OrcSim*)h;
    auto get = [&](auto& sim) {
        const size_t n = sim.n;
        for (size_t i = 0; i < 3 * n; i++) {
            if (x) x[i] = (double)sim.x[i];
            if (v) v[i] = (double)sim.v[i];
            if (accel) accel[i] = (double)sim.a[i];
        }
        for (size_t i = 0; i < n; i++) {
            if (rho) rho[i] = (double)sim.rho[i];
            if (pid) pid[i] = sim.pid[i];
        }
    };
    if (s->fp64)
        get(s->s64);
    else
        get(s->s32);
}

void orc_sim_get_state_f32(void* h, float* x, float* v, float* rho) {
    auto* s = (OrcSim*)h;
    if (s->fp64) return;
    auto& sim = s->s32;
    const size_t n = sim.n;
    if (x) std::memcpy(x, sim.x.data(), 12 * n);
    if (v) std::memcpy(v, sim.v.data(), 12 * n);
    if (rho) std::memcpy(rho, sim.rho.data(), 4 * n);
}

void orc_sim_get_hash(void* h, uint32_t* sorted_keys, uint32_t* perm, uint32_t* offsets) {
    auto* s = (OrcSim*)h;
    auto get = [&](auto& sim) {
        const size_t n = sim.n;
        if (sorted_keys) std::memcpy(sorted_keys, sim.sorted_keys.data(), 4 * n);
        if (perm) std::memcpy(perm, sim.perm.data(), 4 * n);
        if (offsets) std::memcpy(offsets, sim.offsets.data(), 4 * n);
    };
    if (s->fp64)
        get(s->s64);
    else
        get(s->s32);
}

void orc_sim_get_boundary(void* h, double* psi, double* grad) {
    auto* s = (OrcSim*)h;
    auto get = [&](auto& sim) {
        const size_t n = sim.n;
        for (size_t i = 0; i < n; i++) {
            if (psi) psi[i] = (double)sim.psi[i];
            if (grad)
                for (int c = 0; c < 3; c++) grad[3 * i + c] = (double)sim.bgrad[3 * i + c];
        }
    };
    if (s->fp64)
        get(s->s64);
    else
        get(s->s32);
}

// One density/pressure evaluation on the CURRENT order without hashing again is
// not exposed: every call path goes through the full substep, as the shaders do.

// Gauss-Legendre table (for tests)
void orc_gauss_legendre(int n, double* x, double* w) {
    std::vector<double> xs, ws;
    gauss_legendre(n, xs, ws);
    for (int i = 0; i < n; i++) {
        x[i] = xs[i];
        w[i] = ws[i];
    }
}

// Volume map + SDF grids of an analytic shape, following mesh_sdf.cpp:28-45,
// discretization.cpp:16-33, volume_map.cpp:9-54 and gaussian_quadrature.cpp:2293-2333.
int orc_volume_map(int kind, const double* params, const uint32_t* res, double support_radius_d,
                   double tolerance, double margin, float* sdf_out, float* vm_out, float* box_size,
                   float* box_pos, int threads) {
    Shape sh;
    sh.kind = kind;
    for (int i = 0; i < 6; i++) sh.p[i] = params[i];
    const float support_radius = (float)support_radius_d;
    float lo[3], hi[3];
    sh.aabb(lo, hi);
    const double total_margin = tolerance + margin;
    float bsize[3], bpos[3];
    for (int a = 0; a < 3; a++) {
        bsize[a] = (hi[a] - lo[a]) + 2.0f * (float)total_margin;
        bpos[a] = lo[a] - (float)total_margin;
        box_size[a] = bsize[a];
        box_pos[a] = bpos[a];
    }
    float dmin[3], dmax[3];
    double step[3];
    for (int a = 0; a < 3; a++) {
        dmin[a] = bpos[a];
        dmax[a] = bpos[a] + bsize[a];
        step[a] = (double)((dmax[a] - dmin[a]) / (float)(res[a] - 1u));
    }
    std::vector<double> gx, gw;
    gauss_legendre(16, gx, gw);
    KernelD kernel((double)support_radius);
    const double c0 = 0.5 * (double)(support_radius - (-support_radius));
    const double c1 = 0.5 * (double)(support_radius + (-support_radius));
    const float sr2 = support_radius * support_radius;

    auto vm_func = [&](const float* x) -> double {
        const double dist = sh.sdf(x);
        if (dist > 2.0 * support_radius) return 0.0;
        double integral = 0.0;
        for (int i = 0; i < 16; i++) {
            const double wi = gw[i];
            for (int j = 0; j < 16; j++) {
                const double wij = wi * gw[j];
                for (int k = 0; k < 16; k++) {
                    const double wijk = wij * gw[k];
                    const float xi[3] = {(float)(c0 * gx[i] + c1), (float)(c0 * gx[j] + c1),
                                         (float)(c0 * gx[k] + c1)};
                    double val;
                    if ((xi[0] * xi[0] + xi[1] * xi[1]) + xi[2] * xi[2] > sr2) {
                        val = 0.0;
                    } else {
                        const float q[3] = {x[0] + xi[0], x[1] + xi[1], x[2] + xi[2]};
                        const double di = sh.sdf(q);
                        if (di <= 0.0)
                            val = 1.0;
                        else if (di < support_radius)
                            val = kernel.W(di) / kernel.w_zero;
                        else
                            val = 0.0;
                    }
                    integral += wijk * val;
                }
            }
        }
        integral *= c0 * c0 * c0;
        return 0.8 * integral;
    };

    const int64_t total = (int64_t)res[0] * res[1] * res[2];
    (void)threads;
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 0 ? threads : 1)
    for (int64_t idx = 0; idx < total; idx++) {
        const uint32_t i = (uint32_t)(idx % res[0]);
        const uint32_t j = (uint32_t)((idx / res[0]) % res[1]);
        const uint32_t k = (uint32_t)(idx / ((int64_t)res[0] * res[1]));
        const double ep[3] = {step[0] * i + (double)dmin[0], step[1] * j + (double)dmin[1],
                              step[2] * k + (double)dmin[2]};
        const float xf[3] = {(float)ep[0], (float)ep[1], (float)ep[2]};
        sdf_out[idx] = (float)(sh.sdf(xf) - tolerance);
        vm_out[idx] = (float)vm_func(xf);
    }
    return 0;
}

// Pseudonormals of a mesh (face 3, edges 9 per triangle; vertices 3 per vertex).
int orc_mesh_pseudonormals(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt, float* face,
                           float* edge, float* vert) {
    const MeshO m(verts, nv, tris, nt);
    std::copy(m.face.begin(), m.face.end(), face);
    std::copy(m.edge.begin(), m.edge.end(), edge);
    std::copy(m.vert.begin(), m.vert.end(), vert);
    return 0;
}

// DistancePointToTriangle for n (point, triangle) pairs: d2, closest point, feature index.
int orc_point_triangle(const double* p, const double* tri9, uint32_t n, double* d2, double* point, int* feature) {
    for (uint32_t i = 0; i < n; i++) {
        const PointTri r = point_triangle(p + 3 * (size_t)i, tri9 + 9 * (size_t)i, tri9 + 9 * (size_t)i + 3,
                                          tri9 + 9 * (size_t)i + 6);
        d2[i] = r.d2;
        for (int a = 0; a < 3; a++) point[3 * (size_t)i + a] = r.point[a];
        feature[i] = (int)r.feature;
    }
    return 0;
}

int orc_mesh_signed_distance(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt, const float* pts,
                             uint32_t n, double* out, int threads) {
    const MeshO m(verts, nv, tris, nt);
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 0 ? threads : 1)
    for (int64_t i = 0; i < (int64_t)n; i++) out[i] = m.signed_distance(pts + 3 * (size_t)i);
    return 0;
}

// SDF + volume-map grids of a mesh: MeshSDF::Build (mesh_sdf.cpp:28-45) and
// GenerateVolumeMap (volume_map.cpp:9-54) with the exhaustive signed distance.
int orc_mesh_volume_map(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt, const uint32_t* res,
                        double support_radius_d, double tolerance, double margin, float* sdf_out, float* vm_out,
                        float* box_size, float* box_pos, int threads) {
    const MeshO m(verts, nv, tris, nt);
    const float support_radius = (float)support_radius_d;
    const double total_margin = tolerance + margin;
    float dmin[3];
    double step[3];
    for (int a = 0; a < 3; a++) {
        box_size[a] = (m.hi[a] - m.lo[a]) + 2.0f * (float)total_margin;
        box_pos[a] = m.lo[a] - (float)total_margin;
        dmin[a] = box_pos[a];
        const float dmax = box_pos[a] + box_size[a];
        step[a] = (double)((dmax - dmin[a]) / (float)(res[a] - 1u));
    }
    std::vector<double> gx, gw;
    gauss_legendre(16, gx, gw);
    const KernelD kernel((double)support_radius);
    auto sdf = [&](const float* q) { return m.signed_distance(q); };
    const int64_t total = (int64_t)res[0] * res[1] * res[2];
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads > 0 ? threads : 1)
    for (int64_t idx = 0; idx < total; idx++) {
        const uint32_t i = (uint32_t)(idx % res[0]);
        const uint32_t j = (uint32_t)((idx / res[0]) % res[1]);
        const uint32_t k = (uint32_t)(idx / ((int64_t)res[0] * res[1]));
        const float xf[3] = {(float)(step[0] * i + (double)dmin[0]), (float)(step[1] * j + (double)dmin[1]),
                             (float)(step[2] * k + (double)dmin[2])};
        sdf_out[idx] = (float)(m.signed_distance(xf) - tolerance);
        vm_out[idx] = (float)volume_map_node(sdf, xf, support_radius, gx, gw, kernel);
    }
    return 0;
}

// Lague model: handle of an FP32 (fp64 = 0) or FP64 restatement.
struct OrcLague {
    int fp64 = 0;
    LagueSim<float> s32;
    LagueSim<double> s64;
};

void* orc_lague_create(const ParamsC* p, const LagueParamsC* lp, int fp64, int threads) {
    if (!p || !lp || p->n_particles < 0) return nullptr;
    auto* s = new OrcLague();
    s->fp64 = fp64;
    if (fp64) {
        s->s64.init(*p, *lp);
        s->s64.threads = threads < 1 ? 1 : threads;
    } else {
        s->s32.init(*p, *lp);
        s->s32.threads = threads < 1 ? 1 : threads;
    }
    return s;
}
void orc_lague_destroy(void* h) { delete (OrcLague*)h; }

void orc_lague_set_state(void* h, const float* x, const float* v) {
    auto* s = (OrcLague*)h;
    auto set = [&](auto& sim) {
        for (size_t i = 0; i < 3 * (size_t)sim.n; i++) {
            sim.x[i] = x[i];
            sim.v[i] = v[i];
        }
        for (uint32_t i = 0; i < sim.n; i++) sim.pid[i] = i;
    };
    if (s->fp64) set(s->s64); else set(s->s32);
}

void orc_lague_substeps(void* h, uint32_t k) {
    auto* s = (OrcLague*)h;
    for (uint32_t i = 0; i < k; i++) {
        if (s->fp64) s->s64.substep(); else s->s32.substep();
    }
}

// x, v (3 n), densities, near densities (n), ids (n); any may be null
void orc_lague_get_state(void* h, double* x, double* v, double* rho, double* near_rho, uint32_t* pid) {
    auto* s = (OrcLague*)h;
    auto get = [&](auto& sim) {
        for (size_t i = 0; i < 3 * (size_t)sim.n; i++) {
            if (x) x[i] = (double)sim.x[i];
            if (v) v[i] = (double)sim.v[i];
        }
        for (uint32_t i = 0; i < sim.n; i++) {
            if (rho) rho[i] = (double)sim.rho[i];
            if (near_rho) near_rho[i] = (double)sim.near_rho[i];
            if (pid) pid[i] = sim.pid[i];
        }
    };
    if (s->fp64) get(s->s64); else get(s->s32);
}

// Kernel values of the FP32 model (spiky2, spiky3, grad_spiky2, grad_spiky3, poly6) at d.
void orc_lague_kernels(void* h, double d, double* out5) {
    auto& m = ((OrcLague*)h)->s32;
    const float f = (float)d;
    out5[0] = m.spiky2(f);
    out5[1] = m.spiky3(f);
    out5[2] = m.grad_spiky2(f);
    out5[3] = m.grad_spiky3(f);
    out5[4] = m.poly6_w(f);
}

int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
