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
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
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
                }
            }
            density += m.rho0 * psi[i];
            rho[i] = density;
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
    };
    if (s->fp64)
        set(s->s64);
    else
        set(s->s32);
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

int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
