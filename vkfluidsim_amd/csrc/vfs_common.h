// Shared definitions of the gfx950 WCSPH substep engine: error plumbing,
// device-side arithmetic helpers and the launch-parameter structs.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/vfs_c.h"

namespace vfs {

// ------------------------------------------------------------ error plumbing
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define VFS_HIP(call)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return ::vfs::fail(VFS_ERR_HIP, std::string(#call " failed: ") +              \
                                                 hipGetErrorString(e_));                    \
    } while (0)

#define VFS_CHECK(cond, code, msg)                   \
    do {                                             \
        if (!(cond)) return ::vfs::fail(code, msg); \
    } while (0)

#define VFS_TRY(expr)            \
    do {                         \
        int r_ = (expr);         \
        if (r_ != 0) return r_;  \
    } while (0)

// Kernel launches report launch errors through hipGetLastError.
int check_launch(const char* what);

// ------------------------------------------------------------- integer math
// a % d for every 32-bit a and d >= 1 (Lemire, Kaser, Kurz, "Faster remainder
// by direct computation", 2019): M = floor((2^64 - 1) / d) + 1.
struct FastMod {
    uint64_t M;
    uint32_t d;
};

inline FastMod make_fastmod(uint32_t d) {
    FastMod f;
    f.d = d;
    f.M = d ? (UINT64_C(0xFFFFFFFFFFFFFFFF) / d + 1) : 0;
    return f;
}

__device__ __forceinline__ uint32_t fastmod(uint32_t a, FastMod f) {
    const uint64_t low = f.M * (uint64_t)a;
    return (uint32_t)__umul64hi(low, (uint64_t)f.d);
}

// HashCell3D (reference assets/shaders/simulation/spatial_hash/spatial_hash_3d.slang:20-28):
// 50^3 blocks of local cells, block ids mixed with three primes, all mod 2^32.
__device__ __forceinline__ uint32_t hash_cell(int32_t cx, int32_t cy, int32_t cz) {
    const uint32_t ux = (uint32_t)cx + 25u, uy = (uint32_t)cy + 25u, uz = (uint32_t)cz + 25u;
    const uint32_t bx = ux / 50u, by = uy / 50u, bz = uz / 50u;
    const uint32_t lx = ux - 50u * bx, ly = uy - 50u * by, lz = uz - 50u * bz;
    return lx + 50u * (ly + 50u * lz) + (bx * 15823u + by * 9737333u + bz * 440817757u);
}

// (int)floor(v) with saturation and NaN -> 0 (the shader cast is undefined there).
__device__ __forceinline__ int32_t floor_to_cell(float v) {
    const float f = floorf(v);
    if (!(f == f)) return 0;
    if (f <= -2147483648.0f) return INT32_MIN;
    if (f >= 2147483648.0f) return INT32_MAX;
    return (int32_t)f;
}

// --------------------------------------------------------- model constants
// Everything the substep kernels read that is uniform over particles.
struct SimConsts {
    float gx, gy, gz;
    float h, h2, inv_h;    // inv_h = 1 / h (IEEE)
    float cell;            // hash cell size
    float rho0, B, gamma;
    int32_t gamma_int;     // >= 0: integral EOS exponent, -1: general power
    float nu10;            // 2*(3+2)*viscosity (ViscousAccel prefactor)
    float kwall_neg;       // -wall_stiffness
    float dt, mass;
    float sigma;           // cubic spline scale 8/(pi h^3)
    float grad_sigma;      // grad scale 48/(pi h^4) divided by h once more
    float visc_eps_h2;     // 0.01 * h^2
    float rho0_sq;         // rho0 * rho0
    float bb_lo[3], bb_hi[3];
    uint32_t n;            // particles of the scene
    uint32_t nbr_cap;      // neighbour-list slots per particle
    FastMod fm;            // % n
    // Scene segment inside the particle arrays (batched scenes; 0 for one
    // scene): the scene's particles are entries [base, base + n) and its hash
    // keys are base + (hash % n).
    uint32_t base;
    uint32_t obj_begin, obj_count;  // its boundary objects in the object array
    float nu;              // viscosity_strenght (the adjoint's FP64 prefactor 10 nu)
};

// Constants of the Lague (Clavet double-density) model (lague_model.slang,
// kernels_3d.slang:16-60, CalcKernelCoefficients model.cpp:224-233).
struct LagueConsts {
    float gx, gy, gz;
    float h, h2;
    float dt;
    float pf;              // prediction factor 1/120
    float rho0;            // target density
    float pm, npm;         // pressure multipliers
    float visc, damp;      // viscosity, wall damping
    float sp2, sp3;        // SpikyPow2 / SpikyPow3 scales
    float dsp2, dsp3;      // their derivative scales
    float poly6;           // Poly6 scale
    float bb_lo[3], bb_hi[3];
};

// Per-wave scene lookup of batched launches: a batch pads every scene to a
// multiple of 64 entries so that a wave never spans two scenes.
struct SceneTable {
    const SimConsts* consts;      // [n_scenes + 1]; the last entry has n = 0
    const uint32_t* wave_scene;   // scene of each wave of the padded launch
};

// Neighbour-list slots per particle, a multiple of 4 (the lattice workloads
// have ~30 neighbours).
constexpr uint32_t kNeighbourCap = VFS_NEIGHBOUR_CAPACITY;
static_assert(kNeighbourCap % 4 == 0, "lists are stored in chunks of four");
// uint32 words of list storage per particle: kNeighbourCap / 4 chunk rows, then
// one more row that takes the chunk stores of a list past the capacity (never
// read), so that the stores need no branch
constexpr uint32_t kNeighbourWords = kNeighbourCap + 4u;

// Boundary object as the kernels see it (column-major matrices).
struct DevObject {
    float T[16];
    float Rm[16];
    float box_pos[3];
    float box_size[3];
    uint32_t res[3];
    uint32_t pad;
    const float* sdf;
    const float* vm;
};

// Host computation of SimConsts from vfs_params, with the reference's own
// float expressions (model.cpp:225-233, wcsph_with_boundary_model.cpp:84-89,
// wcsph_with_boundary_model.slang:39-41).
SimConsts make_consts(const vfs_params& p);
LagueConsts make_lague_consts(const vfs_params& p, const vfs_lague_params& lp);

}  // namespace vfs
