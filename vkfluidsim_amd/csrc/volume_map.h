// Device side of the boundary-grid generators: the FP64 cubic spline, the
// 16-point Gauss-Legendre table and the volume-map integral of one grid node
// (reference src/util/volume_map.cpp:9-54, src/util/kernel.cpp:10-34,
// src/util/gaussian_quadrature.cpp:2293-2333), for any signed-distance functor
// `sdf(qx, qy, qz) -> double` of a float query point.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vfs {

struct GLTable {
    double x[16];
    double w[16];
};

// Cubic spline of src/util/kernel.cpp:10-34 (FP64).
__device__ __forceinline__ double kernel_w_f64(double r, double radius, double factor) {
    const double q = r / radius;
    if (q <= 1.0) {
        if (q <= 0.5) {
            const double q2 = q * q, q3 = q2 * q;
            return factor * (6.0 * q3 - 6.0 * q2 + 1.0);
        }
        return factor * (2.0 * pow(1.0 - q, 3.0));
    }
    return 0.0;
}

// Grid node position: step * (i, j, k) + domain minimum in double, then the
// float the reference's grid function receives (discretization.cpp:16-33).
__device__ __forceinline__ void grid_node(uint64_t idx, uint32_t rx, uint32_t ry, float d0, float d1, float d2,
                                          double s0, double s1, double s2, float* x) {
    const uint32_t i = (uint32_t)(idx % rx);
    const uint32_t j = (uint32_t)((idx / rx) % ry);
    const uint32_t k = (uint32_t)(idx / ((uint64_t)rx * ry));
    x[0] = (float)(s0 * (double)i + (double)d0);
    x[1] = (float)(s1 * (double)j + (double)d1);
    x[2] = (float)(s2 * (double)k + (double)d2);
}

// 0.8 * integral over the support ball of the boundary indicator (1 inside,
// W(d)/W(0) within one support radius outside, 0 beyond), with the early exit
// of volume_map.cpp:33-35 for nodes more than two radii outside.  `dist` is
// the signed distance at the node.  Summation order follows the reference's
// nested loops.
template <class SDF>
__device__ double volume_map_integral(const SDF& sdf, const float* x, double dist, float sr, const GLTable& gl) {
    if (dist > 2.0 * (double)sr) return 0.0;
    const double radius = (double)sr;
    const double factor = 8.0 / (M_PI * radius * radius * radius);
    const double w_zero = kernel_w_f64(0.0, radius, factor);
    const double c0 = 0.5 * (double)(sr - (-sr));
    const double c1 = 0.5 * (double)(sr + (-sr));
    const float sr2 = sr * sr;
    double integral = 0.0;
    for (int a = 0; a < 16; a++) {
        const double wi = gl.w[a];
        const float xa = (float)(c0 * gl.x[a] + c1);
        for (int bb = 0; bb < 16; bb++) {
            const double wij = wi * gl.w[bb];
            const float xb = (float)(c0 * gl.x[bb] + c1);
            for (int cc = 0; cc < 16; cc++) {
                const double wijk = wij * gl.w[cc];
                const float xc = (float)(c0 * gl.x[cc] + c1);
                double val;
                if ((xa * xa + xb * xb) + xc * xc > sr2) {
                    val = 0.0;
                } else {
                    const double di = sdf(x[0] + xa, x[1] + xb, x[2] + xc);
                    if (di <= 0.0)
                        val = 1.0;
                    else if (di < radius)
                        val = kernel_w_f64(di, radius, factor) / w_zero;
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

}  // namespace vfs
