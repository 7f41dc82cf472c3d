// Test helpers of the device arithmetic in wcsph_math.h, run on the GPU.
// rcp_check compares the force pass's reciprocal (rcp_rn) with the IEEE
// division 1.0f / x for every float bit pattern in a range; mode 1 compares
// the bare hardware reciprocal instead, which differs from the division in
// many places (the comparison can fail).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../vkfluidsim_amd/csrc/wcsph_math.h"

namespace {

__global__ void k_rcp_check(uint32_t lo, uint32_t count, int mode, float one, unsigned long long* bad,
                            uint32_t* first) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const uint32_t bits = lo + t;
    const float x = __uint_as_float(bits);
    const float ref = one / x;  // `one` is a kernel argument: a true division
    const float y = mode == 0 ? vfs::rcp_rn(x) : __builtin_amdgcn_rcpf(x);
    if (__float_as_uint(y) != __float_as_uint(ref)) {
        const unsigned long long k = atomicAdd(bad, 1ull);
        if (k == 0) *first = bits;
    }
}

}  // namespace

// Mismatches over the float bit patterns [lo_bits, hi_bits) (and the first one
// found, or 0); returns 0 or a HIP error code.
extern "C" int rcp_check(uint32_t lo_bits, uint32_t hi_bits, int mode, unsigned long long* mismatches,
                         uint32_t* first) {
    unsigned long long* d_bad = nullptr;
    uint32_t* d_first = nullptr;
    hipError_t e = hipMalloc(&d_bad, sizeof *d_bad);
    if (e == hipSuccess) e = hipMalloc(&d_first, sizeof *d_first);
    if (e == hipSuccess) e = hipMemset(d_bad, 0, sizeof *d_bad);
    if (e == hipSuccess) e = hipMemset(d_first, 0, sizeof *d_first);
    constexpr uint32_t kChunk = 1u << 26;
    for (uint32_t b = lo_bits; e == hipSuccess && b < hi_bits;) {
        const uint32_t n = hi_bits - b < kChunk ? hi_bits - b : kChunk;
        hipLaunchKernelGGL(k_rcp_check, dim3((n + 255u) / 256u), dim3(256), 0, 0, b, n, mode, 1.0f, d_bad, d_first);
        e = hipGetLastError();
        b += n;
    }
    if (e == hipSuccess) e = hipMemcpy(mismatches, d_bad, sizeof *d_bad, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(first, d_first, sizeof *first, hipMemcpyDeviceToHost);
    if (d_bad) (void)hipFree(d_bad);
    if (d_first) (void)hipFree(d_first);
    return (int)e;
}

