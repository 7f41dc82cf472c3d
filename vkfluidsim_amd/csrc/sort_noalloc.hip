// vfs_sort_pairs_noalloc: the stable key sort of vfs_sort_pairs with its
// workspace in a fixed device array of this library instead of a caller- or
// library-allocated buffer, for callers that cannot allocate device memory
// (code under stream capture, serving loops, harnesses that forbid
// hipMalloc).  No counterpart in the reference.
//
// The array, `scratch`, lives in every GPU the library runs on until the
// process ends.  Its address on the current device is taken once per device
// with hipGetSymbolAddress and used for every kernel argument; the array's
// name itself is a host-side stand-in and never reaches a kernel.  All calls
// on one GPU share that GPU's array: calls that overlap on one GPU (two
// streams, two host threads) race and must be serialised by the caller.  The
// sort's preparation kernel clears the words that must start at zero
// (histograms, tile counters, look-back status) on every call, so nothing
// left by an earlier call is read.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "sort_internal.h"

namespace {

// Up to 2^20 keys with any max_value: four 8-bit passes need ping-pong keys
// and values (4 n words), digit histograms, tile counters and look-back
// status (4 * 256 * 256 words), the cell-start rows of the plan (~600 words)
// and the sorted keys and permutation before they are copied back (2 n
// words): 6,555,392 words.  26 MiB are reserved.
constexpr uint32_t kNoAllocMaxKeys = 1u << 20;
constexpr size_t kScratchWords = (size_t)26 << 18;

__device__ __align__(256) uint32_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
uint32_t* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, taken on the first call
// per device.
int scratch_address(uint32_t** out) {
    int dev = -1;
    VFS_HIP(hipGetDevice(&dev));
    VFS_CHECK(dev >= 0 && dev < kMaxDevices, VFS_ERR_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        VFS_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        VFS_CHECK(ptr, VFS_ERR_HIP, "hipGetSymbolAddress returned a null address");
        g_scratch_ptr[dev] = static_cast<uint32_t*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

}  // namespace

using namespace vfs;

extern "C" {

int vfs_sort_pairs_noalloc(uint32_t* items_dev, uint32_t* keys_dev, uint32_t n, uint32_t max_value, void* stream) {
    VFS_CHECK(n <= kNoAllocMaxKeys, VFS_ERR_INVALID,
              "n = " + std::to_string(n) + " exceeds the " + std::to_string(kNoAllocMaxKeys) +
                  " keys of the fixed sort scratch (use vfs_sort_pairs with a workspace)");
    if (n == 0) return 0;
    VFS_CHECK(items_dev && keys_dev, VFS_ERR_INVALID, "null argument");
    const SortPlan p = make_sort_plan(n, max_value, 8, 16, false);
    const size_t need = p.workspace_words + 2 * (size_t)n;
    VFS_CHECK(need <= kScratchWords, VFS_ERR_INVALID,
              "sort plan needs " + std::to_string(need * 4) + " bytes, more than the fixed scratch");
    uint32_t* ws = nullptr;
    VFS_TRY(scratch_address(&ws));
    uint32_t* out_keys = ws + p.workspace_words;
    uint32_t* out_items = out_keys + n;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    VFS_TRY(sort_keys_iota(p, keys_dev, out_keys, out_items, ws, s, nullptr));
    VFS_TRY(launch_copy_words(out_keys, keys_dev, n, s));
    VFS_TRY(launch_copy_words(out_items, items_dev, n, s));
    return 0;
}

}  // extern "C"
