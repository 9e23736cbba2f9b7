// Device scratch of gbi_execute_noalloc: one file-scope array per GPU, held from the library's
// load until the process ends, from which that entry point carves every device buffer of an
// evaluation (inputs, tables, lookup grid, point records, partial sums, outputs), so it calls
// no allocator.
//
// In host code the name `scratch` is only a stand-in for the symbol: its device address is
// looked up once per device, with that device current, and only that address is ever used.
#include "gbi_internal.h"

#include <hip/hip_runtime.h>

#include <mutex>

namespace gbi {

namespace {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

hipError_t noalloc_scratch(void** ptr) {
    *ptr = nullptr;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    return hipSuccess;
}

}  // namespace gbi
