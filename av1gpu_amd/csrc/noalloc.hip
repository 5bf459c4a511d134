// Device scratch of av1mi_encode_noalloc: one file-scope array per device, reserved when the library's code object is
// loaded on that device and held until the process ends.  Host code never passes the array's name to a kernel or a copy
// (on the host it is only a stand-in for the symbol): noalloc_scratch() resolves each device's own address once, with
// that device current, and every later use goes through that address.
#include <hip/hip_runtime.h>
#include <mutex>
#include "../../include/av1mi.h"

namespace av1m {

constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)AV1MI_NOALLOC_SCRATCH_MIB << 20;

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_lock;
void *g_ptr[MAX_DEVICES];
}  // namespace

// The calling thread's current device's copy of the scratch array (resolved once per device): 0 with *ptr and *bytes
// set, or the HIP error.
hipError_t noalloc_scratch(void **ptr, size_t *bytes) {
    *ptr = nullptr;
    *bytes = 0;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_lock);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_ptr[dev] = p;
    }
    *ptr = g_ptr[dev];
    *bytes = NOALLOC_SCRATCH_BYTES;
    return hipSuccess;
}

}  // namespace av1m
