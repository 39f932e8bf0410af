// Host-only stand-in for the HIP runtime calls that cs_alloc.cpp makes, so that the device-memory
// cache's bookkeeping (size classes, idle/busy blocks, deferred releases, limits, the
// out-of-memory retry) can be tested on a machine without a GPU under the address and
// undefined-behaviour sanitizers.  "Device" blocks are host heap blocks with a size header; the
// synchronisation calls do nothing.  Not used by the library build.
#pragma once
#include <atomic>
#include <cstddef>
#include <cstdlib>

typedef int hipError_t;
typedef void* hipStream_t;
enum { hipSuccess = 0, hipErrorOutOfMemory = 2 };

extern std::atomic<long> g_stub_live_bytes;   // bytes held by the stand-in allocator
extern long g_stub_capacity;                  // > 0: hipMalloc fails beyond this many live bytes

inline hipError_t hipMalloc(void** p, size_t n) {
    if (g_stub_capacity > 0 && g_stub_live_bytes.load() + (long)n > g_stub_capacity) return hipErrorOutOfMemory;
    char* b = static_cast<char*>(std::malloc(n + 16));
    if (!b) return hipErrorOutOfMemory;
    *reinterpret_cast<size_t*>(b) = n;
    g_stub_live_bytes += (long)n;
    *p = b + 16;
    return hipSuccess;
}
inline hipError_t hipFree(void* p) {
    if (!p) return hipSuccess;
    char* b = static_cast<char*>(p) - 16;
    g_stub_live_bytes -= (long)*reinterpret_cast<size_t*>(b);
    std::free(b);
    return hipSuccess;
}
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipGetDevice(int* d) {
    *d = 0;
    return hipSuccess;
}
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipMemGetInfo(size_t* free_b, size_t* total_b) {
    *total_b = (size_t)4 << 30;   // a small "device": the cache bound (half) is reached in the test
    *free_b = *total_b;
    return hipSuccess;
}
inline const char* hipGetErrorString(hipError_t e) { return e == hipErrorOutOfMemory ? "out of memory" : "error"; }
