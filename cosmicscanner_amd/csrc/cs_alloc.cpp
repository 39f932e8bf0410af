// Device memory for DevBuf: a per-device cache of large blocks.
//
// A fresh hipMalloc of a multi-GB block costs about 0.1-0.2 s per GB on the MI355X (the driver
// clears new VRAM before handing it out), which dominated the gaps between a serving loop's
// device work: each admitted cohort copies its genotype rows (two layouts, up to ~20 GB) and each
// finished request allocates its result arrays (up to ~2 GB).  Blocks of at least kMinCached
// bytes are therefore rounded up to a size class (eight per octave) and, when released, kept for
// reuse instead of freed, up to a total of kCacheBytes per process (CSAMD_DEVICE_CACHE_BYTES
// overrides it; 0 disables the cache).  A request reuses the smallest cached block of its own or
// a slightly larger class.
//
// Reuse ordering: hipFree synchronises the device, so a released block could never be reused
// while its former owner's kernels still ran.  A cached block keeps that guarantee by a device
// synchronisation when it is handed out again.  When hipMalloc runs out of memory, the device's
// cached blocks are freed and the allocation is retried once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "cs_common.h"

namespace csamd {

namespace {

constexpr size_t kMinCached = (size_t)8 << 20;

struct Cached {
    int device;
    size_t cap;
    void* p;
};

struct Cache {
    std::mutex mu;
    std::vector<Cached> blocks;
    size_t total = 0;
    size_t limit = 0;
    Cache() {
        limit = (size_t)64 << 30;
        if (const char* env = std::getenv("CSAMD_DEVICE_CACHE_BYTES")) limit = std::strtoull(env, nullptr, 10);
    }
};

Cache& cache() {
    static Cache* c = new Cache();   // never destroyed: blocks stay valid until the process ends
    return *c;
}

// Size class of n >= kMinCached: the next (8 + i) 2^(e-3), i = 0..7, at or above n.
size_t size_class(size_t n) {
    int e = 63 - __builtin_clzll((unsigned long long)n);
    const size_t step = (size_t)1 << (e - 3);
    return (n + step - 1) / step * step;
}

void free_cached_on(int device, Cache& c) {
    // caller holds c.mu
    std::vector<Cached> keep;
    for (const Cached& b : c.blocks) {
        if (b.device == device) {
            (void)hipFree(b.p);
            c.total -= b.cap;
        } else {
            keep.push_back(b);
        }
    }
    c.blocks.swap(keep);
}

}  // namespace

void* dev_alloc(size_t n, size_t* cap) {
    if (n == 0) n = 1;
    Cache& c = cache();
    void* p = nullptr;
    if (n < kMinCached || c.limit == 0) {
        CS_HIP(hipMalloc(&p, n));
        *cap = n;
        return p;
    }
    const size_t want = size_class(n);
    int device = 0;
    CS_HIP(hipGetDevice(&device));
    {
        std::lock_guard<std::mutex> lk(c.mu);
        int best = -1;
        for (int i = 0; i < (int)c.blocks.size(); ++i) {
            const Cached& b = c.blocks[(size_t)i];
            if (b.device != device || b.cap < want || b.cap > want + want / 2) continue;
            if (best < 0 || b.cap < c.blocks[(size_t)best].cap) best = i;
        }
        if (best >= 0) {
            const Cached b = c.blocks[(size_t)best];
            c.blocks.erase(c.blocks.begin() + best);
            c.total -= b.cap;
            // the block's former owner may still have work in flight (hipFree would have waited)
            CS_HIP(hipDeviceSynchronize());
            *cap = b.cap;
            return b.p;
        }
    }
    hipError_t e = hipMalloc(&p, want);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        {
            std::lock_guard<std::mutex> lk(c.mu);
            free_cached_on(device, c);
        }
        e = hipMalloc(&p, want);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw RuntimeError(std::string("HIP error: ") + hipGetErrorString(e) + " at hipMalloc (" +
                           std::to_string(want) + " bytes)");
    }
    *cap = want;
    return p;
}

void dev_free(void* p, size_t cap) {
    if (!p) return;
    Cache& c = cache();
    if (cap < kMinCached || c.limit == 0) {
        (void)hipFree(p);
        return;
    }
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) {
        (void)hipFree(p);
        return;
    }
    std::lock_guard<std::mutex> lk(c.mu);
    // make room: the oldest cached blocks go first
    while (!c.blocks.empty() && c.total + cap > c.limit) {
        (void)hipFree(c.blocks.front().p);
        c.total -= c.blocks.front().cap;
        c.blocks.erase(c.blocks.begin());
    }
    if (c.total + cap > c.limit) {
        (void)hipFree(p);
        return;
    }
    c.blocks.push_back(Cached{device, cap, p});
    c.total += cap;
}

void dev_cache_release() {
    Cache& c = cache();
    std::lock_guard<std::mutex> lk(c.mu);
    if (c.blocks.empty()) return;
    (void)hipDeviceSynchronize();
    for (const Cached& b : c.blocks) (void)hipFree(b.p);
    c.blocks.clear();
    c.total = 0;
}

}  // namespace csamd
