// Device memory for DevBuf: a per-device cache of blocks.
//
// A fresh hipMalloc of a multi-GB block costs about 0.1-0.2 s per GB on the MI355X (the driver
// clears new VRAM before handing it out), which dominated the gaps between a serving loop's
// device work: each admitted cohort copies its genotype rows (two layouts, up to ~20 GB) and each
// finished request allocates its result arrays (up to ~2 GB).  Blocks of at least kMinCached
// bytes are therefore rounded up to a size class (eight per octave) and, when released, kept for
// reuse instead of freed, up to half of the device's memory per process (CSAMD_DEVICE_CACHE_BYTES
// sets the bound; 0 disables the cache).  A request reuses the smallest cached block of its own or
// a slightly larger class.  The bound is generous on purpose: a serving loop's cohorts and result
// arrays come in many sizes, and a smaller cache (64 GiB was measured) misses often enough that
// fresh allocations and evictions cost more than the products they feed.
//
// Reuse ordering.  hipFree synchronises the whole device, so a freed block could never be reused
// while its former owner's work still ran.  A block released in the ordinary way keeps that
// guarantee by a device synchronisation when it is handed out again ("busy" blocks).  Two scopes
// avoid the device-wide waits where the caller knows more:
//  - DeferFrees (a thread's section whose device work runs on one or two streams, plus streams
//    that the section's functions synchronise before they return): blocks released inside it, of
//    any size, are held back and enter the cache as "idle" once the section's streams have been
//    synchronised at its end.  An idle block has no work in flight and is handed out with no wait.
//  - IdleFrees: the caller states that no work in flight uses the blocks it releases (their
//    readers and writers have been waited for); they enter the cache as idle at once.
// Inside either scope, small blocks are cached as well (rounded to a class, at least kSmallMin;
// at most kSmallLimit of them in all, the oldest freed first).
// A device-wide wait (hipFree, hipDeviceSynchronize) waits for every stream's work, including a
// concurrent thread's copies, so removing them from a serving step keeps the device fed.
// When hipMalloc runs out of memory, the device's cached blocks are freed and the allocation is
// retried once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "cs_common.h"

namespace csamd {

namespace {

constexpr size_t kMinCached = (size_t)8 << 20;
constexpr size_t kSmallMin = 4096;
constexpr size_t kSmallLimit = (size_t)2 << 30;   // cached blocks below kMinCached, in all

struct Cached {
    int device;
    size_t cap;
    void* p;
    bool idle;   // no work in flight uses it: handed out without a device synchronisation
};

struct Cache {
    std::mutex mu;
    std::vector<Cached> blocks;   // oldest first
    size_t total = 0;
    size_t small_total = 0;
    size_t limit = 0;
    Cache() {
        size_t free_b = 0, total_b = 0;
        limit = hipMemGetInfo(&free_b, &total_b) == hipSuccess && total_b > 0 ? total_b / 2 : (size_t)64 << 30;
        (void)hipGetLastError();
        if (const char* env = std::getenv("CSAMD_DEVICE_CACHE_BYTES")) limit = std::strtoull(env, nullptr, 10);
    }
};

Cache& cache() {
    static Cache* c = new Cache();   // never destroyed: blocks stay valid until the process ends
    return *c;
}

// The calling thread's scopes: 0 none, 1 DeferFrees (releases held in `pending`), 2 IdleFrees.
struct ThreadScope {
    int mode = 0;
    int depth = 0;
    hipStream_t stream = nullptr, stream2 = nullptr;
    bool full_sync = false;
    std::vector<Cached> pending;
};
thread_local ThreadScope t_scope;

// Size class of n: the next (8 + i) 2^(e-3), i = 0..7, at or above n (n >= kSmallMin).
size_t size_class(size_t n) {
    if (n <= kSmallMin) return kSmallMin;
    int e = 63 - __builtin_clzll((unsigned long long)n);
    const size_t step = (size_t)1 << (e - 3);
    return (n + step - 1) / step * step;
}

// Remove block i from the cache's list and totals (caller holds c.mu).
Cached take(Cache& c, size_t i) {
    const Cached b = c.blocks[i];
    c.blocks.erase(c.blocks.begin() + (std::ptrdiff_t)i);
    c.total -= b.cap;
    if (b.cap < kMinCached) c.small_total -= b.cap;
    return b;
}

void free_cached_on(int device, Cache& c) {
    // caller holds c.mu; hipFree waits for the device, so busy blocks are safe to free
    for (size_t i = c.blocks.size(); i-- > 0;)
        if (c.blocks[i].device == device) (void)hipFree(take(c, i).p);
}

// Add a released block to the cache (caller holds c.mu); the oldest blocks make room (for a
// small block, the oldest small blocks).
void insert(Cache& c, const Cached& b) {
    const bool small = b.cap < kMinCached;
    while (small && c.small_total + b.cap > kSmallLimit) {
        size_t i = 0;
        while (i < c.blocks.size() && c.blocks[i].cap >= kMinCached) ++i;
        if (i == c.blocks.size()) break;
        (void)hipFree(take(c, i).p);
    }
    while (!c.blocks.empty() && c.total + b.cap > c.limit) (void)hipFree(take(c, 0).p);
    if (c.total + b.cap > c.limit || (small && c.small_total + b.cap > kSmallLimit)) {
        (void)hipFree(b.p);
        return;
    }
    c.blocks.push_back(b);
    c.total += b.cap;
    if (small) c.small_total += b.cap;
}

}  // namespace

void* dev_alloc(size_t n, size_t* cap) {
    if (n == 0) n = 1;
    Cache& c = cache();
    void* p = nullptr;
    const bool scoped = t_scope.mode != 0;
    if (c.limit == 0 || (n < kMinCached && !scoped)) {
        CS_HIP(hipMalloc(&p, n));
        *cap = n;
        return p;
    }
    const size_t want = size_class(n);
    int device = 0;
    CS_HIP(hipGetDevice(&device));
    {
        std::lock_guard<std::mutex> lk(c.mu);
        // an idle block first; a busy one costs a device synchronisation, which a small block
        // is not worth (hipMalloc of a small block is cheap).  A small request may take a block
        // up to four times its class, a large one up to one and a half.
        const size_t most = want < kMinCached ? 4 * want : want + want / 2;
        int best = -1;
        for (int pass = 0; pass < 2 && best < 0; ++pass) {
            if (pass == 1 && want < kMinCached) break;
            for (int i = 0; i < (int)c.blocks.size(); ++i) {
                const Cached& b = c.blocks[(size_t)i];
                if (b.device != device || b.idle != (pass == 0) || b.cap < want || b.cap > most) continue;
                if (best < 0 || b.cap < c.blocks[(size_t)best].cap) best = i;
            }
        }
        if (best >= 0) {
            const Cached b = take(c, (size_t)best);
            // a busy block's former owner may still have work in flight (hipFree would have waited)
            if (!b.idle) CS_HIP(hipDeviceSynchronize());
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
        // blocks this thread holds back are freed too (hipFree waits for their work)
        for (const Cached& b : t_scope.pending) (void)hipFree(b.p);
        t_scope.pending.clear();
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
    const int mode = t_scope.mode;
    // small blocks released outside a scope go back to the driver
    if (c.limit == 0 || (cap < kMinCached && mode == 0)) {
        (void)hipFree(p);
        return;
    }
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) {
        (void)hipFree(p);
        return;
    }
    if (mode == 1) {
        t_scope.pending.push_back(Cached{device, cap, p, true});
        return;
    }
    std::lock_guard<std::mutex> lk(c.mu);
    insert(c, Cached{device, cap, p, mode == 2});
}

void dev_cache_release() {
    Cache& c = cache();
    std::lock_guard<std::mutex> lk(c.mu);
    if (c.blocks.empty()) return;
    (void)hipDeviceSynchronize();
    for (const Cached& b : c.blocks) (void)hipFree(b.p);
    c.blocks.clear();
    c.total = 0;
    c.small_total = 0;
}

DeferFrees::DeferFrees(hipStream_t st, hipStream_t st2) {
    if (t_scope.depth++ > 0) return;
    t_scope.mode = 1;
    t_scope.stream = st;
    t_scope.stream2 = st2;
    t_scope.full_sync = false;
}

void DeferFrees::sync_device() { t_scope.full_sync = true; }

DeferFrees::~DeferFrees() {
    if (--t_scope.depth > 0) return;
    t_scope.mode = 0;
    std::vector<Cached> held;
    held.swap(t_scope.pending);
    if (held.empty()) return;
    // the section's work on its streams is complete once these return; a failed synchronisation,
    // or a section that asked for it (an error path may leave other streams' work in flight),
    // falls back to the device-wide wait
    bool ok = !t_scope.full_sync && hipStreamSynchronize(t_scope.stream) == hipSuccess &&
              (!t_scope.stream2 || hipStreamSynchronize(t_scope.stream2) == hipSuccess);
    if (!ok) {
        (void)hipGetLastError();
        (void)hipDeviceSynchronize();
    }
    Cache& c = cache();
    std::lock_guard<std::mutex> lk(c.mu);
    for (const Cached& b : held) insert(c, b);
}

IdleFrees::IdleFrees() : prev_(t_scope.mode) {
    if (t_scope.mode == 0) t_scope.mode = 2;
}

IdleFrees::~IdleFrees() { t_scope.mode = prev_; }

}  // namespace csamd
