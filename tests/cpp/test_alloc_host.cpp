// Host test of the device-memory cache (cosmicscanner_amd/csrc/cs_alloc.cpp) against the stand-in
// runtime in hip_host_stub/: several threads allocate, grow, and release DevBufs of random sizes
// (small and large blocks) inside and outside DeferFrees / IdleFrees sections, every byte asked
// for is written, and hipMalloc is optionally capped so that the out-of-memory retry runs.  The
// sanitizers report any use of a released block or a double release; at the end every block must
// have gone back to the runtime after dev_cache_release.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "cs_common.h"

std::atomic<long> g_stub_live_bytes{0};
long g_stub_capacity = 0;

namespace csamd {
void set_error(const std::string&) {}
void clear_error() {}
}  // namespace csamd

using csamd::DevBuf;

static size_t pick_size(std::mt19937_64& rng) {
    switch (rng() % 4) {
        case 0: return 1 + rng() % 5000;
        case 1: return 1 + rng() % (1 << 20);
        case 2: return ((size_t)1 << 20) + rng() % ((size_t)16 << 20);
        default: return ((size_t)8 << 20) + rng() % ((size_t)40 << 20);
    }
}

static void worker(int seed, int rounds, long* failures) {
    std::mt19937_64 rng((unsigned long long)seed);
    std::vector<DevBuf> held;
    for (int it = 0; it < rounds; ++it) {
        auto body = [&] {
            const int ops = 1 + (int)(rng() % 40);
            for (int o = 0; o < ops; ++o) {
                const int r = (int)(rng() % 10);
                if (r < 5 || held.empty()) {
                    const size_t n = pick_size(rng);
                    held.emplace_back(n);
                    std::memset(held.back().p, 0xab, n);
                } else if (r < 8) {
                    held.erase(held.begin() + (std::ptrdiff_t)(rng() % held.size()));
                } else {
                    DevBuf& b = held[rng() % held.size()];
                    const size_t n = 1 + rng() % ((size_t)24 << 20);
                    b.ensure(n);
                    std::memset(b.p, 0xcd, n);
                }
            }
        };
        try {
            switch (rng() % 3) {
                case 0: body(); break;
                case 1: {
                    csamd::DeferFrees d(nullptr);
                    if (rng() % 7 == 0) d.sync_device();
                    body();
                    break;
                }
                default: {
                    csamd::IdleFrees f;
                    body();
                }
            }
        } catch (const csamd::RuntimeError&) {
            ++*failures;   // out of (capped) memory: let go of everything and go on
            held.clear();
        }
        if (held.size() > 64) held.resize(32);
    }
}

// The reuse rule of a DeferFrees section: a block released inside it is not handed out again
// before the section ends (its work may still be in flight), and is reused after it (small blocks
// only by allocations inside a section, as the next step of a serving loop makes them).
static int check_deferred_reuse() {
    int bad = 0;
    for (size_t n : {(size_t)3000, (size_t)1 << 20, (size_t)12 << 20}) {
        void* first = nullptr;
        {
            csamd::DeferFrees d(nullptr);
            DevBuf a(n);
            first = a.p;
            a.release();
            DevBuf b(n);   // same size class: must not be the block released in this section
            if (b.p == first) {
                std::printf("block of %zu bytes reused inside its DeferFrees section\n", n);
                ++bad;
            }
        }
        {
            csamd::DeferFrees next(nullptr);
            DevBuf c(n), e(n);   // after the section the two held blocks are idle and come back
            if (c.p != first && e.p != first) {
                std::printf("block of %zu bytes not reused after its DeferFrees section\n", n);
                ++bad;
            }
        }
    }
    return bad;
}

int main(int argc, char** argv) {
    if (check_deferred_reuse() != 0) return 2;
    const int threads = argc > 1 ? std::atoi(argv[1]) : 4;
    const int rounds = argc > 2 ? std::atoi(argv[2]) : 200;
    if (argc > 3) g_stub_capacity = std::atol(argv[3]);
    std::vector<long> failures((size_t)threads, 0);
    std::vector<std::thread> ts;
    for (int t = 0; t < threads; ++t) ts.emplace_back(worker, 100 + t, rounds, &failures[(size_t)t]);
    for (auto& t : ts) t.join();
    csamd::dev_cache_release();
    long f = 0;
    for (long x : failures) f += x;
    std::printf("live bytes after release: %ld, allocation failures: %ld\n", g_stub_live_bytes.load(), f);
    return g_stub_live_bytes.load() == 0 ? 0 : 1;
}
