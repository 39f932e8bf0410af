// Shared definitions of the MI355X genotype library: error state, device handles, layouts.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cosmicscanner_amd.h"

namespace csamd {

// ---------------------------------------------------------------------------------------------
// Errors: every C entry point catches these and turns them into a status code plus a
// thread-local message (cs_last_error).  The C++ adapter rethrows with the same type and text.
struct RuntimeError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct InvalidArgument : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};

void set_error(const std::string& msg);
void clear_error();

#define CS_HIP(call)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            throw ::csamd::RuntimeError(std::string("HIP error: ") + hipGetErrorString(e_) + \
                                        " at " #call);                                     \
    } while (0)

template <class F>
int guarded(F&& f) {
    try {
        clear_error();
        f();
        return CS_OK;
    } catch (const InvalidArgument& e) {
        set_error(e.what());
        return CS_ERR_INVALID_ARGUMENT;
    } catch (const RuntimeError& e) {
        set_error(e.what());
        return CS_ERR_RUNTIME;
    } catch (const std::bad_alloc&) {
        set_error("out of host memory");
        return CS_ERR_RUNTIME;
    } catch (const std::exception& e) {
        set_error(e.what());
        return CS_ERR_RUNTIME;
    }
}

inline int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

// ---------------------------------------------------------------------------------------------
// Device memory with a per-device cache of blocks (cs_alloc.cpp): dev_alloc returns a block of at
// least n bytes and its capacity, dev_free takes it back (large blocks, and inside the scopes
// below all blocks, are kept for reuse), dev_cache_release frees every cached block.
void* dev_alloc(size_t n, size_t* cap);
void dev_free(void* p, size_t cap);
void dev_cache_release();

// A section of the calling thread whose device work runs on `st` and `st2` (optional), plus other
// streams that the section's functions synchronise before they return.  Blocks released inside
// it are held back and enter the cache at its end, after both streams have been synchronised, as
// idle blocks that are reused with no device-wide wait.  sync_device() makes the end wait for the
// whole device instead (for error paths, which may leave other streams' work in flight).
// Sections nest; the outermost one ends it.
class DeferFrees {
public:
    explicit DeferFrees(hipStream_t st, hipStream_t st2 = nullptr);
    ~DeferFrees();
    void sync_device();
    DeferFrees(const DeferFrees&) = delete;
    DeferFrees& operator=(const DeferFrees&) = delete;
};

// A section of the calling thread that releases only blocks no work in flight uses (their
// kernels and copies have been waited for): they enter the cache as idle blocks at once.
class IdleFrees {
public:
    IdleFrees();
    ~IdleFrees();
    IdleFrees(const IdleFrees&) = delete;
    IdleFrees& operator=(const IdleFrees&) = delete;

private:
    int prev_;
};

// Device buffer (RAII over dev_alloc).  `bytes` is the size asked for; `cap` the block's capacity.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    size_t cap = 0;
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes), cap(o.cap) {
        o.p = nullptr;
        o.bytes = 0;
        o.cap = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            cap = o.cap;
            o.p = nullptr;
            o.bytes = 0;
            o.cap = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t n) {
        release();
        if (n == 0) n = 1;
        p = dev_alloc(n, &cap);
        bytes = n;
    }
    void ensure(size_t n) {
        if (n <= bytes) return;
        if (p && n <= cap) {
            bytes = n;
            return;
        }
        alloc(n);
    }
    void release() {
        if (p) dev_free(p, cap);
        p = nullptr;
        bytes = 0;
        cap = 0;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Padding granules.  Samples and SNPs are both padded to multiples of 256 so that every K block
// of the MFMA kernels (256 packed elements = 64 bytes) and every row tile is whole; padding
// slots hold code 01 (missing), so they contribute to no sum.
constexpr int64_t kPad = 256;
constexpr uint8_t kMissingByte = 0x55;

}  // namespace csamd

// The opaque handles of the C-ABI.
struct cs_bed {
    int device = 0;
    int64_t N = 0, M = 0;      // samples (.fam rows), SNPs (.bim rows)
    int64_t Braw = 0;          // ceil(N/4): bytes per SNP in the file
    int64_t Npad = 0, Mpad = 0;
    int64_t KBs = 0;           // SNP-major copy: 256-sample blocks per row tile (= Npad/256)
    int64_t KBm = 0;           // sample-major copy: 256-SNP blocks per row tile (= Mpad/256)
    csamd::DevBuf snp;         // Mpad x Npad genotypes, SNP-major, MFMA tiles (cs_geno_mfma.h)
    csamd::DevBuf smp;         // Npad x Mpad genotypes, sample-major, MFMA tiles (on demand)
    csamd::DevBuf counts;      // M x 4 int64: all-sample class counts per code
    csamd::DevBuf means;       // M double  (PlinkReader::getMeans)
    csamd::DevBuf inv_stds;    // M double  (PlinkReader::getInvStds)
    bool have_stats = false;
    // Missing genotypes as CSR lists (built on demand): per SNP the missing samples, per sample
    // the missing SNPs, both in increasing order; real rows and columns only.
    csamd::DevBuf miss_snp_off, miss_snp_idx, miss_smp_off, miss_smp_idx;
    int64_t miss_nnz = -1;     // -1: lists not built
    int64_t miss_total = -1;   // missing genotypes from the class counts; -1: not yet summed
    hipStream_t stream = nullptr;
};
