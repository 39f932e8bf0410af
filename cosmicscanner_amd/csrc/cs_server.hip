// Serving loop for scan requests over a resident genotype matrix.
//
// Because K's standardisation uses all-sample statistics (kinship_operator.cpp:64-67,
// plink_reader.cpp:187-239), the kinship matrix of a cohort S is the restriction P_S K P_S^T.
// Each admitted request owns a kinship operator over its cohort.  When the cohort is at most 3/4
// of the samples and the HBM budget holds it, the operator's products run on a compact copy of the
// cohort's genotype rows (gathered once at admission), so a K.V apply reads and multiplies n_S
// instead of N sample rows; otherwise they run on the full matrix through the operator's keep
// list.  A request's columns (its traits, covariates and calibration markers) form one multi-RHS
// PCG over n_S rows, each column with its own recurrence and convergence; one server step
// advances every active request by one iteration (one K.V apply each).  New requests are
// admitted at step boundaries as column capacity and the HBM budget free up (continuous
// batching).  A request whose columns have all converged runs its projection, calibration and
// fused per-SNP scan in the next step and gives its solve buffers back; the SNPs its exact stage selects
// then return as a second phase of columns (their prepared markers, in chunks that fit the free
// capacity and budget), whose solutions give the exact GLS variances.  The budget covers a
// request's compact copy, operator lists, preconditioner and projection from admission to
// completion, its solve buffers and operator workspace while held, and its result arrays from
// admission until the client releases them.
//
// A step first queues the iterations on the server stream, then does the host-heavy work of
// admissions and finished solves (right-hand sides, row copies, projections, scans, exact-stage
// chunks) with its device work on a second stream, so that this work and its host waits overlap
// the iterations instead of leaving the device idle; the step ends with both streams waited for.
//
// A column's arithmetic depends only on its own request (per-column quantisation scales, exact
// integer products, per-column dots with a fixed reduction order, and an operator whose choice of
// compact rows depends only on the cohort and the server's budget), so a request's results are
// the same alone and in any mix.
#include <rocblas/rocblas.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <map>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "cs_common.h"
#include "cs_ops.h"

namespace csamd {

#define CS_BLAS4(call)                                                                 \
    do {                                                                               \
        rocblas_status s_ = (call);                                                    \
        if (s_ != rocblas_status_success)                                              \
            throw ::csamd::RuntimeError(std::string("rocBLAS error ") +                \
                                        rocblas_status_to_string(s_) + " at " #call);  \
    } while (0)

constexpr int kSrvDotBlocks = 128;
// Pinned staging for results delivered to the host: two buffers of this size.
constexpr size_t kStageBytes = (size_t)64 << 20;
// Cohorts up to this fraction of the samples run on a compact copy of their rows.
constexpr double kCompactFrac = 0.75;

// partial[c * kSrvDotBlocks + blk] = sum over the block's rows of A[:, c] B[:, c]
__global__ __launch_bounds__(256) void k_srv_dots(const double* __restrict__ A, const double* __restrict__ B,
                                                  int64_t n, double* __restrict__ partial) {
    __shared__ double sh[256];
    const int64_t c = blockIdx.y;
    const int64_t chunk = (n + kSrvDotBlocks - 1) / kSrvDotBlocks;
    const int64_t r0 = blockIdx.x * chunk, r1 = min(n, r0 + chunk);
    double s = 0.0;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) s += A[c * n + r] * B[c * n + r];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[c * kSrvDotBlocks + blockIdx.x] = sh[0];
}

__global__ void k_srv_reduce(const double* __restrict__ partial, int64_t k, double* __restrict__ out) {
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= k) return;
    double s = 0.0;
    for (int b = 0; b < kSrvDotBlocks; ++b) s += partial[c * kSrvDotBlocks + b];
    out[c] = s;
}

// AP = sigma_g K P + sigma_e P
__global__ void k_srv_apply(const double* __restrict__ KP, const double* __restrict__ P, double sg, double se,
                            int64_t total, double* __restrict__ AP) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x)
        AP[e] = sg * KP[e] + se * P[e];
}

// PCG init of k columns: X = 0, R = B, Z = Minv R, P = Z (Minv: one diagonal for all columns)
__global__ void k_srv_init(const double* __restrict__ Bm, const double* __restrict__ minv, int64_t n, int64_t k,
                           double* __restrict__ X, double* __restrict__ R, double* __restrict__ Z,
                           double* __restrict__ P) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * k;
         e += (int64_t)gridDim.x * blockDim.x) {
        const double r = Bm[e];
        X[e] = 0.0;
        R[e] = r;
        const double z = minv[e % n] * r;
        Z[e] = z;
        P[e] = z;
    }
}

__global__ void k_srv_step(double* __restrict__ X, double* __restrict__ R, double* __restrict__ Z,
                           const double* __restrict__ P, const double* __restrict__ AP,
                           const double* __restrict__ minv, const double* __restrict__ rz,
                           const double* __restrict__ pap, const int32_t* __restrict__ active, int64_t n,
                           int64_t k) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * k;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = e / n;
        if (!active[c]) continue;
        const double alpha = rz[c] / pap[c];
        X[e] += alpha * P[e];
        const double r = R[e] - alpha * AP[e];
        R[e] = r;
        Z[e] = minv[e - c * n] * r;
    }
}

__global__ void k_srv_dir(double* __restrict__ P, const double* __restrict__ Z, const double* __restrict__ rz_new,
                          const double* __restrict__ rz_old, const int32_t* __restrict__ active, int64_t n,
                          int64_t k) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * k;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = e / n;
        if (!active[c]) continue;
        P[e] = Z[e] + (rz_new[c] / rz_old[c]) * P[e];
    }
}

// rz <- rz_new on the active columns
__global__ void k_srv_rz_keep(double* __restrict__ rz, const double* __restrict__ rz_new,
                              const int32_t* __restrict__ active, int64_t k) {
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c < k && active[c]) rz[c] = rz_new[c];
}

}  // namespace csamd

using namespace csamd;

namespace {

enum { REQ_PENDING = 0, REQ_ACTIVE = 1, REQ_DONE = 2, REQ_FAILED = -1 };

// PCG state of the columns a request holds (n x cols each, column-major).
// Pinned host memory for a request's per-iteration exchange with the device (the squared
// residual norms read back, the active-column flags sent), so both copies stay asynchronous.
struct Pinned {
    void* p = nullptr;
    Pinned() = default;
    Pinned(const Pinned&) = delete;
    Pinned& operator=(const Pinned&) = delete;
    Pinned& operator=(Pinned&& o) noexcept {
        if (this != &o) {
            if (p) (void)hipHostFree(p);
            p = o.p;
            o.p = nullptr;
        }
        return *this;
    }
    ~Pinned() {
        if (p) (void)hipHostFree(p);
    }
};

struct Solve {
    int64_t cols = 0;
    int64_t bytes = 0;   // charged to the budget while held (with the operator's workspace)
    DevBuf X, R, Z, P, KP, AP, partial, rz, rzn, pap, rr, bb, act;
    std::vector<double> hbb, hrr;
    std::vector<int32_t> hact;
    Pinned pin;          // cols doubles (residual norms), then cols int32 (active flags)
    size_t pin_cap = 0;  // its capacity (the block comes from, and goes back to, the server's pool)
    bool act_dirty = true;
    double* pin_rr() const { return static_cast<double*>(pin.p); }
    int32_t* pin_act() const { return reinterpret_cast<int32_t*>(static_cast<double*>(pin.p) + cols); }
};

struct Request {
    int64_t id = 0;
    int status = REQ_PENDING;
    std::vector<int32_t> keep;     // the cohort in the caller's order
    int64_t n = 0, T = 0, c = 0, L = 0, k = 0;
    std::vector<double> Y;
    std::vector<float> X;
    double sg = 0, se = 1;
    std::vector<int32_t> cal, snps;
    double min_maf = 0.01, tol = 1e-6, thr = kPscanExactDefault;   // effective (pscan_threshold)
    int max_iter = 200, prec = CS_PREC_INT8_L3;
    // phase 1: the main solve over [Y, X, G_cal]; phase 2: exact-stage chunks of selected markers
    int phase = 1;
    cs_kinship* K = nullptr;       // the request's operator (admission to completion)
    bool compact = false;          // its products run on a compact copy of the cohort's rows
    DevBuf B;                      // right-hand sides of the columns held (n x cols)
    DevBuf minv;                   // Jacobi preconditioner 1 / (sigma_g diag(K) + sigma_e), n
    Solve sv;
    int iterations = 0, chunk_iterations = 0, exact_iterations = 0;
    int64_t hold_bytes = 0;        // compact copy, lists, preconditioner, projection: admission to completion
    int64_t admit_bytes = 0;       // hold + solve buffers + workspace at admission (results excluded)
    int64_t result_bytes = 0;      // result arrays, charged from admission until release
    double t_submit = 0, t_admit = 0, t_done = 0;
    // projection state between the phases, selected rows, next chunk
    PscanProj pj;
    std::vector<int64_t> rows;
    int64_t row_next = 0, chunk = 0;
    // results, in device memory until cs_server_result copies them out (or the request is released)
    int64_t ns = 0;
    DevBuf freq, n_obs, pass, beta, se_out, p, log10p;
    std::string error;
    int fetching = 0;   // cs_server_result calls copying this request's arrays (guarded by mu)

    ~Request() {
        if (K) (void)cs_kinship_destroy(K);
    }
};

}  // namespace

struct cs_server {
    cs_bed* bed = nullptr;
    int64_t N = 0, cap = 0, used = 0;
    int64_t budget = 0, budget_used = 0;
    std::vector<int32_t> variants;
    hipStream_t st = nullptr;     // the requests' iterations
    hipStream_t side = nullptr;   // admissions and finished solves, beside the iterations
    rocblas_handle blas = nullptr;   // on the side stream
    // pinned blocks for the iterations' residual readbacks, reused across requests (pinned memory
    // is slow to allocate and free, and freeing it waits for the device)
    std::vector<std::pair<size_t, void*>> pin_pool;
    std::vector<double> diag;
    std::deque<int64_t> pending;
    std::map<int64_t, std::unique_ptr<Request>> reqs;
    int64_t next_id = 1;
    std::chrono::steady_clock::time_point t0;
    void* stage[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    // mu guards the scheduler state (requests, budget, columns); cs_server_result holds it only to
    // read a request and takes dmu for the copy itself, on its own stream, so one client thread
    // can fetch finished results while another runs the steps.
    std::mutex mu, dmu;
    hipStream_t dst = nullptr;   // result delivery
    ~cs_server() {
        if (st) (void)hipStreamSynchronize(st);
        if (side) (void)hipStreamSynchronize(side);
        if (dst) (void)hipStreamSynchronize(dst);
        reqs.clear();
        for (auto& b : pin_pool) (void)hipHostFree(b.second);
        for (int b = 0; b < 2; ++b) {
            if (stage[b]) (void)hipHostFree(stage[b]);
            if (stage_ev[b]) (void)hipEventDestroy(stage_ev[b]);
        }
        if (blas) (void)rocblas_destroy_handle(blas);
        if (st) (void)hipStreamDestroy(st);
        if (side) (void)hipStreamDestroy(side);
        if (dst) (void)hipStreamDestroy(dst);
    }
    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace {

// memcpy over up to 8 host threads (a fresh destination's page faults are taken in parallel too)
void host_copy(char* dst, const char* src, size_t bytes) {
    const size_t min_part = (size_t)8 << 20;
    const unsigned nt = (unsigned)std::max<size_t>(1, std::min<size_t>(8, bytes / min_part));
    if (nt == 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const size_t part = (bytes + nt - 1) / nt;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t) {
        const size_t a = t * part, b = std::min(bytes, a + part);
        if (a < b) th.emplace_back([=] { std::memcpy(dst + a, src + a, b - a); });
    }
    for (auto& t : th) t.join();
}

// Device arrays to host arrays through the two pinned stages: the copy engine fills one stage
// while the host copies the other out.
struct Download {
    void* dst;
    const void* src;
    size_t bytes;
};
void deliver(cs_server* s, const std::vector<Download>& list) {
    struct Chunk {
        char* dst;
        const char* src;
        size_t len;
    };
    std::vector<Chunk> ch;
    for (const Download& d : list)
        if (d.dst && d.bytes)
            for (size_t o = 0; o < d.bytes; o += kStageBytes)
                ch.push_back({(char*)d.dst + o, (const char*)d.src + o, std::min(kStageBytes, d.bytes - o)});
    for (size_t i = 0; i <= ch.size(); ++i) {
        if (i < ch.size()) {
            const int b = (int)(i & 1);
            CS_HIP(hipMemcpyAsync(s->stage[b], ch[i].src, ch[i].len, hipMemcpyDeviceToHost, s->dst));
            CS_HIP(hipEventRecord(s->stage_ev[b], s->dst));
        }
        if (i > 0) {
            const int b = (int)((i - 1) & 1);
            CS_HIP(hipEventSynchronize(s->stage_ev[b]));
            host_copy(ch[i - 1].dst, (const char*)s->stage[b], ch[i - 1].len);
        }
    }
}

void dots(Solve& v, const double* A, const double* B, int64_t n, DevBuf& out, std::vector<double>* host,
          hipStream_t st) {
    const int64_t k = v.cols;
    k_srv_dots<<<dim3(kSrvDotBlocks, (unsigned)k), 256, 0, st>>>(A, B, n, v.partial.as<double>());
    k_srv_reduce<<<grid_for(k), 256, 0, st>>>(v.partial.as<double>(), k, out.as<double>());
    CS_HIP(hipGetLastError());
    if (host) {
        CS_HIP(hipMemcpyAsync(host->data(), out.p, (size_t)k * 8, hipMemcpyDeviceToHost, st));
        CS_HIP(hipStreamSynchronize(st));
    }
}

// A pinned block of at least `bytes` from the server's pool, and its return.
void pin_take(cs_server* s, Solve& v, size_t bytes) {
    size_t best = s->pin_pool.size();
    for (size_t i = 0; i < s->pin_pool.size(); ++i)
        if (s->pin_pool[i].first >= bytes && (best == s->pin_pool.size() || s->pin_pool[i].first < s->pin_pool[best].first))
            best = i;
    if (best < s->pin_pool.size()) {
        v.pin_cap = s->pin_pool[best].first;
        v.pin.p = s->pin_pool[best].second;
        s->pin_pool.erase(s->pin_pool.begin() + (std::ptrdiff_t)best);
        return;
    }
    const size_t cap = std::max<size_t>(4096, (bytes + 4095) / 4096 * 4096);
    CS_HIP(hipHostMalloc(&v.pin.p, cap, hipHostMallocDefault));
    v.pin_cap = cap;
}
void pin_give(cs_server* s, Solve& v) {
    if (!v.pin.p) return;
    s->pin_pool.push_back({v.pin_cap, v.pin.p});
    v.pin.p = nullptr;
    v.pin_cap = 0;
}

int64_t result_bytes_of(const cs_server* s, const Request& q) {
    const int64_t ns = q.snps.empty() ? s->bed->M : (int64_t)q.snps.size();
    return ns * (4 * q.T + 3) * 8;
}

// Solve buffers of k columns over n rows: X, R, Z, P, K P, A P, the right-hand sides, dot partials.
int64_t solve_bytes(int64_t n, int64_t k) { return 7 * n * k * 8 + k * (kSrvDotBlocks + 8) * 8; }

// Rows the request's products read (padded), and the workspace of an apply of k columns.
int64_t work_bytes(const cs_server* s, int64_t n, bool compact, int64_t k) {
    return kinship_work_bytes(compact ? round_up(n, kPad) : s->bed->Npad, s->bed->Mpad, CS_PREC_INT8_L3, k);
}

// The n x k right-hand sides in q.B become the request's PCG columns: init, squared norms, first r.z.
void place_columns(cs_server* s, Request& q, int64_t k) {
    const int64_t n = q.n;
    Solve& v = q.sv;
    v.cols = k;
    s->used += k;
    for (DevBuf* b : {&v.X, &v.R, &v.Z, &v.P, &v.KP, &v.AP}) b->alloc((size_t)(n * k) * 8);
    v.partial.alloc((size_t)(k * kSrvDotBlocks) * 8);
    for (DevBuf* b : {&v.rz, &v.rzn, &v.pap, &v.rr, &v.bb}) b->alloc((size_t)k * 8);
    v.act.alloc((size_t)k * 4);
    pin_take(s, v, (size_t)k * 12);
    v.hbb.assign((size_t)k, 0.0);
    v.hrr.assign((size_t)k, 0.0);
    v.hact.assign((size_t)k, 0);
    k_srv_init<<<grid_for(n * k), 256, 0, s->side>>>(q.B.as<double>(), q.minv.as<double>(), n, k, v.X.as<double>(),
                                                  v.R.as<double>(), v.Z.as<double>(), v.P.as<double>());
    CS_HIP(hipGetLastError());
    dots(v, v.R.as<double>(), v.Z.as<double>(), n, v.rz, nullptr, s->side);
    dots(v, q.B.as<double>(), q.B.as<double>(), n, v.bb, &v.hbb, s->side);
    for (int64_t c = 0; c < k; ++c) {
        v.hrr[(size_t)c] = v.hbb[(size_t)c];
        v.hact[(size_t)c] = v.hbb[(size_t)c] > 0.0 ? 1 : 0;
    }
    v.act_dirty = true;
    q.chunk_iterations = 0;
}

// Give back the columns, their buffers and the operator's workspace.  The request has no device
// work in flight: its last iteration was waited for in the step that saw it converge, and a
// failure waits for both streams first.
void release_solve(cs_server* s, Request& q) {
    Solve& v = q.sv;
    pin_give(s, v);
    s->used -= v.cols;
    s->budget_used -= v.bytes;
    v = Solve();
    q.B.release();
    if (q.K) kinship_release_workspace(q.K);
}

void free_results(cs_server* s, Request& q) {
    for (DevBuf* b : {&q.freq, &q.n_obs, &q.pass, &q.beta, &q.se_out, &q.p, &q.log10p}) b->release();
    s->budget_used -= q.result_bytes;
    q.result_bytes = 0;
}

// What a request holds from admission to completion: operator, compact copy, preconditioner,
// projection.
void release_hold(cs_server* s, Request& q) {
    if (q.K) {
        (void)cs_kinship_destroy(q.K);
        q.K = nullptr;
    }
    q.minv.release();
    q.pj.R.release();
    q.pj.ZX.release();
    s->budget_used -= q.hold_bytes;
    q.hold_bytes = 0;
}

// A request that failed gives back everything it holds (after its work in flight, on either
// stream, has ended).
void fail(cs_server* s, Request& q, const std::string& msg) {
    (void)hipStreamSynchronize(s->st);
    (void)hipStreamSynchronize(s->side);
    release_solve(s, q);
    free_results(s, q);
    release_hold(s, q);
    q.pj = PscanProj();
    q.status = REQ_FAILED;
    q.error = msg;
    q.t_done = s->now();
}

bool admit(cs_server* s, Request& q) {
    const int64_t n = q.n, res = result_bytes_of(s, q);
    // operator lists (variant multiplicities, keep and its inverse), preconditioner, projection
    const int64_t base = s->bed->M * 4 + (s->N + 1) * 4 + 2 * n * 4 + n * 8 + (q.T + 2 * q.c) * n * 8;
    auto need_of = [&](bool compact, int64_t k) {
        return base + (compact ? bed_rows_copy_bytes(s->bed, n) : 0) + solve_bytes(n, k) +
               work_bytes(s, n, compact, k) + res;
    };
    const bool want_compact = n < s->N && (double)n <= kCompactFrac * (double)s->N;
    if (q.k == 0) {
        // the right-hand sides are prepared only once the request could fit (at most T + c +
        // n_cal columns)
        const int64_t k_est = std::min<int64_t>(q.T + q.c + (int64_t)q.cal.size(), s->cap);
        if (s->used + k_est > s->cap) return false;
        const bool cpt = want_compact && need_of(true, k_est) <= s->budget;
        if (need_of(cpt, k_est) <= s->budget && s->budget_used + need_of(cpt, k_est) > s->budget) return false;
        DevBuf B;
        int64_t L = 0;
        pscan_prepare_rhs(s->bed, q.keep.data(), q.n, q.n, q.Y.data(), q.n, q.T, q.X.data(), q.n, q.c,
                          q.cal.empty() ? nullptr : q.cal.data(), (int64_t)q.cal.size(), q.min_maf, B, L, s->side);
        q.L = L;
        q.k = q.T + q.c + L;
        q.B = std::move(B);
    }
    if (q.k > s->cap) throw InvalidArgument("server: request needs more columns than the server holds");
    if (s->used + q.k > s->cap) return false;
    const bool compact = want_compact && need_of(true, q.k) <= s->budget;   // else the copy never fits
    const int64_t need = need_of(compact, q.k);
    if (need > s->budget) throw InvalidArgument("server: request needs more HBM than the budget holds");
    if (s->budget_used + need > s->budget) return false;
    const int rc = cs_kinship_create(s->bed, q.keep.data(), n, s->variants.empty() ? nullptr : s->variants.data(),
                                     (int64_t)s->variants.size(), &q.K);
    if (rc != CS_OK) throw RuntimeError(cs_last_error());
    q.compact = compact;
    q.hold_bytes = base + (compact ? bed_rows_copy_bytes(s->bed, n) : 0);
    q.sv.bytes = solve_bytes(n, q.k) + work_bytes(s, n, compact, q.k);
    q.result_bytes = res;
    q.admit_bytes = q.hold_bytes + q.sv.bytes;
    s->budget_used += need;
    if (compact) kinship_compact_rows(q.K);
    std::vector<double> mv((size_t)n);
    for (int64_t r = 0; r < n; ++r) mv[(size_t)r] = 1.0 / (q.sg * s->diag[(size_t)q.keep[(size_t)r]] + q.se);
    q.minv.alloc((size_t)n * 8);
    CS_HIP(hipMemcpyAsync(q.minv.p, mv.data(), (size_t)n * 8, hipMemcpyHostToDevice, s->side));
    q.status = REQ_ACTIVE;
    q.t_admit = s->now();
    place_columns(s, q, q.k);
    return true;
}

// Place the next exact-stage chunk of q's selected markers if capacity and budget allow: the
// largest chunk that fits both.
bool admit_chunk(cs_server* s, Request& q) {
    const int64_t left = (int64_t)q.rows.size() - q.row_next;
    const int64_t maxc = std::min<int64_t>(left, s->cap - s->used);
    const int64_t avail = s->budget - s->budget_used;
    auto cost = [&](int64_t sb) { return solve_bytes(q.n, sb) + work_bytes(s, q.n, q.compact, sb); };
    if (maxc < 1 || cost(1) > avail) return false;
    int64_t lo = 1, hi = maxc;   // cost(lo) fits
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo + 1) / 2;
        if (cost(mid) <= avail) lo = mid;
        else hi = mid - 1;
    }
    const int64_t sb = lo;
    std::vector<int32_t> sel((size_t)sb);
    for (int64_t i = 0; i < sb; ++i) {
        const int64_t r = q.rows[(size_t)(q.row_next + i)];
        sel[(size_t)i] = q.snps.empty() ? (int32_t)r : q.snps[(size_t)r];
    }
    q.sv.bytes = cost(sb);
    s->budget_used += q.sv.bytes;
    q.B.alloc((size_t)(q.n * sb) * 8);
    // on the request's compact row copy when it has one (the cohort's rows in order): the same
    // markers from a cohort-sized read
    cs_bed* const rb = q.compact && q.K ? kinship_rows_bed(q.K) : nullptr;
    markers_device(rb ? rb : s->bed, rb ? nullptr : q.keep.data(), q.n, sel.data(), sb, q.min_maf, q.B.as<double>(),
                   q.n, nullptr, nullptr, s->side);
    q.chunk = sb;
    place_columns(s, q, sb);
    return true;
}

// One PCG iteration over the request's columns, enqueued on the server stream: the K.V apply,
// the vector updates and the squared residual norms (copied to pinned memory).  The step waits
// for the stream once, after every active request's iteration is queued.
void iterate_launch(cs_server* s, Request& q) {
    Solve& v = q.sv;
    const int64_t n = q.n, k = v.cols;
    if (v.act_dirty) {
        std::copy(v.hact.begin(), v.hact.end(), v.pin_act());
        CS_HIP(hipMemcpyAsync(v.act.p, v.pin_act(), (size_t)k * 4, hipMemcpyHostToDevice, s->st));
        v.act_dirty = false;
    }
    kinship_apply_async(q.K, v.P.as<double>(), n, k, v.KP.as<double>(), n, s->st);
    k_srv_apply<<<grid_for(n * k), 256, 0, s->st>>>(v.KP.as<double>(), v.P.as<double>(), q.sg, q.se, n * k,
                                                   v.AP.as<double>());
    dots(v, v.P.as<double>(), v.AP.as<double>(), n, v.pap, nullptr, s->st);
    k_srv_step<<<grid_for(n * k), 256, 0, s->st>>>(v.X.as<double>(), v.R.as<double>(), v.Z.as<double>(),
                                                  v.P.as<double>(), v.AP.as<double>(), q.minv.as<double>(),
                                                  v.rz.as<double>(), v.pap.as<double>(), v.act.as<int32_t>(), n, k);
    dots(v, v.R.as<double>(), v.Z.as<double>(), n, v.rzn, nullptr, s->st);
    k_srv_dir<<<grid_for(n * k), 256, 0, s->st>>>(v.P.as<double>(), v.Z.as<double>(), v.rzn.as<double>(),
                                                 v.rz.as<double>(), v.act.as<int32_t>(), n, k);
    k_srv_rz_keep<<<grid_for(k), 256, 0, s->st>>>(v.rz.as<double>(), v.rzn.as<double>(), v.act.as<int32_t>(), k);
    CS_HIP(hipGetLastError());
    dots(v, v.R.as<double>(), v.R.as<double>(), n, v.rr, nullptr, s->st);
    CS_HIP(hipMemcpyAsync(v.pin_rr(), v.rr.p, (size_t)k * 8, hipMemcpyDeviceToHost, s->st));
}

// After the stream wait: the iteration count and each column's convergence.
void iterate_finish(cs_server* s, Request& q) {
    Solve& v = q.sv;
    const int64_t k = v.cols;
    std::copy(v.pin_rr(), v.pin_rr() + k, v.hrr.begin());
    int& its = q.phase == 1 ? q.iterations : q.chunk_iterations;
    ++its;
    // convergence per column (the exact stage solves to its own tolerance)
    const double tol = q.phase == 1 ? q.tol : pscan_exact_tol(q.tol);
    for (int64_t c = 0; c < k; ++c) {
        if (!v.hact[(size_t)c]) continue;
        const double b2 = v.hbb[(size_t)c];
        if (b2 == 0.0 || std::sqrt(v.hrr[(size_t)c] / b2) <= tol || its >= q.max_iter) {
            v.hact[(size_t)c] = 0;
            v.act_dirty = true;
        }
    }
}

bool live(const Request& q) {
    for (int32_t a : q.sv.hact)
        if (a) return true;
    return false;
}

// The request's columns have all converged: end of phase 1 (projection, calibration, fused scan,
// selection of the exact rows) or of an exact-stage chunk.
void advance(cs_server* s, Request& q) {
    const int64_t n = q.n;
    if (q.phase == 1) {
        const int64_t ns = q.snps.empty() ? s->bed->M : (int64_t)q.snps.size();
        q.ns = ns;
        q.freq.alloc((size_t)ns * 8);
        q.n_obs.alloc((size_t)ns * 8);
        q.pass.alloc((size_t)ns * 4);
        q.beta.alloc((size_t)(ns * q.T) * 8);
        q.se_out.alloc((size_t)(ns * q.T) * 8);
        q.p.alloc((size_t)(ns * q.T) * 8);
        q.log10p.alloc((size_t)(ns * q.T) * 8);
        pscan_project(n, q.T, q.c, q.L, q.B.as<double>(), q.sv.X.as<double>(), q.pj, s->side, s->blas);
        release_solve(s, q);
        // the scan reads the request's compact row copy when it has one (the cohort's rows in the
        // order of R): class counts and g^T R products from a cohort-sized matrix, the same
        // integers as through the keep list over all samples
        cs_bed* const rb = q.compact && q.K ? kinship_rows_bed(q.K) : nullptr;
        pscan_scan(rb ? rb : s->bed, rb ? nullptr : q.keep.data(), q.pj, q.snps.empty() ? nullptr : q.snps.data(), ns, q.min_maf, q.prec,
                   q.freq.as<double>(), q.n_obs.as<int64_t>(), q.pass.as<int32_t>(), q.beta.as<double>(),
                   q.se_out.as<double>(), q.p.as<double>(), q.log10p.as<double>(), true);
        q.rows = pscan_select(q.log10p.as<double>(), q.pass.as<int32_t>(), ns, q.T, q.thr, true, s->side);
        q.row_next = 0;
        q.phase = 2;
    } else {
        const int64_t sb = q.chunk;
        // the recurrence residual stands for B - V Z (its drift from the true residual is at the
        // operator's quantisation level)
        pscan_exact(q.pj, sb, q.rows.data() + q.row_next, q.B.as<double>(), q.sv.X.as<double>(),
                    q.sv.R.as<double>(), true, q.beta.as<double>(), q.se_out.as<double>(), q.p.as<double>(),
                    q.log10p.as<double>(), true, s->side, s->blas);
        q.exact_iterations += q.chunk_iterations;
        q.row_next += sb;
        release_solve(s, q);
    }
    if (q.row_next >= (int64_t)q.rows.size()) {
        CS_HIP(hipStreamSynchronize(s->side));
        release_hold(s, q);
        q.status = REQ_DONE;
        q.t_done = s->now();
    }
}

}  // namespace

extern "C" {

int cs_server_create(cs_bed* bed, const int32_t* variants, int64_t n_variants, int64_t max_columns,
                     int64_t hbm_budget_bytes, cs_server** out) {
    return guarded([&] {
        if (!bed || !out) throw InvalidArgument("cs_server_create: null argument");
        if (max_columns < 1) throw InvalidArgument("cs_server_create: max_columns must be positive");
        auto s = std::make_unique<cs_server>();
        s->t0 = std::chrono::steady_clock::now();
        s->bed = bed;
        s->N = bed->N;
        s->cap = max_columns;
        s->budget = hbm_budget_bytes > 0 ? hbm_budget_bytes : (int64_t)1 << 40;
        if (variants && n_variants > 0) s->variants.assign(variants, variants + n_variants);
        CS_HIP(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking));
        CS_HIP(hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking));
        CS_HIP(hipStreamCreateWithFlags(&s->dst, hipStreamNonBlocking));
        for (int b = 0; b < 2; ++b) {
            CS_HIP(hipHostMalloc(&s->stage[b], kStageBytes, hipHostMallocDefault));
            CS_HIP(hipEventCreateWithFlags(&s->stage_ev[b], hipEventDisableTiming));
        }
        CS_BLAS4(rocblas_create_handle(&s->blas));
        CS_BLAS4(rocblas_set_stream(s->blas, s->side));
        // diag(K) over all samples (the requests' Jacobi preconditioners gather from it)
        cs_kinship* K = nullptr;
        if (cs_kinship_create(bed, nullptr, 0, variants, n_variants, &K) != CS_OK) throw RuntimeError(cs_last_error());
        s->diag.resize((size_t)s->N);
        const int rc = cs_kinship_diag(K, s->diag.data());
        const std::string err = rc == CS_OK ? "" : cs_last_error();
        (void)cs_kinship_destroy(K);
        if (rc != CS_OK) throw RuntimeError(err);
        *out = s.release();
    });
}

int cs_server_destroy(cs_server* s) {
    return guarded([&] {
        if (s) {
            std::lock_guard<std::mutex> lk(s->mu);
            for (auto& kv : s->reqs)
                if (kv.second->fetching) throw InvalidArgument("cs_server_destroy: results are being fetched");
        }
        delete s;
        dev_cache_release();   // the requests' compact copies and result arrays
    });
}

int cs_server_submit(cs_server* s, const cs_request* r, int64_t* id) {
    return guarded([&] {
        if (!s || !r || !r->Y || (r->n_cov > 0 && !r->X)) throw InvalidArgument("cs_server_submit: null argument");
        std::lock_guard<std::mutex> lk(s->mu);
        if (r->n_traits < 1 || r->n_cov < 1 || r->n_cov > 32)
            throw InvalidArgument("cs_server_submit: need n_traits >= 1 and 1 <= n_cov <= 32");
        if (!(r->sigma2_e > 0.0) || !(r->sigma2_g >= 0.0))
            throw InvalidArgument("cs_server_submit: need sigma2_g >= 0 and sigma2_e > 0");
        auto q = std::make_unique<Request>();
        q->id = s->next_id++;
        if (r->keep && r->n_keep > 0) {
            q->keep.assign(r->keep, r->keep + r->n_keep);
            std::vector<uint8_t> seen((size_t)s->N, 0);
            for (int32_t i : q->keep) {
                if (i < 0 || i >= s->N) throw RuntimeError("server: sample index out of range");
                if (seen[(size_t)i]++) throw InvalidArgument("server: a cohort lists a sample twice");
            }
        } else {
            q->keep.resize((size_t)s->N);
            for (int64_t i = 0; i < s->N; ++i) q->keep[(size_t)i] = (int32_t)i;
        }
        q->n = (int64_t)q->keep.size();
        if (r->ldy < q->n || r->ldx < q->n) throw RuntimeError("server: dimension mismatch");
        q->T = r->n_traits;
        q->thr = pscan_threshold(r->exact_log10p, q->T);
        q->c = r->n_cov;
        q->Y.resize((size_t)(q->n * q->T));
        for (int64_t t = 0; t < q->T; ++t)
            for (int64_t i = 0; i < q->n; ++i) q->Y[(size_t)(t * q->n + i)] = r->Y[t * r->ldy + i];
        q->X.resize((size_t)(q->n * q->c));
        for (int64_t a = 0; a < q->c; ++a)
            for (int64_t i = 0; i < q->n; ++i) q->X[(size_t)(a * q->n + i)] = r->X[a * r->ldx + i];
        q->sg = r->sigma2_g;
        q->se = r->sigma2_e;
        if (r->cal_snps && r->n_cal > 0) q->cal.assign(r->cal_snps, r->cal_snps + r->n_cal);
        if (r->snps && r->n_snps > 0) q->snps.assign(r->snps, r->snps + r->n_snps);
        for (int32_t j : q->cal)
            if (j < 0 || j >= s->bed->M) throw RuntimeError("server: variant index out of range");
        for (int32_t j : q->snps)
            if (j < 0 || j >= s->bed->M) throw RuntimeError("server: variant index out of range");
        q->min_maf = r->min_maf;
        q->tol = r->rel_tol > 0.0 ? r->rel_tol : 1e-6;
        q->max_iter = r->max_iter > 0 ? r->max_iter : 200;
        q->prec = r->prec ? r->prec : CS_PREC_INT8_L3;

        q->t_submit = s->now();
        if (id) *id = q->id;
        s->pending.push_back(q->id);
        s->reqs[q->id] = std::move(q);
    });
}

int cs_server_step(cs_server* s, int32_t* n_active, int32_t* n_pending, int32_t* n_completed) {
    return guarded([&] {
        if (!s) throw InvalidArgument("cs_server_step: null handle");
        std::lock_guard<std::mutex> lk(s->mu);
        // The step's device work runs on the server stream and the side stream; the operator,
        // row-copy and scan functions it calls synchronise their own streams before they return.
        // Device memory released during the step is therefore reused after one wait on the two
        // streams at its end, with no device-wide wait (which would also wait for a client's result
        // copies).
        DeferFrees defer(s->st, s->side);
        int32_t completed = 0;
        // 1. one PCG iteration of every request with live columns, queued back to back on the
        //    server stream and waited for once (step 3)
        std::vector<Request*> ran;
        for (auto& kv : s->reqs) {
            Request& q = *kv.second;
            if (q.status != REQ_ACTIVE || q.sv.cols == 0 || !live(q)) continue;
            try {
                iterate_launch(s, q);
                ran.push_back(&q);
            } catch (const std::exception& e) {
                defer.sync_device();
                fail(s, q, e.what());
            }
        }
        // 2. beside the iterations, with device work on the side stream: requests whose columns
        //    converged in an earlier step (projection, calibration, scan; exact-stage results),
        //    then exact-stage chunks of active requests (they hold results already) in id order,
        //    then admission in arrival order (a request that does not fit waits; later ones may
        //    pass it)
        for (auto& kv : s->reqs) {
            Request& q = *kv.second;
            if (q.status != REQ_ACTIVE || q.sv.cols == 0 || live(q)) continue;
            try {
                advance(s, q);
                if (q.status == REQ_DONE) ++completed;
            } catch (const std::exception& e) {
                defer.sync_device();
                fail(s, q, e.what());
            }
        }
        for (auto& kv : s->reqs) {
            Request& q = *kv.second;
            if (q.status != REQ_ACTIVE || q.phase != 2 || q.sv.cols > 0) continue;
            try {
                (void)admit_chunk(s, q);
            } catch (const std::exception& e) {
                defer.sync_device();
                fail(s, q, e.what());
            }
        }
        for (auto it = s->pending.begin(); it != s->pending.end();) {
            Request& q = *s->reqs[*it];
            bool ok = false;
            try {
                ok = admit(s, q);
            } catch (const std::exception& e) {
                defer.sync_device();
                fail(s, q, e.what());
                it = s->pending.erase(it);
                continue;
            }
            if (ok) it = s->pending.erase(it);
            else ++it;
        }
        // 3. the iterations' residual norms and each column's convergence (converged requests
        //    advance in the next step); the columns placed in step 2 iterate from the next step
        if (!ran.empty()) CS_HIP(hipStreamSynchronize(s->st));
        for (Request* q : ran) iterate_finish(s, *q);
        CS_HIP(hipStreamSynchronize(s->side));
        int32_t na = 0;
        for (auto& kv : s->reqs) na += kv.second->status == REQ_ACTIVE ? 1 : 0;
        if (n_active) *n_active = na;
        if (n_pending) *n_pending = (int32_t)s->pending.size();
        if (n_completed) *n_completed = completed;
    });
}

int cs_server_result(cs_server* s, int64_t id, int32_t* status, double* freq, int64_t* n_obs, int32_t* pass,
                     double* beta, double* se, double* p, double* log10p, cs_request_stats* stats) {
    return guarded([&] {
        if (!s) throw InvalidArgument("cs_server_result: null handle");
        std::vector<Download> list;
        Request* qp = nullptr;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            auto it = s->reqs.find(id);
            if (it == s->reqs.end()) throw InvalidArgument("cs_server_result: unknown request");
            Request& q = *it->second;
            if (status) *status = q.status;
            if (stats) {
                stats->t_submit = q.t_submit;
                stats->t_admit = q.t_admit;
                stats->t_done = q.t_done;
                stats->iterations = q.iterations;
                stats->gamma = q.pj.gamma;
                stats->n_calibration = q.pj.ncal;
                stats->n_columns = q.k;
                stats->n_exact = (int64_t)q.rows.size();
                stats->exact_iterations = q.exact_iterations;
                stats->calibration_rel_sd = q.pj.rel_sd;
                stats->dlog10p_bound_below = pscan_bound_below(q.thr, q.pj.rel_sd);
                stats->admit_bytes = q.admit_bytes;
                stats->compact_rows = q.compact ? 1 : 0;
            }
            const bool wants = freq || n_obs || pass || beta || se || p || log10p;
            if (q.status == REQ_FAILED && wants) throw RuntimeError("request failed: " + q.error);
            if (q.status != REQ_DONE || !wants) return;
            const size_t ns = (size_t)q.ns, nt = (size_t)(q.ns * q.T);
            list = {{freq, q.freq.p, ns * 8},   {n_obs, q.n_obs.p, ns * 8}, {pass, q.pass.p, ns * 4},
                    {beta, q.beta.p, nt * 8},   {se, q.se_out.p, nt * 8},  {p, q.p.p, nt * 8},
                    {log10p, q.log10p.p, nt * 8}};
            ++q.fetching;
            qp = &q;
        }
        // a finished request's arrays are read only here, and it stays in the map while fetching
        // (cs_server_release refuses it), so the copy runs outside the scheduler lock
        struct Done {
            cs_server* s;
            Request* q;
            ~Done() {
                std::lock_guard<std::mutex> lk(s->mu);
                --q->fetching;
            }
        } done{s, qp};
        std::lock_guard<std::mutex> dl(s->dmu);
        deliver(s, list);
    });
}

int cs_server_release(cs_server* s, int64_t id) {
    return guarded([&] {
        if (!s) throw InvalidArgument("cs_server_release: null handle");
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(id);
        if (it == s->reqs.end()) return;
        if (it->second->status == REQ_ACTIVE || it->second->status == REQ_PENDING)
            throw InvalidArgument("cs_server_release: request still running");
        if (it->second->fetching) throw InvalidArgument("cs_server_release: results are being fetched");
        // a finished request's arrays were written before it completed (the server stream was
        // waited for) and its fetches have ended: nothing in flight uses them
        IdleFrees idle;
        free_results(s, *it->second);
        s->reqs.erase(it);
    });
}

}  // extern "C"
