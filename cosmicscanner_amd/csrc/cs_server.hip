// Serving loop for scan requests over a resident genotype matrix.
//
// One kinship operator K over all samples (and the chosen variants) stays resident.  Because K's
// standardisation uses all-sample statistics (kinship_operator.cpp:64-67, plink_reader.cpp:187-239),
// the kinship matrix of a cohort S is the restriction P_S K P_S^T: a request's system
// (sigma_g K_S + sigma_e I) x = b is solved in the full sample space with b zero outside S and
// the operator masked to S.  Every request contributes columns (its traits, its covariates and its
// calibration markers) to one shared multi-RHS PCG: one K.V apply per iteration serves every
// admitted request, each column keeping its own variance components, preconditioner, mask,
// recurrence and convergence.  New requests are admitted at iteration boundaries as column slots
// and the HBM budget free up (continuous batching).  A request whose columns have all converged
// runs its projection, calibration and fused per-SNP scan at once and gives its slots back; the
// SNPs its exact stage selects then return to the batch as a second phase of columns (their
// prepared markers, in chunks that fit the free slots and budget), whose solutions give the exact
// GLS variances.  The budget covers the solve buffers while held and the result arrays from
// admission until the client releases them.
//
// A column's arithmetic is independent of the other columns in the batch (per-column
// quantisation scales, exact integer products, per-column dots with a fixed reduction order), so a
// request's results are the same alone and in any mix.
#include <rocblas/rocblas.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <map>
#include <memory>
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

// AP = sigma_g[c] mask(K P) + sigma_e[c] P   (mask: minv != 0 on the column's cohort)
__global__ void k_srv_apply(const double* __restrict__ KP, const double* __restrict__ P,
                            const double* __restrict__ minv, const double* __restrict__ sg,
                            const double* __restrict__ se, int64_t n, int64_t k, double* __restrict__ AP) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * k;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = e / n;
        AP[e] = (minv[e] != 0.0 ? sg[c] * KP[e] : 0.0) + se[c] * P[e];
    }
}

// column init: X = 0, R = B, Z = Minv R, P = Z on columns [c0, c0 + k)
__global__ void k_srv_init(const double* __restrict__ Bm, const double* __restrict__ minv, int64_t n, int64_t c0,
                           int64_t k, double* __restrict__ X, double* __restrict__ R, double* __restrict__ Z,
                           double* __restrict__ P) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * k;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t g = c0 * n + e;
        const double r = Bm[g];
        X[g] = 0.0;
        R[g] = r;
        const double z = minv[g] * r;
        Z[g] = z;
        P[g] = z;
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
        Z[e] = minv[e] * r;
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

// dst[c * n + keep[r]] = src[c * nk + r]  (scatter compact cohort rows into full-sample columns)
__global__ void k_srv_scatter(const double* __restrict__ src, int64_t nk, const int32_t* __restrict__ keep,
                              int64_t n, int64_t k, double* __restrict__ dst) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nk * k;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = e / nk, r = e - c * nk;
        dst[c * n + keep[r]] = src[e];
    }
}

__global__ void k_srv_gather(const double* __restrict__ src, int64_t n, const int32_t* __restrict__ keep,
                             int64_t nk, int64_t k, double* __restrict__ dst) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nk * k;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = e / nk, r = e - c * nk;
        dst[e] = src[c * n + keep[r]];
    }
}

}  // namespace csamd

using namespace csamd;

namespace {

enum { REQ_PENDING = 0, REQ_ACTIVE = 1, REQ_DONE = 2, REQ_FAILED = -1 };

struct Request {
    int64_t id = 0;
    int status = REQ_PENDING;
    std::vector<int32_t> keep;     // sorted-order not required; the cohort in the caller's order
    int64_t n = 0, T = 0, c = 0, L = 0, k = 0;
    std::vector<double> Y;
    std::vector<float> X;
    double sg = 0, se = 1;
    std::vector<int32_t> cal, snps;
    double min_maf = 0.01, tol = 1e-6, thr = kPscanExactDefault;   // effective (pscan_threshold)
    int max_iter = 200, prec = CS_PREC_INT8_L3;
    // phase 1: the main solve over [Y, X, G_cal]; phase 2: exact-stage chunks of selected markers
    int phase = 1;
    DevBuf B;                      // compact RHS of the columns in the slots (n x cols)
    DevBuf dkeep;
    int64_t slot0 = -1, cols = 0;  // slots held (slot0 < 0: none)
    int iterations = 0, chunk_iterations = 0, exact_iterations = 0;
    int64_t bytes = 0;             // solve buffers charged to the budget while held
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
};

}  // namespace

struct cs_server {
    cs_bed* bed = nullptr;
    cs_kinship* K = nullptr;
    int64_t N = 0, cap = 0, used = 0;
    int64_t budget = 0, budget_used = 0;
    hipStream_t st = nullptr;
    rocblas_handle blas = nullptr;
    DevBuf Xs, R, Z, P, KP, AP, Bs, minv, dsg, dse, act, partial, rz, rzn, pap, rr, bb;
    std::vector<double> diag, hsg, hse, hbb, hrr;
    std::vector<int32_t> hact;
    std::vector<int64_t> slot_req;
    std::deque<int64_t> pending;
    std::map<int64_t, std::unique_ptr<Request>> reqs;
    int64_t next_id = 1;
    std::chrono::steady_clock::time_point t0;
    ~cs_server() {
        if (K) cs_kinship_destroy(K);
        if (blas) (void)rocblas_destroy_handle(blas);
        if (st) (void)hipStreamDestroy(st);
    }
    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace {

void dots(cs_server* s, const double* A, const double* B, int64_t k, DevBuf& out, std::vector<double>* host) {
    if (k == 0) return;
    k_srv_dots<<<dim3(kSrvDotBlocks, (unsigned)k), 256, 0, s->st>>>(A, B, s->N, s->partial.as<double>());
    k_srv_reduce<<<grid_for(k), 256, 0, s->st>>>(s->partial.as<double>(), k, out.as<double>());
    CS_HIP(hipGetLastError());
    if (host) {
        CS_HIP(hipMemcpyAsync(host->data(), out.p, (size_t)k * 8, hipMemcpyDeviceToHost, s->st));
        CS_HIP(hipStreamSynchronize(s->st));
    }
}

int64_t find_slots(cs_server* s, int64_t k) {
    int64_t run = 0;
    for (int64_t c = 0; c < s->cap; ++c) {
        run = s->slot_req[(size_t)c] < 0 ? run + 1 : 0;
        if (run == k) return c - k + 1;
    }
    return -1;
}

int64_t longest_free_run(const cs_server* s) {
    int64_t run = 0, best = 0;
    for (int64_t c = 0; c < s->cap; ++c) {
        run = s->slot_req[(size_t)c] < 0 ? run + 1 : 0;
        best = std::max(best, run);
    }
    return best;
}

// Put the k compact cohort columns of q.B into the slots [s0, s0 + k): scatter into full-sample
// columns, preconditioner / mask, variance components, PCG init, squared norms and first r.z.
void place_columns(cs_server* s, Request& q, int64_t s0, int64_t k) {
    const int64_t N = s->N;
    q.slot0 = s0;
    q.cols = k;
    for (int64_t c = 0; c < k; ++c) s->slot_req[(size_t)(s0 + c)] = q.id;
    CS_HIP(hipMemsetAsync(s->Bs.as<double>() + s0 * N, 0, (size_t)(N * k) * 8, s->st));
    k_srv_scatter<<<grid_for(q.n * k), 256, 0, s->st>>>(q.B.as<double>(), q.n, q.dkeep.as<int32_t>(), N, k,
                                                       s->Bs.as<double>() + s0 * N);
    std::vector<double> mv((size_t)(N * k), 0.0);
    for (int64_t c = 0; c < k; ++c)
        for (int64_t r = 0; r < q.n; ++r) {
            const int64_t i = q.keep[(size_t)r];
            mv[(size_t)(c * N + i)] = 1.0 / (q.sg * s->diag[(size_t)i] + q.se);
        }
    CS_HIP(hipMemcpyAsync(s->minv.as<double>() + s0 * N, mv.data(), mv.size() * 8, hipMemcpyHostToDevice, s->st));
    for (int64_t c = 0; c < k; ++c) {
        s->hsg[(size_t)(s0 + c)] = q.sg;
        s->hse[(size_t)(s0 + c)] = q.se;
    }
    CS_HIP(hipMemcpyAsync(s->dsg.p, s->hsg.data(), (size_t)s->cap * 8, hipMemcpyHostToDevice, s->st));
    CS_HIP(hipMemcpyAsync(s->dse.p, s->hse.data(), (size_t)s->cap * 8, hipMemcpyHostToDevice, s->st));
    k_srv_init<<<grid_for(N * k), 256, 0, s->st>>>(s->Bs.as<double>(), s->minv.as<double>(), N, s0, k,
                                                  s->Xs.as<double>(), s->R.as<double>(), s->Z.as<double>(),
                                                  s->P.as<double>());
    CS_HIP(hipGetLastError());
    const double* Bc = s->Bs.as<double>() + s0 * N;
    k_srv_dots<<<dim3(kSrvDotBlocks, (unsigned)k), 256, 0, s->st>>>(Bc, Bc, N, s->partial.as<double>());
    k_srv_reduce<<<grid_for(k), 256, 0, s->st>>>(s->partial.as<double>(), k, s->bb.as<double>() + s0);
    k_srv_dots<<<dim3(kSrvDotBlocks, (unsigned)k), 256, 0, s->st>>>(s->R.as<double>() + s0 * N, s->Z.as<double>() + s0 * N,
                                                                   N, s->partial.as<double>());
    k_srv_reduce<<<grid_for(k), 256, 0, s->st>>>(s->partial.as<double>(), k, s->rz.as<double>() + s0);
    CS_HIP(hipGetLastError());
    CS_HIP(hipMemcpyAsync(s->hbb.data() + s0, s->bb.as<double>() + s0, (size_t)k * 8, hipMemcpyDeviceToHost, s->st));
    CS_HIP(hipStreamSynchronize(s->st));
    for (int64_t c = 0; c < k; ++c) {
        s->hrr[(size_t)(s0 + c)] = s->hbb[(size_t)(s0 + c)];
        s->hact[(size_t)(s0 + c)] = s->hbb[(size_t)(s0 + c)] > 0.0 ? 1 : 0;
    }
    s->used = std::max(s->used, s0 + k);
    q.chunk_iterations = 0;
}

// Free the request's slots (P and the mask zeroed so idle columns add nothing to the shared
// apply) and its solve buffers; safe to call on any path, including failures.
void release_slots(cs_server* s, Request& q) {
    if (q.slot0 >= 0) {
        const int64_t N = s->N;
        (void)hipMemsetAsync(s->P.as<double>() + q.slot0 * N, 0, (size_t)(N * q.cols) * 8, s->st);
        (void)hipMemsetAsync(s->minv.as<double>() + q.slot0 * N, 0, (size_t)(N * q.cols) * 8, s->st);
        (void)hipStreamSynchronize(s->st);
        for (int64_t c = 0; c < q.cols; ++c) {
            s->slot_req[(size_t)(q.slot0 + c)] = -1;
            s->hact[(size_t)(q.slot0 + c)] = 0;
        }
        while (s->used > 0 && s->slot_req[(size_t)(s->used - 1)] < 0) --s->used;
        q.slot0 = -1;
        q.cols = 0;
    }
    s->budget_used -= q.bytes;
    q.bytes = 0;
    q.B.release();
}

void free_results(cs_server* s, Request& q) {
    for (DevBuf* b : {&q.freq, &q.n_obs, &q.pass, &q.beta, &q.se_out, &q.p, &q.log10p}) b->release();
    s->budget_used -= q.result_bytes;
    q.result_bytes = 0;
}

// A request that failed gives back everything it holds.
void fail(cs_server* s, Request& q, const std::string& msg) {
    release_slots(s, q);
    free_results(s, q);
    q.dkeep.release();
    q.pj = PscanProj();
    q.status = REQ_FAILED;
    q.error = msg;
    q.t_done = s->now();
}

int64_t result_bytes_of(const cs_server* s, const Request& q) {
    const int64_t ns = q.snps.empty() ? s->bed->M : (int64_t)q.snps.size();
    return ns * (4 * q.T + 3) * 8;
}

bool admit(cs_server* s, Request& q) {
    if (q.k == 0) {
        DevBuf B;
        int64_t L = 0;
        pscan_prepare_rhs(s->bed, q.keep.data(), q.n, q.n, q.Y.data(), q.n, q.T, q.X.data(), q.n, q.c,
                          q.cal.empty() ? nullptr : q.cal.data(), (int64_t)q.cal.size(), q.min_maf, B, L, s->st);
        q.L = L;
        q.k = q.T + q.c + L;
        q.B = std::move(B);
    }
    if (q.k > s->cap) throw InvalidArgument("server: request needs more columns than the server holds");
    // solve buffers (B, the gathered solution, the projection) and the result arrays
    const int64_t need = (int64_t)q.B.bytes * 3, res = result_bytes_of(s, q);
    if (s->budget_used + need + res > s->budget) return false;
    const int64_t s0 = find_slots(s, q.k);
    if (s0 < 0) return false;
    q.dkeep.alloc((size_t)q.n * 4);
    CS_HIP(hipMemcpyAsync(q.dkeep.p, q.keep.data(), (size_t)q.n * 4, hipMemcpyHostToDevice, s->st));
    q.bytes = need;
    q.result_bytes = res;
    s->budget_used += need + res;
    place_columns(s, q, s0, q.k);
    q.status = REQ_ACTIVE;
    q.t_admit = s->now();
    return true;
}

// Place the next exact-stage chunk of q's selected markers if slots and budget allow: as many
// rows as fit the longest free slot run and the budget left (prepared markers, gathered solution
// and residual: three n x chunk FP64 arrays).
bool admit_chunk(cs_server* s, Request& q) {
    const int64_t left = (int64_t)q.rows.size() - q.row_next;
    const int64_t per_col = 3 * q.n * 8;
    int64_t sb = std::min<int64_t>(left, longest_free_run(s));
    sb = std::min<int64_t>(sb, (s->budget - s->budget_used) / per_col);
    if (sb < 1) return false;
    const int64_t s0 = find_slots(s, sb);
    std::vector<int32_t> sel((size_t)sb);
    for (int64_t i = 0; i < sb; ++i) {
        const int64_t r = q.rows[(size_t)(q.row_next + i)];
        sel[(size_t)i] = q.snps.empty() ? (int32_t)r : q.snps[(size_t)r];
    }
    q.B.alloc((size_t)(q.n * sb) * 8);
    q.bytes = per_col * sb;
    s->budget_used += q.bytes;
    markers_device(s->bed, q.keep.data(), q.n, sel.data(), sb, q.min_maf, q.B.as<double>(), q.n, nullptr, nullptr,
                   s->st);
    q.chunk = sb;
    place_columns(s, q, s0, sb);
    return true;
}

// The request's columns in the slots have all converged: end of phase 1 (projection,
// calibration, fused scan, selection of the exact rows) or of an exact-stage chunk.
void advance(cs_server* s, Request& q) {
    const int64_t N = s->N, n = q.n;
    if (q.phase == 1) {
        const int64_t k = q.k;
        DevBuf Zc((size_t)(n * k) * 8);
        k_srv_gather<<<grid_for(n * k), 256, 0, s->st>>>(s->Xs.as<double>() + q.slot0 * N, N, q.dkeep.as<int32_t>(),
                                                        n, k, Zc.as<double>());
        CS_HIP(hipGetLastError());
        CS_HIP(hipStreamSynchronize(s->st));
        const int64_t ns = q.snps.empty() ? s->bed->M : (int64_t)q.snps.size();
        q.ns = ns;
        q.freq.alloc((size_t)ns * 8);
        q.n_obs.alloc((size_t)ns * 8);
        q.pass.alloc((size_t)ns * 4);
        q.beta.alloc((size_t)(ns * q.T) * 8);
        q.se_out.alloc((size_t)(ns * q.T) * 8);
        q.p.alloc((size_t)(ns * q.T) * 8);
        q.log10p.alloc((size_t)(ns * q.T) * 8);
        pscan_project(n, q.T, q.c, q.L, q.B.as<double>(), Zc.as<double>(), q.pj, s->st, s->blas);
        Zc.release();
        release_slots(s, q);
        pscan_scan(s->bed, q.keep.data(), q.pj, q.snps.empty() ? nullptr : q.snps.data(), ns, q.min_maf, q.prec,
                   q.freq.as<double>(), q.n_obs.as<int64_t>(), q.pass.as<int32_t>(), q.beta.as<double>(),
                   q.se_out.as<double>(), q.p.as<double>(), q.log10p.as<double>(), true);
        q.rows = pscan_select(q.log10p.as<double>(), q.pass.as<int32_t>(), ns, q.T, q.thr, true, s->st);
        q.row_next = 0;
        q.phase = 2;
    } else {
        const int64_t sb = q.chunk;
        DevBuf Zc((size_t)(n * sb) * 8), Rc((size_t)(n * sb) * 8);
        k_srv_gather<<<grid_for(n * sb), 256, 0, s->st>>>(s->Xs.as<double>() + q.slot0 * N, N, q.dkeep.as<int32_t>(),
                                                         n, sb, Zc.as<double>());
        k_srv_gather<<<grid_for(n * sb), 256, 0, s->st>>>(s->R.as<double>() + q.slot0 * N, N, q.dkeep.as<int32_t>(),
                                                         n, sb, Rc.as<double>());
        CS_HIP(hipGetLastError());
        // the recurrence residual stands for B - V Z (its drift from the true residual is at the
        // operator's quantisation level)
        pscan_exact(q.pj, sb, q.rows.data() + q.row_next, q.B.as<double>(), Zc.as<double>(), Rc.as<double>(), true,
                    q.beta.as<double>(), q.se_out.as<double>(), q.p.as<double>(), q.log10p.as<double>(), true, s->st,
                    s->blas);
        q.exact_iterations += q.chunk_iterations;
        q.row_next += sb;
        release_slots(s, q);
    }
    if (q.row_next >= (int64_t)q.rows.size()) {
        q.dkeep.release();
        q.pj.R.release();
        q.pj.ZX.release();
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
        const int rc = cs_kinship_create(bed, nullptr, 0, variants, n_variants, &s->K);
        if (rc != CS_OK) throw RuntimeError(cs_last_error());
        CS_HIP(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking));
        CS_BLAS4(rocblas_create_handle(&s->blas));
        CS_BLAS4(rocblas_set_stream(s->blas, s->st));
        const size_t cols = (size_t)(s->N * s->cap) * 8;
        for (DevBuf* b : {&s->Xs, &s->R, &s->Z, &s->P, &s->KP, &s->AP, &s->Bs, &s->minv}) {
            b->alloc(cols);
            CS_HIP(hipMemsetAsync(b->p, 0, cols, s->st));
        }
        for (DevBuf* b : {&s->dsg, &s->dse, &s->rz, &s->rzn, &s->pap, &s->rr, &s->bb}) b->alloc((size_t)s->cap * 8);
        s->act.alloc((size_t)s->cap * 4);
        s->partial.alloc((size_t)(s->cap * kSrvDotBlocks) * 8);
        s->diag.resize((size_t)s->N);
        if (cs_kinship_diag(s->K, s->diag.data()) != CS_OK) throw RuntimeError(cs_last_error());
        s->hsg.assign((size_t)s->cap, 0.0);
        s->hse.assign((size_t)s->cap, 1.0);
        s->hbb.assign((size_t)s->cap, 0.0);
        s->hrr.assign((size_t)s->cap, 0.0);
        s->hact.assign((size_t)s->cap, 0);
        s->slot_req.assign((size_t)s->cap, -1);
        CS_HIP(hipStreamSynchronize(s->st));
        *out = s.release();
    });
}

int cs_server_destroy(cs_server* s) {
    return guarded([&] { delete s; });
}

int cs_server_submit(cs_server* s, const cs_request* r, int64_t* id) {
    return guarded([&] {
        if (!s || !r || !r->Y || (r->n_cov > 0 && !r->X)) throw InvalidArgument("cs_server_submit: null argument");
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
        int32_t completed = 0;
        // exact-stage chunks of active requests first (they hold results already), in id order
        for (auto& kv : s->reqs) {
            Request& q = *kv.second;
            if (q.status != REQ_ACTIVE || q.phase != 2 || q.slot0 >= 0) continue;
            try {
                (void)admit_chunk(s, q);
            } catch (const std::exception& e) {
                fail(s, q, e.what());
            }
        }
        // admission in arrival order (a request that does not fit waits; later ones may pass it)
        for (auto it = s->pending.begin(); it != s->pending.end();) {
            Request& q = *s->reqs[*it];
            bool ok = false;
            try {
                ok = admit(s, q);
            } catch (const std::exception& e) {
                fail(s, q, e.what());
                it = s->pending.erase(it);
                continue;
            }
            if (ok) it = s->pending.erase(it);
            else ++it;
        }
        const int64_t N = s->N, k = s->used;
        bool any = false;
        for (int64_t c = 0; c < k; ++c) any = any || s->hact[(size_t)c];
        if (any) {
            CS_HIP(hipMemcpyAsync(s->act.p, s->hact.data(), (size_t)k * 4, hipMemcpyHostToDevice, s->st));
            CS_HIP(hipStreamSynchronize(s->st));
            const int rc = cs_kinship_apply_device(s->K, s->P.as<double>(), N, k, s->KP.as<double>(), N);
            if (rc != CS_OK) throw RuntimeError(cs_last_error());
            k_srv_apply<<<grid_for(N * k), 256, 0, s->st>>>(s->KP.as<double>(), s->P.as<double>(), s->minv.as<double>(),
                                                           s->dsg.as<double>(), s->dse.as<double>(), N, k,
                                                           s->AP.as<double>());
            dots(s, s->P.as<double>(), s->AP.as<double>(), k, s->pap, nullptr);
            k_srv_step<<<grid_for(N * k), 256, 0, s->st>>>(s->Xs.as<double>(), s->R.as<double>(), s->Z.as<double>(),
                                                          s->P.as<double>(), s->AP.as<double>(), s->minv.as<double>(),
                                                          s->rz.as<double>(), s->pap.as<double>(), s->act.as<int32_t>(),
                                                          N, k);
            dots(s, s->R.as<double>(), s->Z.as<double>(), k, s->rzn, nullptr);
            k_srv_dir<<<grid_for(N * k), 256, 0, s->st>>>(s->P.as<double>(), s->Z.as<double>(), s->rzn.as<double>(),
                                                         s->rz.as<double>(), s->act.as<int32_t>(), N, k);
            CS_HIP(hipGetLastError());
            // rz <- rz_new on the active columns only
            std::vector<double> rzn((size_t)k), rzo((size_t)k);
            CS_HIP(hipMemcpyAsync(rzn.data(), s->rzn.p, (size_t)k * 8, hipMemcpyDeviceToHost, s->st));
            CS_HIP(hipMemcpyAsync(rzo.data(), s->rz.p, (size_t)k * 8, hipMemcpyDeviceToHost, s->st));
            dots(s, s->R.as<double>(), s->R.as<double>(), k, s->rr, &s->hrr);
            for (int64_t c = 0; c < k; ++c)
                if (s->hact[(size_t)c]) rzo[(size_t)c] = rzn[(size_t)c];
            CS_HIP(hipMemcpyAsync(s->rz.p, rzo.data(), (size_t)k * 8, hipMemcpyHostToDevice, s->st));
            CS_HIP(hipStreamSynchronize(s->st));
            // per-request iteration counts and convergence (the exact stage solves to its own
            // tolerance)
            std::map<int64_t, bool> stepped;
            for (int64_t c = 0; c < k; ++c)
                if (s->hact[(size_t)c] && s->slot_req[(size_t)c] >= 0) stepped[s->slot_req[(size_t)c]] = true;
            for (auto& kv : stepped) {
                Request& q = *s->reqs[kv.first];
                if (q.phase == 1) q.iterations++;
                else q.chunk_iterations++;
            }
            for (int64_t c = 0; c < k; ++c) {
                if (!s->hact[(size_t)c] || s->slot_req[(size_t)c] < 0) continue;
                const Request& q = *s->reqs[s->slot_req[(size_t)c]];
                const double b2 = s->hbb[(size_t)c];
                const double tol = q.phase == 1 ? q.tol : pscan_exact_tol(q.tol);
                const int its = q.phase == 1 ? q.iterations : q.chunk_iterations;
                if (b2 == 0.0 || std::sqrt(s->hrr[(size_t)c] / b2) <= tol || its >= q.max_iter)
                    s->hact[(size_t)c] = 0;
            }
        }
        // advance requests whose columns have all converged
        for (auto& kv : s->reqs) {
            Request& q = *kv.second;
            if (q.status != REQ_ACTIVE || q.slot0 < 0) continue;
            bool live = false;
            for (int64_t c = 0; c < q.cols; ++c) live = live || s->hact[(size_t)(q.slot0 + c)];
            if (live) continue;
            try {
                advance(s, q);
                if (q.status == REQ_DONE) ++completed;
            } catch (const std::exception& e) {
                fail(s, q, e.what());
            }
        }
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
        auto it = s->reqs.find(id);
        if (it == s->reqs.end()) throw InvalidArgument("cs_server_result: unknown request");
        const Request& q = *it->second;
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
        }
        const bool wants = freq || n_obs || pass || beta || se || p || log10p;
        if (q.status == REQ_FAILED && wants) throw RuntimeError("request failed: " + q.error);
        if (q.status != REQ_DONE) return;
        auto cp = [&](void* dst, const DevBuf& v, size_t bytes) {
            if (dst && bytes) CS_HIP(hipMemcpyAsync(dst, v.p, bytes, hipMemcpyDeviceToHost, s->st));
        };
        const size_t ns = (size_t)q.ns, nt = (size_t)(q.ns * q.T);
        cp(freq, q.freq, ns * 8);
        cp(n_obs, q.n_obs, ns * 8);
        cp(pass, q.pass, ns * 4);
        cp(beta, q.beta, nt * 8);
        cp(se, q.se_out, nt * 8);
        cp(p, q.p, nt * 8);
        cp(log10p, q.log10p, nt * 8);
        CS_HIP(hipStreamSynchronize(s->st));
    });
}

int cs_server_release(cs_server* s, int64_t id) {
    return guarded([&] {
        if (!s) throw InvalidArgument("cs_server_release: null handle");
        auto it = s->reqs.find(id);
        if (it == s->reqs.end()) return;
        if (it->second->status == REQ_ACTIVE || it->second->status == REQ_PENDING)
            throw InvalidArgument("cs_server_release: request still running");
        free_results(s, *it->second);
        s->reqs.erase(it);
    });
}

}  // extern "C"
