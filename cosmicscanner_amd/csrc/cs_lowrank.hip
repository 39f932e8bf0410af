// Randomized low-rank null model on the matrix-free kinship operator and the low-rank scan.
//
// LmmEngine::prepareLowRankMatrixFree (src/lmm_engine.cpp:897-1095), and MatrixFreeVce::
// prepareMatrixFree + runBrentNullModel (src/matrix_free_vce.cpp:32-235) which is the same
// computation with the trace threshold switched off (its fallback role in
// GwasScanner::runLocoScan, src/gwas_scanner.cpp:322-372):
//   Omega ~ N(0,1) from std::mt19937_64(seed) and std::normal_distribution<double> (the same
//   standard-library generator as the reference, column by column), Y = K Omega, two power
//   iterations Q = qr(Y), Y = K Q, final Q, B = Q^T K Q, eigendecomposition, the leading
//   `rank` eigenpairs U = Q v (descending), trace-fraction rank selection, residual eigenvalue
//   r0 = (tr K - sum lambda) / (n - rank), REML over log eta on an 81-point grid on [-12, 12] then
//   golden section (80 iterations, 1e-8), sigma_e = qf / df, sigma_g = eta sigma_e.
//   K.V on the GPU (int8 MFMA operator), CholeskyQR2 (Householder when ill-conditioned), eigen
//   and GEMMs with rocSOLVER / rocBLAS.  The REML
//   objective is evaluated from the small projections U^T X, U^T y, X^T X, X^T y, y^T y (exact
//   algebra of the reference's R^-1 = base I + U diag(1/(1+eta lambda) - base) U^T).
// LmmEngine::scanLowRank (src/lmm_engine.cpp:1269-1361): V^-1 = base I + U diag(corr) U^T with
//   base = 1 / max(1e-12, sigma_g r0 + sigma_e), corr = 1 / (sigma_g lambda + sigma_e) - base;
//   score = g^T V^-1 (y - X beta), var = g^T V^-1 g - (X^T V^-1 g)^T (X^T V^-1 X)^-1 (X^T V^-1 g),
//   through the fused scan with R = [P y, U, V^-1 X].
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <random>
#include <vector>

#include "cs_blas.h"
#include "cs_common.h"
#include "cs_golden.h"
#include "cs_ops.h"

namespace csamd {

#define CS_BLAS3(call)                                                                 \
    do {                                                                               \
        rocblas_status s_ = (call);                                                    \
        if (s_ != rocblas_status_success)                                              \
            throw ::csamd::RuntimeError(std::string("rocBLAS/rocSOLVER error ") +      \
                                        rocblas_status_to_string(s_) + " at " #call);  \
    } while (0)

// Cholesky of a small SPD matrix: inverse (column-major) and log det; false if not PD.
static bool chol_inv_logdet(const std::vector<double>& A, int64_t c, std::vector<double>& inv, double& logdet) {
    std::vector<double> L((size_t)(c * c), 0.0);
    logdet = 0.0;
    for (int64_t j = 0; j < c; ++j) {
        double d = A[(size_t)(j * c + j)];
        for (int64_t k = 0; k < j; ++k) d -= L[(size_t)(k * c + j)] * L[(size_t)(k * c + j)];
        if (!(d > 0.0) || !std::isfinite(d)) return false;
        const double ljj = std::sqrt(d);
        L[(size_t)(j * c + j)] = ljj;
        logdet += 2.0 * std::log(ljj);
        for (int64_t i = j + 1; i < c; ++i) {
            double v = A[(size_t)(j * c + i)];
            for (int64_t k = 0; k < j; ++k) v -= L[(size_t)(k * c + i)] * L[(size_t)(k * c + j)];
            L[(size_t)(j * c + i)] = v / ljj;
        }
    }
    std::vector<double> Li((size_t)(c * c), 0.0);
    for (int64_t j = 0; j < c; ++j) {
        Li[(size_t)(j * c + j)] = 1.0 / L[(size_t)(j * c + j)];
        for (int64_t i = j + 1; i < c; ++i) {
            double v = 0.0;
            for (int64_t k = j; k < i; ++k) v -= L[(size_t)(k * c + i)] * Li[(size_t)(j * c + k)];
            Li[(size_t)(j * c + i)] = v / L[(size_t)(i * c + i)];
        }
    }
    inv.assign((size_t)(c * c), 0.0);
    for (int64_t i = 0; i < c; ++i)
        for (int64_t j = 0; j < c; ++j) {
            double v = 0.0;
            for (int64_t k = std::max(i, j); k < c; ++k) v += Li[(size_t)(i * c + k)] * Li[(size_t)(j * c + k)];
            inv[(size_t)(j * c + i)] = v;
        }
    return true;
}

__global__ void k_scale_cols_add(const double* __restrict__ base_src, double base, int64_t n, int64_t k,
                                 double* __restrict__ out) {
    // out += base * src   (both n x k compact)
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * k;
         e += (int64_t)gridDim.x * blockDim.x)
        out[e] += base * base_src[e];
}

// M = (L^-1)^T as a full s x s matrix (zeros below the diagonal) from the lower L^-1 in Li.
__global__ void k_upper_from_lower_t(const double* __restrict__ Li, int64_t s, double* __restrict__ M) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= s * s) return;
    const int64_t i = e % s, j = e / s;   // M(i, j), column-major
    M[e] = i <= j ? Li[j + i * s] : 0.0;
}

}  // namespace csamd

using namespace csamd;

struct cs_lowrank {
    int64_t n = 0, c = 0, r = 0, sketch = 0;
    rocblas_handle blas = nullptr;
    hipStream_t st = nullptr;
    DevBuf U;        // n x r (descending eigenvalues)
    DevBuf X, y;     // n x c, n (double)
    std::vector<double> lambda, UtX, Uty, XtX, Xty;
    double yty = 0.0, r0 = 0.0, trace = 0.0;
    double sg = NAN, se = NAN, delta = NAN, log_lik = NAN;
    ~cs_lowrank() {
        if (blas) (void)rocblas_destroy_handle(blas);
        if (st) (void)hipStreamDestroy(st);
    }
};

namespace {

void kv(cs_kinship* K, const double* dX, int64_t n, int64_t k, double* dY) {
    const int rc = cs_kinship_apply_device(K, dX, n, k, dY, n);
    if (rc == CS_ERR_INVALID_ARGUMENT) throw InvalidArgument(cs_last_error());
    if (rc != CS_OK) throw RuntimeError(cs_last_error());
}

// Thin Q of an n x s matrix in place.  CholeskyQR2: twice G = A^T A, G = L L^T (potrf),
// A <- A L^-T (trtri + one GEMM) -- GEMM-shaped work -- unless the Cholesky fails or its
// diagonal spans more than 1e6 (cond(A) above ~1e6, where A^T A loses too much), then
// Householder QR (geqrf + orgqr) of the current A.  Every orthonormal basis of the sample spans the same subspace, so the
// eigenpairs of Q^T K Q and U = Q V do not depend on the choice.
void householder_q(cs_lowrank* h, double* A, int64_t n, int64_t s, DevBuf& tau) {
    CS_BLAS3(rocsolver_dgeqrf(h->blas, (int)n, (int)s, A, (int)n, tau.as<double>()));
    CS_BLAS3(rocsolver_dorgqr(h->blas, (int)n, (int)s, (int)s, A, (int)n, tau.as<double>()));
}

void thin_q(cs_lowrank* h, double* A, int64_t n, int64_t s, DevBuf& tau, double* scratch) {
    const double one = 1.0, zero = 0.0;
    DevBuf G((size_t)(s * s) * 8), Mt((size_t)(s * s) * 8), info(sizeof(rocblas_int));
    std::vector<double> dg((size_t)s);
    for (int pass = 0; pass < 2; ++pass) {
        gemm_tn_tall(h->blas, h->st, A, A, n, s, s, G.as<double>());
        CS_BLAS3(rocsolver_dpotrf(h->blas, rocblas_fill_lower, (int)s, G.as<double>(), (int)s,
                                  info.as<rocblas_int>()));
        rocblas_int inf = 0;
        CS_HIP(hipMemcpyAsync(&inf, info.p, sizeof(inf), hipMemcpyDeviceToHost, h->st));
        CS_HIP(hipMemcpy2DAsync(dg.data(), 8, G.p, (size_t)(s + 1) * 8, 8, (size_t)s, hipMemcpyDeviceToHost, h->st));
        CS_HIP(hipStreamSynchronize(h->st));
        double lo = std::numeric_limits<double>::infinity(), hi = 0.0;
        for (double v : dg) {
            lo = std::min(lo, std::fabs(v));
            hi = std::max(hi, std::fabs(v));
        }
        if (inf != 0 || !(lo > 0.0) || !(hi <= 1e6 * lo)) {
            householder_q(h, A, n, s, tau);
            return;
        }
        // A <- A L^-T as one GEMM with the explicit upper-triangular (L^-1)^T
        CS_BLAS3(rocsolver_dtrtri(h->blas, rocblas_fill_lower, rocblas_diagonal_non_unit, (int)s, G.as<double>(),
                                  (int)s, info.as<rocblas_int>()));
        k_upper_from_lower_t<<<grid_for(s * s), 256, 0, h->st>>>(G.as<double>(), s, Mt.as<double>());
        CS_HIP(hipGetLastError());
        CS_BLAS3(rocblas_dgemm(h->blas, rocblas_operation_none, rocblas_operation_none, (int)n, (int)s, (int)s, &one,
                               A, (int)n, Mt.as<double>(), (int)s, &zero, scratch, (int)n));
        CS_HIP(hipMemcpyAsync(A, scratch, (size_t)(n * s) * 8, hipMemcpyDeviceToDevice, h->st));
    }
    CS_HIP(hipStreamSynchronize(h->st));
}

// Omega of the last prepare (host copy): LOCO runs ask for the same seed and size once per
// chromosome, and drawing n s normals serially takes ~1.6 s at n = 200k, s = 266.
struct OmegaCache {
    std::mutex mu;
    int64_t seed = 0, n = -1, s = -1;
    std::vector<double> v;
};
OmegaCache& omega_cache() {
    static OmegaCache c;
    return c;
}

// REML objective of the residual-augmented low-rank model at log(eta) (lmm_engine.cpp:1001-1023),
// from the small projections.  Returns +inf where the reference does.
double lowrank_objective(const cs_lowrank* h, double log_eta, double* quadratic) {
    const int64_t n = h->n, c = h->c, r = h->r;
    const double eta = std::exp(log_eta);
    const double base = 1.0 / (1.0 + eta * h->r0);
    std::vector<double> corr((size_t)r);
    for (int64_t l = 0; l < r; ++l) corr[(size_t)l] = 1.0 / (1.0 + eta * h->lambda[(size_t)l]) - base;
    // information = X^T R^-1 X, X^T R^-1 y
    std::vector<double> info((size_t)(c * c)), xry((size_t)c);
    for (int64_t a = 0; a < c; ++a) {
        for (int64_t b = 0; b < c; ++b) {
            double v = base * h->XtX[(size_t)(b * c + a)];
            for (int64_t l = 0; l < r; ++l)
                v += h->UtX[(size_t)(a * r + l)] * corr[(size_t)l] * h->UtX[(size_t)(b * r + l)];
            info[(size_t)(b * c + a)] = v;
        }
        double v = base * h->Xty[(size_t)a];
        for (int64_t l = 0; l < r; ++l) v += h->UtX[(size_t)(a * r + l)] * corr[(size_t)l] * h->Uty[(size_t)l];
        xry[(size_t)a] = v;
    }
    std::vector<double> inv;
    double logdet_x = 0.0;
    if (!chol_inv_logdet(info, c, inv, logdet_x)) return std::numeric_limits<double>::infinity();
    std::vector<double> beta((size_t)c, 0.0);
    for (int64_t a = 0; a < c; ++a)
        for (int64_t b = 0; b < c; ++b) beta[(size_t)a] += inv[(size_t)(b * c + a)] * xry[(size_t)b];
    // residual quadratic form: base ||res||^2 + sum corr (U^T res)^2
    double rr = h->yty;
    for (int64_t a = 0; a < c; ++a) {
        rr -= 2.0 * beta[(size_t)a] * h->Xty[(size_t)a];
        for (int64_t b = 0; b < c; ++b) rr += beta[(size_t)a] * h->XtX[(size_t)(b * c + a)] * beta[(size_t)b];
    }
    double qf = base * rr;
    for (int64_t l = 0; l < r; ++l) {
        double ur = h->Uty[(size_t)l];
        for (int64_t a = 0; a < c; ++a) ur -= h->UtX[(size_t)(a * r + l)] * beta[(size_t)a];
        qf += corr[(size_t)l] * ur * ur;
    }
    if (!(qf > 0.0) || !std::isfinite(qf)) return std::numeric_limits<double>::infinity();
    if (quadratic) *quadratic = qf;
    const double df = (double)(n - c);
    double logdet_r = (double)(n - r) * std::log1p(eta * h->r0);
    for (int64_t l = 0; l < r; ++l) logdet_r += std::log(1.0 + eta * h->lambda[(size_t)l]);
    return logdet_r + logdet_x + df * std::log(qf / df);
}

}  // namespace

extern "C" {

int cs_lowrank_prepare(cs_kinship* K, const double* y, const float* X, int64_t ldx, int64_t c, int rank,
                       int oversample, int64_t seed, double trace_threshold, int min_rank,
                       int no_fallback, cs_lowrank** out, cs_lowrank_diag* diag) {
    return guarded([&] {
        if (!K || !y || !out || (c > 0 && !X)) throw InvalidArgument("cs_lowrank_prepare: null argument");
        int64_t n = 0;
        if (cs_kinship_rows(K, &n) != CS_OK) throw RuntimeError(cs_last_error());
        if (rank <= 0 || rank >= n || (c > 0 && ldx < n))
            throw InvalidArgument("Low-rank LMM requires valid dimensions and 0 < rank < sample count");
        if (n - c <= 0) throw InvalidArgument("Low-rank LMM has no residual degrees of freedom");
        const auto t0 = std::chrono::steady_clock::now();
        auto h = std::make_unique<cs_lowrank>();
        h->n = n;
        h->c = c;
        CS_HIP(hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking));
        CS_BLAS3(rocblas_create_handle(&h->blas));
        CS_BLAS3(rocblas_set_stream(h->blas, h->st));
        const int64_t s = std::min<int64_t>(n, rank + std::max(0, oversample));
        h->sketch = s;
        // Omega, column by column, from the standard generators (as the reference fills it)
        DevBuf A((size_t)(n * s) * 8), Yb((size_t)(n * s) * 8), tau((size_t)s * 8);
        {
            OmegaCache& oc = omega_cache();
            std::lock_guard<std::mutex> lock(oc.mu);
            if (oc.seed != seed || oc.n != n || oc.s != s) {
                oc.n = -1;
                oc.v.assign((size_t)(n * s), 0.0);
                std::mt19937_64 gen((uint64_t)seed);
                std::normal_distribution<double> normal(0.0, 1.0);
                for (int64_t j = 0; j < s; ++j)
                    for (int64_t i = 0; i < n; ++i) oc.v[(size_t)(j * n + i)] = normal(gen);
                oc.seed = seed;
                oc.n = n;
                oc.s = s;
            }
            CS_HIP(hipMemcpyAsync(A.p, oc.v.data(), oc.v.size() * 8, hipMemcpyHostToDevice, h->st));
            CS_HIP(hipStreamSynchronize(h->st));
        }
        kv(K, A.as<double>(), n, s, Yb.as<double>());            // sample = K Omega
        for (int it = 0; it < 2; ++it) {
            thin_q(h.get(), Yb.as<double>(), n, s, tau, A.as<double>());
            CS_HIP(hipStreamSynchronize(h->st));
            kv(K, Yb.as<double>(), n, s, A.as<double>());        // sample = K q
            std::swap(A.p, Yb.p);
            std::swap(A.bytes, Yb.bytes);
        }
        thin_q(h.get(), Yb.as<double>(), n, s, tau, A.as<double>());   // q
        CS_HIP(hipStreamSynchronize(h->st));
        kv(K, Yb.as<double>(), n, s, A.as<double>());            // kq
        const double one = 1.0, zero = 0.0;
        DevBuf B((size_t)(s * s) * 8), ev((size_t)s * 8), E((size_t)s * 8), info(sizeof(rocblas_int));
        gemm_tn_tall(h->blas, h->st, Yb.as<double>(), A.as<double>(), n, s, s, B.as<double>());
        CS_BLAS3(rocsolver_dsyevd(h->blas, rocblas_evect_original, rocblas_fill_lower, (int)s, B.as<double>(), (int)s,
                                  ev.as<double>(), E.as<double>(), info.as<rocblas_int>()));
        std::vector<double> evals((size_t)s);
        rocblas_int inf = 0;
        CS_HIP(hipMemcpyAsync(evals.data(), ev.p, (size_t)s * 8, hipMemcpyDeviceToHost, h->st));
        CS_HIP(hipMemcpyAsync(&inf, info.p, sizeof(inf), hipMemcpyDeviceToHost, h->st));
        CS_HIP(hipStreamSynchronize(h->st));
        if (inf != 0) throw RuntimeError("Low-rank GRM eigensolver failed");
        // leading `rank` pairs, descending
        std::vector<double> lead((size_t)rank);
        for (int j = 0; j < rank; ++j) lead[(size_t)j] = std::max(0.0, evals[(size_t)(s - 1 - j)]);
        double total_trace = 0.0;
        if (cs_kinship_trace(K, &total_trace) != CS_OK) throw RuntimeError(cs_last_error());
        int selected = rank;
        bool threshold_met = trace_threshold <= 0.0;
        if (trace_threshold > 0.0 && total_trace > 0.0) {
            double retained = 0.0;
            const int first_allowed = std::min(rank, std::max(1, min_rank));
            for (int j = 0; j < rank; ++j) {
                retained += lead[(size_t)j];
                if (j + 1 >= first_allowed && retained / total_trace >= trace_threshold) {
                    selected = j + 1;
                    threshold_met = true;
                    break;
                }
            }
        }
        const bool fell_back = !threshold_met && !no_fallback;
        if (fell_back) selected = rank;   // MatrixFreeVce with the threshold off keeps every pair
        h->r = selected;
        h->lambda.assign(lead.begin(), lead.begin() + selected);
        // U = q V[:, s-1 .. s-selected]: reverse the eigenvector columns into a contiguous block
        DevBuf Vrev((size_t)(s * selected) * 8);
        for (int j = 0; j < selected; ++j)
            CS_HIP(hipMemcpyAsync(Vrev.as<double>() + (int64_t)j * s, B.as<double>() + (s - 1 - j) * s, (size_t)s * 8,
                                  hipMemcpyDeviceToDevice, h->st));
        h->U.alloc((size_t)(n * selected) * 8);
        CS_BLAS3(rocblas_dgemm(h->blas, rocblas_operation_none, rocblas_operation_none, (int)n, selected, (int)s, &one,
                               Yb.as<double>(), (int)n, Vrev.as<double>(), (int)s, &zero, h->U.as<double>(), (int)n));
        double lsum = 0.0;
        for (double v : h->lambda) lsum += v;
        h->trace = total_trace;
        h->r0 = std::max(0.0, (total_trace - lsum) / (double)(n - selected));
        // projections of y and X
        std::vector<double> Xd((size_t)(n * std::max<int64_t>(c, 1)));
        for (int64_t a = 0; a < c; ++a)
            for (int64_t i = 0; i < n; ++i) Xd[(size_t)(a * n + i)] = (double)X[a * ldx + i];
        h->X.alloc((size_t)(n * std::max<int64_t>(c, 1)) * 8);
        h->y.alloc((size_t)n * 8);
        CS_HIP(hipMemcpyAsync(h->X.p, Xd.data(), (size_t)(n * c) * 8, hipMemcpyHostToDevice, h->st));
        CS_HIP(hipMemcpyAsync(h->y.p, y, (size_t)n * 8, hipMemcpyHostToDevice, h->st));
        const int64_t r = selected;
        DevBuf UtXd((size_t)(r * std::max<int64_t>(c, 1)) * 8), Utyd((size_t)r * 8);
        if (c > 0) gemm_tn_tall(h->blas, h->st, h->U.as<double>(), h->X.as<double>(), n, r, c, UtXd.as<double>());
        gemm_tn_tall(h->blas, h->st, h->U.as<double>(), h->y.as<double>(), n, r, 1, Utyd.as<double>());
        h->UtX.resize((size_t)(r * c));
        h->Uty.resize((size_t)r);
        if (c > 0) CS_HIP(hipMemcpyAsync(h->UtX.data(), UtXd.p, (size_t)(r * c) * 8, hipMemcpyDeviceToHost, h->st));
        CS_HIP(hipMemcpyAsync(h->Uty.data(), Utyd.p, (size_t)r * 8, hipMemcpyDeviceToHost, h->st));
        CS_HIP(hipStreamSynchronize(h->st));
        // X^T X, X^T y, y^T y in FP64 on the host copies (n (c+1)^2 flops)
        h->XtX.assign((size_t)(c * c), 0.0);
        h->Xty.assign((size_t)c, 0.0);
        h->yty = 0.0;
        for (int64_t i = 0; i < n; ++i) h->yty += y[i] * y[i];
        for (int64_t a = 0; a < c; ++a) {
            const double* xa = &Xd[(size_t)(a * n)];
            double v = 0.0;
            for (int64_t i = 0; i < n; ++i) v += xa[i] * y[i];
            h->Xty[(size_t)a] = v;
            for (int64_t b = 0; b <= a; ++b) {
                const double* xb = &Xd[(size_t)(b * n)];
                double w = 0.0;
                for (int64_t i = 0; i < n; ++i) w += xa[i] * xb[i];
                h->XtX[(size_t)(b * c + a)] = w;
                h->XtX[(size_t)(a * c + b)] = w;
            }
        }
        // REML: 81-point grid on [-12, 12], then golden section (<= 80 iterations, 1e-8)
        const int grid = 81;
        int best_index = 0;
        double best = lowrank_objective(h.get(), -12.0, nullptr);
        for (int i = 1; i < grid; ++i) {
            const double cand = -12.0 + 24.0 * i / (double)(grid - 1);
            const double v = lowrank_objective(h.get(), cand, nullptr);
            if (v < best) {
                best = v;
                best_index = i;
            }
        }
        double left = -12.0 + 24.0 * std::max(0, best_index - 1) / (double)(grid - 1);
        double right = -12.0 + 24.0 * std::min(grid - 1, best_index + 1) / (double)(grid - 1);
        const GoldenSpec spec{(std::sqrt(5.0) - 1.0) / 2.0, false, 1e-8, 80, false};
        const double best_log_eta =
            golden_section_min([&](double x) { return lowrank_objective(h.get(), x, nullptr); }, left, right, spec);
        double quadratic = 0.0;
        best = lowrank_objective(h.get(), best_log_eta, &quadratic);
        if (!std::isfinite(best)) throw RuntimeError("Low-rank REML optimization failed");
        h->delta = std::exp(-best_log_eta);
        h->se = quadratic / (double)(n - c);
        h->sg = std::exp(best_log_eta) * h->se;
        h->log_lik = -0.5 * best;
        if (diag) {
            std::memset(diag, 0, sizeof(*diag));
            diag->requested_rank = rank;
            diag->effective_rank = (int)selected;
            diag->sketch_size = (int)s;
            diag->total_rank = (int)n;
            diag->residual_eigenvalue = h->r0;
            diag->trace = total_trace;
            diag->trace_fraction = total_trace > 0.0 ? lsum / total_trace : 0.0;
            diag->trace_threshold_met = threshold_met ? 1 : 0;
            diag->fell_back = fell_back ? 1 : 0;
            // Frobenius diagnostic of the reference (lmm_engine.cpp:979-990)
            double sq = 0.0, cross = 0.0;
            for (double v : h->lambda) {
                sq += v * v;
                cross += (v - h->r0) * v;
            }
            const double approx = sq + (double)(n - selected) * h->r0 * h->r0;
            cross += h->r0 * total_trace;
            const double resid = std::max(0.0, total_trace * total_trace + approx - 2.0 * cross);
            diag->relative_frobenius_error =
                std::sqrt(resid) / std::max(total_trace, std::numeric_limits<double>::min());
            diag->sigma2_g = h->sg;
            diag->sigma2_e = h->se;
            diag->delta = h->delta;
            diag->decomposition_seconds =
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            diag->estimated_peak_bytes = 8LL * (2LL * n * s + n * (long long)selected + s * s);
        }
        *out = h.release();
    });
}

int cs_lowrank_destroy(cs_lowrank* h) {
    return guarded([&] { delete h; });
}

int cs_lowrank_get(cs_lowrank* h, int64_t* rank, double* lambda, double* U, int64_t ldu, double* r0,
                   double* sigma2_g, double* sigma2_e) {
    return guarded([&] {
        if (!h) throw InvalidArgument("cs_lowrank_get: null handle");
        if (rank) *rank = h->r;
        if (lambda) std::copy(h->lambda.begin(), h->lambda.end(), lambda);
        if (U) {
            if (ldu < h->n) throw RuntimeError("cs_lowrank_get: dimension mismatch");
            CS_HIP(hipMemcpy2DAsync(U, (size_t)ldu * 8, h->U.p, (size_t)h->n * 8, (size_t)h->n * 8, (size_t)h->r,
                                    hipMemcpyDeviceToHost, h->st));
            CS_HIP(hipStreamSynchronize(h->st));
        }
        if (r0) *r0 = h->r0;
        if (sigma2_g) *sigma2_g = h->sg;
        if (sigma2_e) *sigma2_e = h->se;
    });
}

int cs_lowrank_scan(cs_lowrank* h, cs_bed* bed, const int32_t* keep, int64_t n_keep, const int32_t* snps,
                    int64_t n_snps, double min_maf, int prec, double* freq, int64_t* n_obs, int32_t* pass,
                    double* beta, double* se, double* p, double* log10p) {
    return guarded([&] {
        if (!h || !bed) throw InvalidArgument("cs_lowrank_scan: null argument");
        const int64_t n = h->n, c = h->c, r = h->r;
        if ((keep && n_keep > 0 ? n_keep : bed->N) != n) throw RuntimeError("scan: dimension mismatch");
        if (c > 32) throw InvalidArgument("low-rank scan: at most 32 covariate columns");
        const double base = 1.0 / std::max(1e-12, h->sg * h->r0 + h->se);
        std::vector<double> corr((size_t)r);
        for (int64_t l = 0; l < r; ++l) corr[(size_t)l] = 1.0 / (h->sg * h->lambda[(size_t)l] + h->se) - base;
        // information = X^T V^-1 X and beta (lmm_engine.cpp:1275-1283)
        std::vector<double> info((size_t)(c * c)), xvy((size_t)c);
        for (int64_t a = 0; a < c; ++a) {
            for (int64_t b = 0; b < c; ++b) {
                double v = base * h->XtX[(size_t)(b * c + a)];
                for (int64_t l = 0; l < r; ++l)
                    v += h->UtX[(size_t)(a * r + l)] * corr[(size_t)l] * h->UtX[(size_t)(b * r + l)];
                info[(size_t)(b * c + a)] = v;
            }
            double v = base * h->Xty[(size_t)a];
            for (int64_t l = 0; l < r; ++l) v += h->UtX[(size_t)(a * r + l)] * corr[(size_t)l] * h->Uty[(size_t)l];
            xvy[(size_t)a] = v;
        }
        std::vector<double> inv;
        double ld = 0.0;
        if (c > 0 && !chol_inv_logdet(info, c, inv, ld))
            throw RuntimeError("Low-rank LMM fixed-effect information matrix is singular");
        std::vector<double> b((size_t)c, 0.0);
        for (int64_t a = 0; a < c; ++a)
            for (int64_t q = 0; q < c; ++q) b[(size_t)a] += inv[(size_t)(q * c + a)] * xvy[(size_t)q];
        // U^T residual = U^T y - U^T X beta ;  coefficient vectors for U
        std::vector<double> wres((size_t)r), WX((size_t)(r * std::max<int64_t>(c, 1)));
        for (int64_t l = 0; l < r; ++l) {
            double ur = h->Uty[(size_t)l];
            for (int64_t a = 0; a < c; ++a) ur -= h->UtX[(size_t)(a * r + l)] * b[(size_t)a];
            wres[(size_t)l] = corr[(size_t)l] * ur;
            for (int64_t a = 0; a < c; ++a) WX[(size_t)(a * r + l)] = corr[(size_t)l] * h->UtX[(size_t)(a * r + l)];
        }
        // R = [P y, U, V^-1 X] on the device
        const int64_t nr = 1 + r + c;
        DevBuf R((size_t)(n * nr) * 8), dw((size_t)r * 8), dWX((size_t)(r * std::max<int64_t>(c, 1)) * 8),
            dbeta((size_t)std::max<int64_t>(c, 1) * 8);
        double* Rp = R.as<double>();
        CS_HIP(hipMemcpyAsync(dw.p, wres.data(), (size_t)r * 8, hipMemcpyHostToDevice, h->st));
        if (c > 0) {
            CS_HIP(hipMemcpyAsync(dWX.p, WX.data(), (size_t)(r * c) * 8, hipMemcpyHostToDevice, h->st));
            CS_HIP(hipMemcpyAsync(dbeta.p, b.data(), (size_t)c * 8, hipMemcpyHostToDevice, h->st));
        }
        const double one = 1.0, zero = 0.0, mone = -1.0;
        // res = y - X beta ; P y = U w + base res
        DevBuf res((size_t)n * 8);
        CS_HIP(hipMemcpyAsync(res.p, h->y.p, (size_t)n * 8, hipMemcpyDeviceToDevice, h->st));
        if (c > 0)
            CS_BLAS3(rocblas_dgemv(h->blas, rocblas_operation_none, (int)n, (int)c, &mone, h->X.as<double>(), (int)n,
                                   dbeta.as<double>(), 1, &one, res.as<double>(), 1));
        CS_BLAS3(rocblas_dgemv(h->blas, rocblas_operation_none, (int)n, (int)r, &one, h->U.as<double>(), (int)n,
                               dw.as<double>(), 1, &zero, Rp, 1));
        k_scale_cols_add<<<grid_for(n), 256, 0, h->st>>>(res.as<double>(), base, n, 1, Rp);
        CS_HIP(hipGetLastError());
        CS_HIP(hipMemcpyAsync(Rp + n, h->U.p, (size_t)(n * r) * 8, hipMemcpyDeviceToDevice, h->st));
        if (c > 0) {
            CS_BLAS3(rocblas_dgemm(h->blas, rocblas_operation_none, rocblas_operation_none, (int)n, (int)c, (int)r,
                                   &one, h->U.as<double>(), (int)n, dWX.as<double>(), (int)r, &zero,
                                   Rp + (1 + r) * n, (int)n));
            k_scale_cols_add<<<grid_for(n * c), 256, 0, h->st>>>(h->X.as<double>(), base, n, c, Rp + (1 + r) * n);
            CS_HIP(hipGetLastError());
        }
        CS_HIP(hipStreamSynchronize(h->st));
        cs_scan_model m{1, base, 1, 1 + r, corr.data(), 1 + r, 1 + r + c, c > 0 ? inv.data() : nullptr};
        ScanArgs a;
        a.keep = (keep && n_keep > 0) ? keep : nullptr;
        a.n = n;
        a.R = Rp;
        a.ldr = n;
        a.nr = nr;
        a.min_maf = min_maf;
        a.snps = (snps && n_snps > 0) ? snps : nullptr;
        a.n_snps = a.snps ? n_snps : bed->M;
        a.prec = prec;
        a.device_R = true;
        ScanOut o;
        o.freq = freq; o.n_obs = n_obs; o.pass = pass;
        o.beta = beta; o.se = se; o.p = p; o.log10p = log10p;
        run_scan(bed, a, &m, o);
    });
}

}  // extern "C"
