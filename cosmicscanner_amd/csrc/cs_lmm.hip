// Dense LMM null model on the device and the rotated per-SNP scan that consumes it.
//
//   EigenVce::prepare with covariate projection (libCosmicCore/src/eigen_vce.cpp:19-49):
//     Householder QR of X (n x c), A = Q[:, c:], G_reml = A^T G A, eigendecomposition
//     G_reml = U diag(D) U^T (D clamped at 0), W_proj = A U, Uty = U^T A^T y.   rocSOLVER / rocBLAS;
//     A^T G A, A^T y and A U come from the c reflectors directly (O(c n^2), no explicit Q).
//   VCESolver::solveDiagonal (libCosmicCore/src/vce.cpp:3073-3171), the default "brent" null
//     solver: REML golden section on log10(delta) in [-5, 5] while |c - d| > 1e-5; one workgroup
//     evaluates every likelihood with a fixed-order reduction.
//   LmmEngine::scan (src/lmm_engine.cpp:1363-1525): score = (W_proj^T g)^T (V^-1 Uty),
//     var = sum_l (W_proj^T g)_l^2 / (sigma_g D_l + sigma_e), through the fused scan kernel with
//     R = [W_proj V^-1 Uty, W_proj] and quadratic weights V^-1.
//   The dense LOCO GRM of GwasScanner::runLocoScan (src/gwas_scanner.cpp:481-495).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cmath>
#include <memory>
#include <vector>

#include "cs_blas.h"
#include "cs_common.h"
#include "cs_golden.h"
#include "cs_ops.h"

namespace csamd {

#define CS_BLAS(call)                                                                   \
    do {                                                                                \
        rocblas_status s_ = (call);                                                     \
        if (s_ != rocblas_status_success)                                               \
            throw ::csamd::RuntimeError(std::string("rocBLAS/rocSOLVER error ") +      \
                                        rocblas_status_to_string(s_) + " at " #call);   \
    } while (0)

__global__ void k_clamp_nonneg(double* __restrict__ d, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n && d[i] < 0.0) d[i] = 0.0;
}

// Two sums over the eigen-system for one delta: sum log(D + delta), sum Uty^2 / (D + delta);
// each thread strides, then a fixed tree -- the same order on every call.
__device__ void reml_sums(const double* __restrict__ D, const double* __restrict__ Uty, int64_t ne,
                          double delta, double* sh, double& slog, double& squad) {
    double a = 0.0, b = 0.0;
    for (int64_t i = threadIdx.x; i < ne; i += blockDim.x) {
        const double v = D[i] + delta;
        a += log(v);
        b += Uty[i] * Uty[i] / v;
    }
    sh[threadIdx.x] = a;
    sh[blockDim.x + threadIdx.x] = b;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sh[threadIdx.x] += sh[threadIdx.x + w];
            sh[blockDim.x + threadIdx.x] += sh[blockDim.x + threadIdx.x + w];
        }
        __syncthreads();
    }
    slog = sh[0];
    squad = sh[blockDim.x];
    __syncthreads();
}

// solveDiagonal with an empty UtX (covariates projected out): n_eff = ne,
// -logL = 0.5 (sum log V + n_eff log(quad / n_eff) + n_eff).  out = {sigma_g, sigma_e, delta, evals}.
__global__ __launch_bounds__(256) void k_reml_golden(const double* __restrict__ D,
                                                     const double* __restrict__ Uty, int64_t ne,
                                                     double* __restrict__ out) {
    __shared__ double sh[512];
    const double n_eff = (double)ne;
    int evals = 0;
    auto negll = [&](double delta) -> double {
        double slog, squad;
        reml_sums(D, Uty, ne, delta, sh, slog, squad);
        ++evals;
        return 0.5 * (slog + n_eff * log(squad / n_eff) + n_eff);
    };
    const GoldenSpec spec{(sqrt(5.0) + 1.0) / 2.0, true, 1e-5, -1, true};
    const double best = pow(10.0, golden_section_min([&](double x) { return negll(pow(10.0, x)); }, -5.0, 5.0, spec));
    double slog, squad;
    reml_sums(D, Uty, ne, best, sh, slog, squad);
    if (threadIdx.x == 0) {
        const double vg = squad / n_eff;
        out[0] = vg;
        out[1] = vg * best;
        out[2] = best;
        out[3] = evals;
    }
}

// P_y = Uty / (sigma_g D + sigma_e) and the quadratic weights 1 / (sigma_g D + sigma_e)
// (lmm_engine.cpp:1372-1373, 1404).
__global__ void k_vinv(const double* __restrict__ D, const double* __restrict__ Uty, int64_t ne,
                       double sg, double se, double* __restrict__ vinv, double* __restrict__ py) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= ne) return;
    const double v = 1.0 / (sg * D[i] + se);
    vinv[i] = v;
    py[i] = Uty[i] * v;
}

// Dense LOCO GRM over the keep list, in the reference's operation order:
// G_global = raw / S (GrmBuilder's normalisation), G_sub = G_global[keep, keep], and with
// S_chr > 0 and S_loco = S - S_chr > 1e-6: G_sub * (S / S_loco) - raw_chr[keep, keep] / S_loco.
__global__ void k_loco(const double* __restrict__ raw, const double* __restrict__ raw_chr, int64_t N,
                       const int32_t* __restrict__ keep, int64_t n, double S, double S_loco,
                       bool loco, double* __restrict__ out) {
    const double ratio = S / S_loco;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e % n, j = e / n;
        const int64_t src = (int64_t)keep[j] * N + keep[i];
        const double g = raw[src] / S;
        out[e] = loco ? g * ratio - raw_chr[src] / S_loco : g;
    }
}

// LOCO GRM of the multivariate driver (src/models/multivariate_model.cpp:1040-1063):
// (raw / norm) w1 - (raw_chr / norm_chr) w2, w1 = S / S_loco, w2 = S_chr / S_loco.
__global__ void k_loco_scaled(const double* __restrict__ raw, const double* __restrict__ raw_chr, int64_t N,
                              const int32_t* __restrict__ keep, int64_t n, double norm, double norm_chr,
                              double w1, double w2, bool loco, double* __restrict__ out) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e % n, j = e / n;
        const int64_t src = (int64_t)keep[j] * N + keep[i];
        const double g = raw[src] / norm;
        out[e] = loco ? g * w1 - (raw_chr[src] / norm_chr) * w2 : g;
    }
}

}  // namespace csamd

using namespace csamd;

struct cs_lmm {
    int64_t n = 0, c = 0, ne = 0;
    rocblas_handle blas = nullptr;
    hipStream_t st = nullptr;
    DevBuf Q, tau;   // Householder reflectors of X (n x c, geqrf) and their scalars; unused when c == 0
    std::vector<double> tau_h;
    DevBuf y;        // n
    DevBuf G;        // working copy of G (n x n): projected, then the eigenvectors at (c, c)
    DevBuf D, E, Uty, t, vinv, py;
    DevBuf R;        // n x (1 + ne): [W_proj P_y, W_proj]
    DevBuf info, res;
    bool prepared = false, have_null = false;
    double sg = NAN, se = NAN, delta = NAN;
    int evals = 0;
    ~cs_lmm() {
        if (blas) (void)rocblas_destroy_handle(blas);
        if (st) (void)hipStreamDestroy(st);
    }
};

static void lmm_prepare(cs_lmm* h, const double* G, int64_t ldg, bool g_dev) {
    const int64_t n = h->n, ne = h->ne, c = h->c;
    if (ldg < n) throw RuntimeError("EigenVce::prepare: Dimension mismatch");
    const double one = 1.0, zero = 0.0;
    h->G.ensure((size_t)(n * n) * 8);
    double* Gw = h->G.as<double>();
    CS_HIP(hipMemcpy2DAsync(Gw, (size_t)n * 8, G, (size_t)ldg * 8, (size_t)n * 8, (size_t)n,
                            g_dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, h->st));
    h->D.ensure((size_t)ne * 8);
    h->E.ensure((size_t)ne * 8);
    h->Uty.ensure((size_t)ne * 8);
    h->t.ensure((size_t)n * 8);
    h->R.ensure((size_t)(n * (ne + 1)) * 8);
    double* W = h->R.as<double>() + n;
    CS_HIP(hipMemcpyAsync(h->t.p, h->y.p, (size_t)n * 8, hipMemcpyDeviceToDevice, h->st));
    if (c > 0) {
        // G_reml = A^T G A (A = Q[:, c:]) in place from the reflectors; t = Q^T y (rows c: = A^T y)
        householder_project_sym(h->blas, h->st, h->Q.as<double>(), h->tau_h.data(), n, c, Gw);
        CS_BLAS(rocsolver_dormqr(h->blas, rocblas_side_left, rocblas_operation_transpose, (int)n, 1, (int)c,
                                 h->Q.as<double>(), (int)n, h->tau.as<double>(), h->t.as<double>(), (int)n));
    }
    double* U = Gw + c * n + c;   // ne x ne, ld n
    const bool large = eig_use_large(ne);
    if (large) {
        sym_eig_large(h->blas, h->st, ne, U, n, h->D.as<double>(), eig_large_leaf());
        CS_HIP(hipMemsetAsync(h->info.p, 0, sizeof(rocblas_int), h->st));
    } else {
        CS_BLAS(rocsolver_dsyevd(h->blas, rocblas_evect_original, rocblas_fill_lower, (int)ne, U, (int)n,
                                 h->D.as<double>(), h->E.as<double>(), h->info.as<rocblas_int>()));
    }
    k_clamp_nonneg<<<grid_for(ne), 256, 0, h->st>>>(h->D.as<double>(), ne);
    CS_HIP(hipGetLastError());
    // Uty = U^T (A^T y); W_proj = A U = Q [0; U]
    CS_BLAS(rocblas_dgemv_64(h->blas, rocblas_operation_transpose, ne, ne, &one, U, n, h->t.as<double>() + c, 1,
                             &zero, h->Uty.as<double>(), 1));
    if (c > 0) CS_HIP(hipMemset2DAsync(W, (size_t)n * 8, 0, (size_t)c * 8, (size_t)ne, h->st));
    CS_HIP(hipMemcpy2DAsync(W + c, (size_t)n * 8, U, (size_t)n * 8, (size_t)ne * 8, (size_t)ne,
                            hipMemcpyDeviceToDevice, h->st));
    if (c > 0 && large)
        apply_reflectors(h->blas, h->st, h->Q.as<double>(), n, h->tau_h.data(), n, c, false, W, n, ne);
    else if (c > 0)
        CS_BLAS(rocsolver_dormqr(h->blas, rocblas_side_left, rocblas_operation_none, (int)n, (int)ne, (int)c,
                                 h->Q.as<double>(), (int)n, h->tau.as<double>(), W, (int)n));
    rocblas_int info = 0;
    CS_HIP(hipMemcpyAsync(&info, h->info.p, sizeof(info), hipMemcpyDeviceToHost, h->st));
    CS_HIP(hipStreamSynchronize(h->st));
    if (info != 0) throw RuntimeError("EigenVce::prepare: eigendecomposition did not converge");
    h->prepared = true;
    h->have_null = false;
}

extern "C" {

int cs_lmm_create(const double* y, const float* X, int64_t ldx, int64_t n, int64_t c, cs_lmm** out) {
    return guarded([&] {
        if (!y || !out || (c > 0 && !X)) throw InvalidArgument("cs_lmm_create: null argument");
        if (n < 2 || c < 0 || c >= n || (c > 0 && ldx < n))
            throw RuntimeError("EigenVce::prepare: Dimension mismatch");
        auto h = std::make_unique<cs_lmm>();
        h->n = n;
        h->c = c;
        h->ne = n - c;
        CS_HIP(hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking));
        CS_BLAS(rocblas_create_handle(&h->blas));
        CS_BLAS(rocblas_set_stream(h->blas, h->st));
        h->info.alloc(sizeof(rocblas_int));
        h->res.alloc(4 * sizeof(double));
        h->y.alloc((size_t)n * 8);
        CS_HIP(hipMemcpyAsync(h->y.p, y, (size_t)n * 8, hipMemcpyHostToDevice, h->st));
        if (c > 0) {
            // X is float32 (univariate_model.cpp:368-372); the QR runs on X.cast<double>()
            std::vector<double> Xd((size_t)(n * c));
            for (int64_t j = 0; j < c; ++j)
                for (int64_t i = 0; i < n; ++i) Xd[(size_t)(j * n + i)] = (double)X[j * ldx + i];
            h->Q.alloc((size_t)(n * c) * 8);
            h->tau.alloc((size_t)c * 8);
            CS_HIP(hipMemcpyAsync(h->Q.p, Xd.data(), Xd.size() * 8, hipMemcpyHostToDevice, h->st));
            CS_BLAS(rocsolver_dgeqrf(h->blas, (int)n, (int)c, h->Q.as<double>(), (int)n, h->tau.as<double>()));
            h->tau_h.resize((size_t)c);
            CS_HIP(hipMemcpyAsync(h->tau_h.data(), h->tau.p, (size_t)c * 8, hipMemcpyDeviceToHost, h->st));
            CS_HIP(hipStreamSynchronize(h->st));
        }
        CS_HIP(hipStreamSynchronize(h->st));
        *out = h.release();
    });
}

int cs_lmm_destroy(cs_lmm* h) {
    return guarded([&] { delete h; });
}

int cs_lmm_prepare_dense(cs_lmm* h, const double* G, int64_t ldg, int g_on_device) {
    return guarded([&] {
        if (!h || !G) throw InvalidArgument("cs_lmm_prepare_dense: null argument");
        lmm_prepare(h, G, ldg, g_on_device != 0);
    });
}

int cs_lmm_null_brent(cs_lmm* h, double* sigma2_g, double* sigma2_e, double* delta) {
    return guarded([&] {
        if (!h) throw InvalidArgument("cs_lmm_null_brent: null handle");
        if (!h->prepared) throw RuntimeError("null model: prepare() has not run");
        k_reml_golden<<<1, 256, 0, h->st>>>(h->D.as<double>(), h->Uty.as<double>(), h->ne,
                                            h->res.as<double>());
        CS_HIP(hipGetLastError());
        double r[4];
        CS_HIP(hipMemcpyAsync(r, h->res.p, sizeof(r), hipMemcpyDeviceToHost, h->st));
        CS_HIP(hipStreamSynchronize(h->st));
        h->sg = r[0];
        h->se = r[1];
        h->delta = r[2];
        h->evals = (int)r[3];
        // null_delta = sigma_e / sigma_g (eigen_vce.cpp:105)
        h->have_null = true;
        if (sigma2_g) *sigma2_g = h->sg;
        if (sigma2_e) *sigma2_e = h->se;
        if (delta) *delta = h->se / h->sg;
    });
}

int cs_lmm_set_variance(cs_lmm* h, double sigma2_g, double sigma2_e) {
    return guarded([&] {
        if (!h) throw InvalidArgument("cs_lmm_set_variance: null handle");
        if (!h->prepared) throw RuntimeError("null model: prepare() has not run");
        h->sg = sigma2_g;
        h->se = sigma2_e;
        h->have_null = true;
    });
}

int cs_lmm_get_eigen(cs_lmm* h, double* D, double* Uty, double* W_proj, int64_t ldw) {
    return guarded([&] {
        if (!h) throw InvalidArgument("cs_lmm_get_eigen: null handle");
        if (!h->prepared) throw RuntimeError("null model: prepare() has not run");
        if (D) CS_HIP(hipMemcpyAsync(D, h->D.p, (size_t)h->ne * 8, hipMemcpyDeviceToHost, h->st));
        if (Uty) CS_HIP(hipMemcpyAsync(Uty, h->Uty.p, (size_t)h->ne * 8, hipMemcpyDeviceToHost, h->st));
        if (W_proj) {
            if (ldw < h->n) throw RuntimeError("cs_lmm_get_eigen: dimension mismatch");
            CS_HIP(hipMemcpy2DAsync(W_proj, (size_t)ldw * 8, h->R.as<double>() + h->n, (size_t)h->n * 8,
                                    (size_t)h->n * 8, (size_t)h->ne, hipMemcpyDeviceToHost, h->st));
        }
        CS_HIP(hipStreamSynchronize(h->st));
    });
}

int cs_lmm_scan_rotated(cs_lmm* h, cs_bed* bed, const int32_t* keep, int64_t n_keep,
                        const int32_t* snps, int64_t n_snps, double min_maf, int prec, double* freq,
                        int64_t* n_obs, int32_t* pass, double* beta, double* se, double* p,
                        double* log10p) {
    return guarded([&] {
        if (!h || !bed) throw InvalidArgument("cs_lmm_scan_rotated: null argument");
        if (!h->have_null) throw RuntimeError("scan: the null model has not been fitted");
        const int64_t n = h->n, ne = h->ne;
        if ((keep && n_keep > 0 ? n_keep : bed->N) != n) throw RuntimeError("scan: dimension mismatch");
        h->vinv.ensure((size_t)ne * 8);
        h->py.ensure((size_t)ne * 8);
        k_vinv<<<grid_for(ne), 256, 0, h->st>>>(h->D.as<double>(), h->Uty.as<double>(), ne, h->sg, h->se,
                                                h->vinv.as<double>(), h->py.as<double>());
        CS_HIP(hipGetLastError());
        const double one = 1.0, zero = 0.0;
        CS_BLAS(rocblas_dgemv_64(h->blas, rocblas_operation_none, n, ne, &one, h->R.as<double>() + n, n,
                                 h->py.as<double>(), 1, &zero, h->R.as<double>(), 1));
        std::vector<double> w((size_t)ne);
        CS_HIP(hipMemcpyAsync(w.data(), h->vinv.p, (size_t)ne * 8, hipMemcpyDeviceToHost, h->st));
        CS_HIP(hipStreamSynchronize(h->st));
        cs_scan_model m{1, 0.0, 1, 1 + ne, w.data(), 0, 0, nullptr};
        ScanArgs a;
        a.keep = (keep && n_keep > 0) ? keep : nullptr;
        a.n = n;
        a.R = h->R.as<double>();
        a.ldr = n;
        a.nr = 1 + ne;
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

int cs_grm_loco_subset(const double* d_raw, const double* d_raw_chr, int64_t N, const int32_t* keep,
                       int64_t n_keep, double S_global, double S_chr, double* d_out) {
    return guarded([&] {
        if (!d_raw || !d_out) throw InvalidArgument("cs_grm_loco_subset: null argument");
        std::vector<int32_t> all;
        if (!keep || n_keep <= 0) {
            all.resize((size_t)N);
            for (int64_t i = 0; i < N; ++i) all[(size_t)i] = (int32_t)i;
            keep = all.data();
            n_keep = N;
        }
        for (int64_t i = 0; i < n_keep; ++i)
            if (keep[i] < 0 || keep[i] >= N) throw RuntimeError("LOCO GRM: sample index out of range");
        const double S_loco = S_global - S_chr;
        const bool loco = d_raw_chr && S_chr > 0.0 && S_loco > 1e-6;
        DevBuf dk((size_t)n_keep * 4);
        CS_HIP(hipMemcpy(dk.p, keep, (size_t)n_keep * 4, hipMemcpyHostToDevice));
        k_loco<<<grid_for(n_keep * n_keep), 256>>>(d_raw, d_raw_chr, N, dk.as<int32_t>(), n_keep, S_global,
                                                   S_loco, loco, d_out);
        CS_HIP(hipGetLastError());
        CS_HIP(hipDeviceSynchronize());
    });
}

int cs_grm_loco_subset_scaled(const double* d_raw, const double* d_raw_chr, int64_t N, const int32_t* keep,
                              int64_t n_keep, double norm, double norm_chr, double S, double S_chr,
                              double* d_out) {
    return guarded([&] {
        if (!d_raw || !d_out) throw InvalidArgument("cs_grm_loco_subset_scaled: null argument");
        std::vector<int32_t> all;
        if (!keep || n_keep <= 0) {
            all.resize((size_t)N);
            for (int64_t i = 0; i < N; ++i) all[(size_t)i] = (int32_t)i;
            keep = all.data();
            n_keep = N;
        }
        for (int64_t i = 0; i < n_keep; ++i)
            if (keep[i] < 0 || keep[i] >= N) throw RuntimeError("LOCO GRM: sample index out of range");
        const double S_loco = S - S_chr;
        const bool loco = d_raw_chr && S_chr > 0.0 && S_loco > 1e-6;
        DevBuf dk((size_t)n_keep * 4);
        CS_HIP(hipMemcpy(dk.p, keep, (size_t)n_keep * 4, hipMemcpyHostToDevice));
        k_loco_scaled<<<grid_for(n_keep * n_keep), 256>>>(d_raw, d_raw_chr, N, dk.as<int32_t>(), n_keep,
                                                          norm > 0 ? norm : 1.0, norm_chr > 0 ? norm_chr : 1.0,
                                                          loco ? S / S_loco : 1.0, loco ? S_chr / S_loco : 0.0,
                                                          loco, d_out);
        CS_HIP(hipGetLastError());
        CS_HIP(hipDeviceSynchronize());
    });
}

}  // extern "C"
