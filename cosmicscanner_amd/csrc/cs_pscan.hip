// PCG-projected per-SNP scan with a matrix-free kinship operator (one LOCO operator per call).
//
// With V = sigma_g K + sigma_e I and P = V^-1 - V^-1 X (X^T V^-1 X)^-1 X^T V^-1 (the GLS
// projection of LmmEngine::scan, src/lmm_engine.cpp:1363-1525, written without an
// eigendecomposition):
//   1. one multi-RHS PCG solve V [Z_Y, Z_X, Z_C] = [Y, X, G_cal]  (cs_pcg_solve_cols);
//   2. C = (X^T Z_X)^-1 and P Y = Z_Y - Z_X C X^T Z_Y (rocBLAS; the c x c inverse on the host);
//   3. calibration (the grammar-gamma / BOLT-LMM device): for calibration markers g,
//      gamma = mean g^T P g / g~^T g~ with g~ = g projected off X, g^T P g = g^T Z_C -
//      (X^T Z_C)^T C (X^T Z_C);
//   4. fused scan: score_t = g^T P y_t and var = gamma g~^T g~ = gamma (g^T g - q^T (X^T X)^-1 q),
//      q = X^T g, for every SNP -- the reference's dead matrix-free scan used var = tau n with
//      tau = tr(V^-1)/n (lmm_engine.cpp:241-708); this build calibrates per marker instead.
#include <rocblas/rocblas.h>

#include <cmath>
#include <vector>

#include "cs_blas.h"
#include "cs_common.h"
#include "cs_ops.h"

namespace csamd {

#define CS_BLAS2(call)                                                                 \
    do {                                                                               \
        rocblas_status s_ = (call);                                                    \
        if (s_ != rocblas_status_success)                                              \
            throw ::csamd::RuntimeError(std::string("rocBLAS error ") +                \
                                        rocblas_status_to_string(s_) + " at " #call);  \
    } while (0)

// Cholesky inverse of a small SPD matrix (column-major, in place); false if not positive definite.
static bool spd_inverse(std::vector<double>& A, int64_t c) {
    std::vector<double> L((size_t)(c * c), 0.0);
    for (int64_t j = 0; j < c; ++j) {
        double d = A[(size_t)(j * c + j)];
        for (int64_t k = 0; k < j; ++k) d -= L[(size_t)(k * c + j)] * L[(size_t)(k * c + j)];
        if (!(d > 0.0)) return false;
        const double ljj = std::sqrt(d);
        L[(size_t)(j * c + j)] = ljj;
        for (int64_t i = j + 1; i < c; ++i) {
            double v = A[(size_t)(j * c + i)];
            for (int64_t k = 0; k < j; ++k) v -= L[(size_t)(k * c + i)] * L[(size_t)(k * c + j)];
            L[(size_t)(j * c + i)] = v / ljj;
        }
    }
    // inverse via L^-1
    std::vector<double> Li((size_t)(c * c), 0.0);
    for (int64_t j = 0; j < c; ++j) {
        Li[(size_t)(j * c + j)] = 1.0 / L[(size_t)(j * c + j)];
        for (int64_t i = j + 1; i < c; ++i) {
            double v = 0.0;
            for (int64_t k = j; k < i; ++k) v -= L[(size_t)(k * c + i)] * Li[(size_t)(j * c + k)];
            Li[(size_t)(j * c + i)] = v / L[(size_t)(i * c + i)];
        }
    }
    for (int64_t i = 0; i < c; ++i)
        for (int64_t j = 0; j < c; ++j) {
            double v = 0.0;
            for (int64_t k = std::max(i, j); k < c; ++k) v += Li[(size_t)(i * c + k)] * Li[(size_t)(j * c + k)];
            A[(size_t)(j * c + i)] = v;
        }
    return true;
}

static double quad_form(const std::vector<double>& C, const double* q, int64_t c) {
    double s = 0.0;
    for (int64_t a = 0; a < c; ++a)
        for (int64_t b = 0; b < c; ++b) s += q[a] * C[(size_t)(b * c + a)] * q[b];
    return s;
}

// RHS of the projected scan over the keep list: B = [Y, X, G_cal] (n x (T + c + L)) on the
// device, L = the calibration markers that pass the filter.
void pscan_prepare_rhs(cs_bed* bed, const int32_t* keep, int64_t n_keep, int64_t n, const double* Y, int64_t ldy,
                       int64_t T, const float* X, int64_t ldx, int64_t c, const int32_t* cal_snps, int64_t n_cal,
                       double min_maf, DevBuf& B, int64_t& L, hipStream_t st) {
    std::vector<int32_t> cpass((size_t)std::max<int64_t>(1, n_cal));
    DevBuf Gc((size_t)(n * std::max<int64_t>(1, n_cal)) * 8);
    if (n_cal > 0)
        markers_device(bed, (keep && n_keep > 0) ? keep : nullptr, n_keep, cal_snps, n_cal, min_maf, Gc.as<double>(),
                       n, cpass.data(), nullptr, st);
    std::vector<int64_t> cidx;
    for (int64_t i = 0; i < n_cal; ++i)
        if (cpass[(size_t)i]) cidx.push_back(i);
    L = (int64_t)cidx.size();
    const int64_t k = T + c + L;
    std::vector<double> host((size_t)(n * (T + c)));
    for (int64_t t = 0; t < T; ++t)
        for (int64_t i = 0; i < n; ++i) host[(size_t)(t * n + i)] = Y[t * ldy + i];
    for (int64_t a = 0; a < c; ++a)
        for (int64_t i = 0; i < n; ++i) host[(size_t)((T + a) * n + i)] = (double)X[a * ldx + i];
    B.alloc((size_t)(n * k) * 8);
    CS_HIP(hipMemcpyAsync(B.p, host.data(), host.size() * 8, hipMemcpyHostToDevice, st));
    for (int64_t l = 0; l < L; ++l)
        CS_HIP(hipMemcpyAsync(B.as<double>() + (T + c + l) * n, Gc.as<double>() + cidx[(size_t)l] * n,
                              (size_t)n * 8, hipMemcpyDeviceToDevice, st));
    CS_HIP(hipStreamSynchronize(st));
}

// Projection, calibration and fused scan from the solved Z = V^-1 B (both n x k on the device).
void pscan_finish(cs_bed* bed, const int32_t* keep, int64_t n_keep, int64_t n, int64_t T, int64_t c, int64_t L,
                  const double* B, const double* Z, const int32_t* snps, int64_t n_snps, double min_maf, int prec,
                  double* freq, int64_t* n_obs, int32_t* pass, double* beta, double* se, double* p, double* log10p,
                  double* gamma_out, int64_t* ncal_out, hipStream_t st, rocblas_handle blas, bool device_out) {
    const int64_t k = T + c + L;
    const double one = 1.0, mone = -1.0;
    const double* dX = B + T * n;
    DevBuf XtZ((size_t)(c * k) * 8), XtG((size_t)(c * std::max<int64_t>(1, L)) * 8), XtX((size_t)(c * c) * 8);
    gemm_tn_tall(blas, st, dX, Z, n, c, k, XtZ.as<double>());
    gemm_tn_tall(blas, st, dX, dX, n, c, c, XtX.as<double>());
    if (L > 0) gemm_tn_tall(blas, st, dX, B + (T + c) * n, n, c, L, XtG.as<double>());
    std::vector<double> hXtZ((size_t)(c * k)), hXtX((size_t)(c * c)), hXtG((size_t)(c * std::max<int64_t>(1, L)));
    std::vector<double> gZg((size_t)std::max<int64_t>(1, L)), gg((size_t)std::max<int64_t>(1, L));
    CS_BLAS2(rocblas_set_pointer_mode(blas, rocblas_pointer_mode_host));
    for (int64_t l = 0; l < L; ++l) {
        const double* g = B + (T + c + l) * n;
        CS_BLAS2(rocblas_ddot(blas, (int)n, g, 1, Z + (T + c + l) * n, 1, &gZg[(size_t)l]));
        CS_BLAS2(rocblas_ddot(blas, (int)n, g, 1, g, 1, &gg[(size_t)l]));
    }
    CS_HIP(hipMemcpyAsync(hXtZ.data(), XtZ.p, hXtZ.size() * 8, hipMemcpyDeviceToHost, st));
    CS_HIP(hipMemcpyAsync(hXtX.data(), XtX.p, hXtX.size() * 8, hipMemcpyDeviceToHost, st));
    if (L > 0) CS_HIP(hipMemcpyAsync(hXtG.data(), XtG.p, (size_t)(c * L) * 8, hipMemcpyDeviceToHost, st));
    CS_HIP(hipStreamSynchronize(st));
    // C = (X^T V^-1 X)^-1, (X^T X)^-1
    std::vector<double> Cm((size_t)(c * c)), XtXi = hXtX;
    for (int64_t b2 = 0; b2 < c; ++b2)
        for (int64_t a = 0; a < c; ++a) Cm[(size_t)(b2 * c + a)] = hXtZ[(size_t)((T + b2) * c + a)];
    if (!spd_inverse(Cm, c) || !spd_inverse(XtXi, c))
        throw RuntimeError("projected scan: covariate matrix is not positive definite");
    // P Y = Z_Y - Z_X (C X^T Z_Y)
    std::vector<double> W((size_t)(c * T), 0.0);
    for (int64_t t = 0; t < T; ++t)
        for (int64_t a = 0; a < c; ++a) {
            double v = 0.0;
            for (int64_t b2 = 0; b2 < c; ++b2) v += Cm[(size_t)(b2 * c + a)] * hXtZ[(size_t)(t * c + b2)];
            W[(size_t)(t * c + a)] = v;
        }
    DevBuf dW((size_t)(c * T) * 8), R((size_t)(n * (T + c)) * 8);
    CS_HIP(hipMemcpyAsync(dW.p, W.data(), W.size() * 8, hipMemcpyHostToDevice, st));
    CS_HIP(hipMemcpyAsync(R.p, Z, (size_t)(n * T) * 8, hipMemcpyDeviceToDevice, st));
    CS_BLAS2(rocblas_dgemm(blas, rocblas_operation_none, rocblas_operation_none, (int)n, (int)T, (int)c, &mone,
                           Z + T * n, (int)n, dW.as<double>(), (int)c, &one, R.as<double>(), (int)n));
    CS_HIP(hipMemcpyAsync(R.as<double>() + T * n, dX, (size_t)(n * c) * 8, hipMemcpyDeviceToDevice, st));
    // calibration
    double gsum = 0.0;
    int64_t gcnt = 0;
    for (int64_t l = 0; l < L; ++l) {
        const double* q = &hXtG[(size_t)(l * c)];
        const double* qv = &hXtZ[(size_t)((T + c + l) * c)];
        const double gPg = gZg[(size_t)l] - quad_form(Cm, qv, c);
        const double gtg = gg[(size_t)l] - quad_form(XtXi, q, c);
        if (gtg > 1e-8 * gg[(size_t)l] && gtg > 0.0) {
            gsum += gPg / gtg;
            ++gcnt;
        }
    }
    if (gcnt == 0) throw RuntimeError("projected scan: no usable calibration marker");
    const double gamma = gsum / (double)gcnt;
    std::vector<double> covC((size_t)(c * c));
    for (size_t e = 0; e < covC.size(); ++e) covC[e] = gamma * XtXi[e];
    CS_HIP(hipStreamSynchronize(st));
    cs_scan_model m{T, gamma, 0, 0, nullptr, T, T + c, covC.data()};
    ScanArgs a;
    a.keep = (keep && n_keep > 0) ? keep : nullptr;
    a.n = n;
    a.R = R.as<double>();
    a.ldr = n;
    a.nr = T + c;
    a.min_maf = min_maf;
    a.snps = (snps && n_snps > 0) ? snps : nullptr;
    a.n_snps = a.snps ? n_snps : bed->M;
    a.prec = prec;
    a.device_R = true;
    ScanOut o;
    o.freq = freq; o.n_obs = n_obs; o.pass = pass;
    o.beta = beta; o.se = se; o.p = p; o.log10p = log10p;
    o.device = device_out;
    run_scan(bed, a, &m, o);
    if (gamma_out) *gamma_out = gamma;
    if (ncal_out) *ncal_out = gcnt;
}

}  // namespace csamd

using namespace csamd;

extern "C" int cs_pcg_projected_scan(cs_kinship* K, cs_bed* bed, const int32_t* keep, int64_t n_keep,
                                     const double* Y, int64_t ldy, int64_t T, const float* X, int64_t ldx,
                                     int64_t c, double sigma_g, double sigma_e, const int32_t* cal_snps,
                                     int64_t n_cal, const int32_t* snps, int64_t n_snps, double min_maf,
                                     double rel_tol, int max_iter, int prec, double* freq, int64_t* n_obs,
                                     int32_t* pass, double* beta, double* se, double* p, double* log10p,
                                     cs_pscan_info* info) {
    return guarded([&] {
        if (!K || !bed || !Y || (c > 0 && !X)) throw InvalidArgument("cs_pcg_projected_scan: null argument");
        if (T < 1 || c < 1 || c > 32) throw InvalidArgument("cs_pcg_projected_scan: need T >= 1 and 1 <= c <= 32");
        int64_t n = 0;
        if (cs_kinship_rows(K, &n) != CS_OK) throw RuntimeError(cs_last_error());
        const int64_t nk = (keep && n_keep > 0) ? n_keep : bed->N;
        if (nk != n || ldy < n || ldx < n) throw RuntimeError("cs_pcg_projected_scan: dimension mismatch");
        hipStream_t st;
        CS_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        rocblas_handle blas = nullptr;
        try {
            CS_BLAS2(rocblas_create_handle(&blas));
            CS_BLAS2(rocblas_set_stream(blas, st));
            DevBuf B;
            int64_t L = 0;
            pscan_prepare_rhs(bed, keep, n_keep, n, Y, ldy, T, X, ldx, c, cal_snps, n_cal, min_maf, B, L, st);
            const int64_t k = T + c + L;
            DevBuf Z((size_t)(n * k) * 8);
            int iters = 0;
            std::vector<double> res((size_t)k);
            const int rc = cs_pcg_solve(K, sigma_g, sigma_e, B.as<double>(), n, k, Z.as<double>(), n, rel_tol, max_iter,
                                        1, 0, &iters, res.data());
            if (rc != CS_OK) throw RuntimeError(cs_last_error());
            double gamma = 0.0;
            int64_t gcnt = 0;
            pscan_finish(bed, keep, n_keep, n, T, c, L, B.as<double>(), Z.as<double>(), snps, n_snps, min_maf, prec,
                         freq, n_obs, pass, beta, se, p, log10p, &gamma, &gcnt, st, blas);
            if (info) {
                info->gamma = gamma;
                info->n_calibration = gcnt;
                info->pcg_iterations = iters;
                double mx = 0.0;
                for (double r : res) mx = std::max(mx, r);
                info->max_rel_residual = mx;
            }
        } catch (...) {
            if (blas) (void)rocblas_destroy_handle(blas);
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
            throw;
        }
        (void)rocblas_destroy_handle(blas);
        CS_HIP(hipStreamDestroy(st));
    });
}
