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
//      tau = tr(V^-1)/n (lmm_engine.cpp:241-708); this build calibrates per marker instead;
//   5. exact stage: the calibrated variance is an estimate whose error grows with the statistic, so
//      every SNP whose largest -log10 P over the traits reaches a threshold is recomputed with its
//      exact GLS variance g^T P g = g^T V^-1 g - (X^T V^-1 g)^T C (X^T V^-1 g): a second multi-RHS
//      PCG solve V Z_g = G_sel over the selected prepared markers, and score_t = g^T P y_t from an
//      FP64 product with P Y.  Below the threshold the calibrated statistics stand, with an error
//      bound derived from the spread of the calibration markers' ratios.
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include "cs_blas.h"
#include "cs_common.h"
#include "cs_ops.h"
#include "cs_stats.h"

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
    // B = [Y, X, G_cal]: Y's columns straight from the caller's array, X widened to doubles
    std::vector<double> hx((size_t)(n * c));
    for (int64_t a = 0; a < c; ++a)
        for (int64_t i = 0; i < n; ++i) hx[(size_t)(a * n + i)] = (double)X[a * ldx + i];
    B.alloc((size_t)(n * k) * 8);
    if (ldy == n) {
        CS_HIP(hipMemcpyAsync(B.p, Y, (size_t)(n * T) * 8, hipMemcpyHostToDevice, st));
    } else {
        CS_HIP(hipMemcpy2DAsync(B.p, (size_t)n * 8, Y, (size_t)ldy * 8, (size_t)n * 8, (size_t)T,
                                hipMemcpyHostToDevice, st));
    }
    if (c > 0) CS_HIP(hipMemcpyAsync(B.as<double>() + T * n, hx.data(), hx.size() * 8, hipMemcpyHostToDevice, st));
    for (int64_t l = 0; l < L; ++l)
        CS_HIP(hipMemcpyAsync(B.as<double>() + (T + c + l) * n, Gc.as<double>() + cidx[(size_t)l] * n,
                              (size_t)n * 8, hipMemcpyDeviceToDevice, st));
    CS_HIP(hipStreamSynchronize(st));
}

// Projection and calibration from the solved Z = V^-1 B (both n x k on the device, k = T + c + L).
void pscan_project(int64_t n, int64_t T, int64_t c, int64_t L, const double* B, const double* Z, PscanProj& pj,
                   hipStream_t st, rocblas_handle blas) {
    const int64_t k = T + c + L;
    const double one = 1.0, mone = -1.0;
    const double* dX = B + T * n;
    pj.n = n;
    pj.T = T;
    pj.c = c;
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
    pj.C.assign((size_t)(c * c), 0.0);
    pj.XtXi = hXtX;
    for (int64_t b2 = 0; b2 < c; ++b2)
        for (int64_t a = 0; a < c; ++a) pj.C[(size_t)(b2 * c + a)] = hXtZ[(size_t)((T + b2) * c + a)];
    if (!spd_inverse(pj.C, c) || !spd_inverse(pj.XtXi, c))
        throw RuntimeError("projected scan: covariate matrix is not positive definite");
    // P Y = Z_Y - Z_X (C X^T Z_Y); R = [P Y, X]
    std::vector<double> W((size_t)(c * T), 0.0);
    for (int64_t t = 0; t < T; ++t)
        for (int64_t a = 0; a < c; ++a) {
            double v = 0.0;
            for (int64_t b2 = 0; b2 < c; ++b2) v += pj.C[(size_t)(b2 * c + a)] * hXtZ[(size_t)(t * c + b2)];
            W[(size_t)(t * c + a)] = v;
        }
    DevBuf dW((size_t)(c * T) * 8);
    pj.R.alloc((size_t)(n * (T + c)) * 8);
    CS_HIP(hipMemcpyAsync(dW.p, W.data(), W.size() * 8, hipMemcpyHostToDevice, st));
    CS_HIP(hipMemcpyAsync(pj.R.p, Z, (size_t)(n * T) * 8, hipMemcpyDeviceToDevice, st));
    CS_BLAS2(rocblas_dgemm(blas, rocblas_operation_none, rocblas_operation_none, (int)n, (int)T, (int)c, &mone,
                           Z + T * n, (int)n, dW.as<double>(), (int)c, &one, pj.R.as<double>(), (int)n));
    CS_HIP(hipMemcpyAsync(pj.R.as<double>() + T * n, dX, (size_t)(n * c) * 8, hipMemcpyDeviceToDevice, st));
    pj.ZX.alloc((size_t)(n * c) * 8);
    CS_HIP(hipMemcpyAsync(pj.ZX.p, Z + T * n, (size_t)(n * c) * 8, hipMemcpyDeviceToDevice, st));
    // calibration: gamma = mean of g^T P g / g~^T g~ over the usable markers, and the relative
    // standard deviation of those ratios (how far one marker's ratio strays from gamma)
    std::vector<double> ratios;
    for (int64_t l = 0; l < L; ++l) {
        const double* q = &hXtG[(size_t)(l * c)];
        const double* qv = &hXtZ[(size_t)((T + c + l) * c)];
        const double gPg = gZg[(size_t)l] - quad_form(pj.C, qv, c);
        const double gtg = gg[(size_t)l] - quad_form(pj.XtXi, q, c);
        if (gtg > 1e-8 * gg[(size_t)l] && gtg > 0.0) ratios.push_back(gPg / gtg);
    }
    if (ratios.empty()) throw RuntimeError("projected scan: no usable calibration marker");
    double gsum = 0.0;
    for (double r : ratios) gsum += r;
    pj.gamma = gsum / (double)ratios.size();
    pj.ncal = (int64_t)ratios.size();
    pj.rel_sd = std::numeric_limits<double>::quiet_NaN();
    if (ratios.size() >= 2) {
        double ss = 0.0;
        for (double r : ratios) ss += (r - pj.gamma) * (r - pj.gamma);
        pj.rel_sd = std::sqrt(ss / (double)(ratios.size() - 1)) / pj.gamma;
    }
    CS_HIP(hipStreamSynchronize(st));
}

// Fused scan of the calibrated statistics: score_t = g^T P y_t, var = gamma g~^T g~.
void pscan_scan(cs_bed* bed, const int32_t* keep, const PscanProj& pj, const int32_t* snps, int64_t n_snps,
                double min_maf, int prec, double* freq, int64_t* n_obs, int32_t* pass, double* beta, double* se,
                double* p, double* log10p, bool device_out) {
    const int64_t T = pj.T, c = pj.c;
    std::vector<double> covC((size_t)(c * c));
    for (size_t e = 0; e < covC.size(); ++e) covC[e] = pj.gamma * pj.XtXi[e];
    cs_scan_model m{T, pj.gamma, 0, 0, nullptr, T, T + c, covC.data()};
    ScanArgs a;
    a.keep = keep;
    a.n = pj.n;
    a.R = pj.R.as<double>();
    a.ldr = pj.n;
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
}

double pscan_threshold(double exact_log10p, int64_t T) {
    if (std::isnan(exact_log10p)) throw InvalidArgument("projected scan: exact_log10p is NaN");
    if (std::isinf(exact_log10p) && exact_log10p > 0.0) return exact_log10p;
    if (exact_log10p < 0.0) return -1.0;
    const double level = exact_log10p == 0.0 ? kPscanExactDefault : exact_log10p;
    return level + std::log10((double)std::max<int64_t>(1, T));
}

// flags[i] = 1 for rows that pass and whose largest -log10 P over the traits reaches thr
// (thr < 0: every tested row)
__global__ void k_pscan_flags(const double* __restrict__ log10p, const int32_t* __restrict__ pass, int64_t ns,
                              int64_t T, double thr, int32_t* __restrict__ flags) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= ns) return;
    int f = 0;
    if (pass[i]) {
        const double* r = log10p + i * T;
        if (thr < 0.0) {
            f = isnan(r[0]) ? 0 : 1;
        } else {
            for (int64_t t = 0; t < T; ++t) f |= (r[t] >= thr) ? 1 : 0;
        }
    }
    flags[i] = f;
}

std::vector<int64_t> pscan_select(const double* log10p, const int32_t* pass, int64_t ns, int64_t T, double thr,
                                  bool device, hipStream_t st) {
    std::vector<int64_t> rows;
    if (std::isinf(thr) || ns == 0) return rows;
    std::vector<int32_t> flags((size_t)ns, 0);
    if (device) {
        DevBuf df((size_t)ns * 4);
        k_pscan_flags<<<grid_for(ns), 256, 0, st>>>(log10p, pass, ns, T, thr, df.as<int32_t>());
        CS_HIP(hipGetLastError());
        CS_HIP(hipMemcpyAsync(flags.data(), df.p, (size_t)ns * 4, hipMemcpyDeviceToHost, st));
        CS_HIP(hipStreamSynchronize(st));
    } else {
        for (int64_t i = 0; i < ns; ++i) {
            if (!pass[i]) continue;
            const double* r = log10p + i * T;
            if (thr < 0.0) {
                flags[(size_t)i] = std::isnan(r[0]) ? 0 : 1;
            } else {
                for (int64_t t = 0; t < T; ++t) flags[(size_t)i] |= (r[t] >= thr) ? 1 : 0;
            }
        }
    }
    for (int64_t i = 0; i < ns; ++i)
        if (flags[(size_t)i]) rows.push_back(i);
    return rows;
}

// Statistics of the exact variances: scores S x T (column-major), var[S]; row s goes to output
// row rows[s] (rows NULL: row s) of the row-major ns x T outputs.  Same formulas and "not
// tested" rule as the fused scan epilogue.
__global__ void k_pscan_exact_stats(const double* __restrict__ scores, const double* __restrict__ var, int64_t S,
                                    int64_t T, const int64_t* __restrict__ rows, double* __restrict__ beta,
                                    double* __restrict__ se, double* __restrict__ p, double* __restrict__ log10p) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= S * T) return;
    const int64_t s = e % S, t = e / S;
    const int64_t o = (rows ? rows[s] : s) * T + t;
    const double v = var[s];
    if (!(v > 1e-12) || !isfinite(v)) {
        beta[o] = se[o] = p[o] = log10p[o] = __builtin_nan("");
        return;
    }
    const double score = scores[t * S + s];
    const double sd = sqrt(v);
    const double stat = score / sd;
    const double pv = d_chisq_prob(stat * stat, 1);
    beta[o] = score / v;
    se[o] = 1.0 / sd;
    p[o] = pv;
    log10p[o] = pv > 0.0 ? -log10(pv) : 300.0;
}

// The per-column dots of the exact stage, one workgroup per selected marker s and a fixed
// reduction order, so a column's values do not depend on how many columns share the launch:
// scores[t S + s] = g.PY_t, q[s c + a] = g.ZX_a, gz[s] = g.z, za[s] = z.aux.
__global__ __launch_bounds__(256) void k_pscan_dots(const double* __restrict__ G, const double* __restrict__ Zg,
                                                   const double* __restrict__ Aux, const double* __restrict__ PY,
                                                   const double* __restrict__ ZX, int64_t n, int64_t S, int T, int c,
                                                   double* __restrict__ scores, double* __restrict__ q,
                                                   double* __restrict__ gz, double* __restrict__ za) {
    constexpr int NB = 8;
    __shared__ double sh[NB][256];
    const int64_t s = blockIdx.x;
    const double* g = G + s * n;
    const double* z = Zg + s * n;
    const double* x = Aux + s * n;
    const int nv = T + c + 2;
    for (int v0 = 0; v0 < nv; v0 += NB) {
        double acc[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j) acc[j] = 0.0;
        for (int64_t r = threadIdx.x; r < n; r += 256) {
            const double gr = g[r], zr = z[r];
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                const int v = v0 + j;
                if (v < T) acc[j] += gr * PY[(int64_t)v * n + r];
                else if (v < T + c) acc[j] += gr * ZX[(int64_t)(v - T) * n + r];
                else if (v == T + c) acc[j] += gr * zr;
                else if (v == T + c + 1) acc[j] += zr * x[r];
            }
        }
#pragma unroll
        for (int j = 0; j < NB; ++j) sh[j][threadIdx.x] = acc[j];
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) {
#pragma unroll
                for (int j = 0; j < NB; ++j) sh[j][threadIdx.x] += sh[j][threadIdx.x + w];
            }
            __syncthreads();
        }
        if (threadIdx.x < NB) {
            const int v = v0 + (int)threadIdx.x;
            const double val = sh[threadIdx.x][0];
            if (v < T) scores[(int64_t)v * S + s] = val;
            else if (v < T + c) q[s * c + (v - T)] = val;
            else if (v == T + c) gz[s] = val;
            else if (v == T + c + 1) za[s] = val;
        }
        __syncthreads();
    }
}

void pscan_exact(const PscanProj& pj, int64_t S, const int64_t* rows, const double* dG, const double* dZg,
                 const double* dAux, bool aux_is_residual, double* beta, double* se, double* p, double* log10p,
                 bool device_out, hipStream_t st, rocblas_handle blas) {
    if (S <= 0) return;
    const int64_t n = pj.n, T = pj.T, c = pj.c;
    DevBuf scores((size_t)(S * T) * 8), XtZg((size_t)(c * S) * 8), gz((size_t)S * 8), za((size_t)S * 8),
        dvar((size_t)S * 8);
    // score_t = g^T P y_t; q = (V^-1 X)^T g = X^T V^-1 g; g^T z; z^T aux
    k_pscan_dots<<<(unsigned)S, 256, 0, st>>>(dG, dZg, dAux, pj.R.as<double>(), pj.ZX.as<double>(), n, S, (int)T,
                                              (int)c, scores.as<double>(), XtZg.as<double>(), gz.as<double>(),
                                              za.as<double>());
    CS_HIP(hipGetLastError());
    std::vector<double> hq((size_t)(c * S)), hg((size_t)S), ha((size_t)S), var((size_t)S);
    CS_HIP(hipMemcpyAsync(hq.data(), XtZg.p, hq.size() * 8, hipMemcpyDeviceToHost, st));
    CS_HIP(hipMemcpyAsync(hg.data(), gz.p, hg.size() * 8, hipMemcpyDeviceToHost, st));
    CS_HIP(hipMemcpyAsync(ha.data(), za.p, ha.size() * 8, hipMemcpyDeviceToHost, st));
    CS_HIP(hipStreamSynchronize(st));
    // g^T P g = g^T V^-1 g - q^T (X^T V^-1 X)^-1 q, g^T V^-1 g = g^T z + z^T r = 2 g^T z - z^T V z
    for (int64_t s = 0; s < S; ++s) {
        const double gvg = aux_is_residual ? hg[(size_t)s] + ha[(size_t)s] : 2.0 * hg[(size_t)s] - ha[(size_t)s];
        var[(size_t)s] = gvg - quad_form(pj.C, &hq[(size_t)(s * c)], c);
    }
    CS_HIP(hipMemcpyAsync(dvar.p, var.data(), (size_t)S * 8, hipMemcpyHostToDevice, st));
    if (device_out) {
        DevBuf drows((size_t)S * 8);
        CS_HIP(hipMemcpyAsync(drows.p, rows, (size_t)S * 8, hipMemcpyHostToDevice, st));
        k_pscan_exact_stats<<<grid_for(S * T), 256, 0, st>>>(scores.as<double>(), dvar.as<double>(), S, T,
                                                             drows.as<int64_t>(), beta, se, p, log10p);
        CS_HIP(hipGetLastError());
        CS_HIP(hipStreamSynchronize(st));
        return;
    }
    DevBuf ob((size_t)(S * T) * 8), os((size_t)(S * T) * 8), op((size_t)(S * T) * 8), ol((size_t)(S * T) * 8);
    k_pscan_exact_stats<<<grid_for(S * T), 256, 0, st>>>(scores.as<double>(), dvar.as<double>(), S, T, nullptr,
                                                         ob.as<double>(), os.as<double>(), op.as<double>(),
                                                         ol.as<double>());
    CS_HIP(hipGetLastError());
    std::vector<double> hb((size_t)(S * T)), hs(hb.size()), hp(hb.size()), hl(hb.size());
    CS_HIP(hipMemcpyAsync(hb.data(), ob.p, hb.size() * 8, hipMemcpyDeviceToHost, st));
    CS_HIP(hipMemcpyAsync(hs.data(), os.p, hs.size() * 8, hipMemcpyDeviceToHost, st));
    CS_HIP(hipMemcpyAsync(hp.data(), op.p, hp.size() * 8, hipMemcpyDeviceToHost, st));
    CS_HIP(hipMemcpyAsync(hl.data(), ol.p, hl.size() * 8, hipMemcpyDeviceToHost, st));
    CS_HIP(hipStreamSynchronize(st));
    for (int64_t s = 0; s < S; ++s)
        for (int64_t t = 0; t < T; ++t) {
            const size_t src = (size_t)(s * T + t), dst = (size_t)(rows[s] * T + t);
            beta[dst] = hb[src];
            se[dst] = hs[src];
            p[dst] = hp[src];
            log10p[dst] = hl[src];
        }
}

// Bound on |d(-log10 P)| of the calibrated statistics of SNPs below the threshold: a relative
// error e of the calibrated variance scales the chi^2_1 statistic by 1 / (1 + e), and the change
// in -log10 P grows with the statistic, so its largest value below the threshold is the change at
// the statistic s_thr whose -log10 P equals the threshold, for e = 4 rel_sd (rel_sd: the spread
// of the calibration markers' ratios).  0 when every tested SNP is exact; NaN when the stage is
// off or fewer than two markers calibrate.
double pscan_bound_below(double thr, double rel_sd) {
    if (thr < 0.0) return 0.0;
    if (std::isinf(thr) || !(rel_sd >= 0.0)) return std::numeric_limits<double>::quiet_NaN();
    auto nl = [](double s) {
        const double pv = d_chisq_prob(s, 1);
        return pv > 0.0 ? -std::log10(pv) : 300.0;
    };
    double lo = 0.0, hi = 4000.0;
    for (int i = 0; i < 200 && hi - lo > 1e-12 * hi; ++i) {
        const double mid = 0.5 * (lo + hi);
        (nl(mid) < thr ? lo : hi) = mid;
    }
    const double e = 4.0 * rel_sd;
    return std::max(nl(hi * (1.0 + e)) - thr, thr - nl(hi / (1.0 + e)));
}

// AZ = sg KZ + se Z over n x k (in place in KZ)
__global__ void k_pscan_vz(double* __restrict__ KZ, const double* __restrict__ Z, int64_t nk, double sg, double se) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nk; e += (int64_t)gridDim.x * blockDim.x)
        KZ[e] = sg * KZ[e] + se * Z[e];
}

// Exact stage of the standalone projected scan: prepared markers of the selected rows, one
// multi-RHS PCG solve per block of columns (tolerance pscan_exact_tol), one more apply for V Z,
// exact statistics written over the calibrated ones.
static void pscan_exact_solve(cs_kinship* K, cs_bed* bed, const int32_t* keep, const PscanProj& pj,
                              const int32_t* snps, const std::vector<int64_t>& rows, double sigma_g, double sigma_e,
                              double min_maf, double rel_tol, int max_iter, double* beta, double* se, double* p,
                              double* log10p, int* iters_out, hipStream_t st, rocblas_handle blas) {
    const int64_t n = pj.n, S = (int64_t)rows.size();
    // columns per solve: the PCG holds about eight n x block FP64 arrays; keep them near 4 GiB
    const int64_t block = std::max<int64_t>(16, std::min<int64_t>(S, ((int64_t)4 << 30) / (64 * std::max<int64_t>(1, n))));
    const double tol = pscan_exact_tol(rel_tol);
    int it_max = 0;
    for (int64_t s0 = 0; s0 < S; s0 += block) {
        const int64_t sb = std::min(block, S - s0);
        std::vector<int32_t> sel((size_t)sb);
        for (int64_t s = 0; s < sb; ++s) {
            const int64_t r = rows[(size_t)(s0 + s)];
            sel[(size_t)s] = snps ? snps[r] : (int32_t)r;
        }
        DevBuf dG((size_t)(n * sb) * 8), dZ((size_t)(n * sb) * 8), dVZ((size_t)(n * sb) * 8);
        markers_device(bed, keep, n, sel.data(), sb, min_maf, dG.as<double>(), n, nullptr, nullptr, st);
        int it = 0;
        if (cs_pcg_solve(K, sigma_g, sigma_e, dG.as<double>(), n, sb, dZ.as<double>(), n, tol, max_iter, 1, 0, &it,
                         nullptr) != CS_OK)
            throw RuntimeError(cs_last_error());
        it_max = std::max(it_max, it);
        if (cs_kinship_apply_device(K, dZ.as<double>(), n, sb, dVZ.as<double>(), n) != CS_OK)
            throw RuntimeError(cs_last_error());
        k_pscan_vz<<<grid_for(n * sb), 256, 0, st>>>(dVZ.as<double>(), dZ.as<double>(), n * sb, sigma_g, sigma_e);
        CS_HIP(hipGetLastError());
        pscan_exact(pj, sb, rows.data() + s0, dG.as<double>(), dZ.as<double>(), dVZ.as<double>(), false, beta, se, p,
                    log10p, false, st, blas);
    }
    if (iters_out) *iters_out = it_max;
}

}  // namespace csamd

using namespace csamd;

extern "C" int cs_pcg_projected_scan(cs_kinship* K, cs_bed* bed, const int32_t* keep, int64_t n_keep,
                                     const double* Y, int64_t ldy, int64_t T, const float* X, int64_t ldx,
                                     int64_t c, double sigma_g, double sigma_e, const int32_t* cal_snps,
                                     int64_t n_cal, const int32_t* snps, int64_t n_snps, double min_maf,
                                     double rel_tol, int max_iter, int prec, double exact_log10p, double* freq,
                                     int64_t* n_obs, int32_t* pass, double* beta, double* se, double* p,
                                     double* log10p, cs_pscan_info* info) {
    return guarded([&] {
        if (!K || !bed || !Y || (c > 0 && !X)) throw InvalidArgument("cs_pcg_projected_scan: null argument");
        if (!freq || !n_obs || !pass || !beta || !se || !p || !log10p)
            throw InvalidArgument("cs_pcg_projected_scan: null output array");
        if (T < 1 || c < 1 || c > 32) throw InvalidArgument("cs_pcg_projected_scan: need T >= 1 and 1 <= c <= 32");
        const double thr = pscan_threshold(exact_log10p, T);
        int64_t n = 0;
        if (cs_kinship_rows(K, &n) != CS_OK) throw RuntimeError(cs_last_error());
        const int64_t nk = (keep && n_keep > 0) ? n_keep : bed->N;
        if (nk != n || ldy < n || ldx < n) throw RuntimeError("cs_pcg_projected_scan: dimension mismatch");
        const int32_t* kp = (keep && n_keep > 0) ? keep : nullptr;
        const int32_t* sp = (snps && n_snps > 0) ? snps : nullptr;
        const int64_t ns = sp ? n_snps : bed->M;
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
            PscanProj pj;
            pscan_project(n, T, c, L, B.as<double>(), Z.as<double>(), pj, st, blas);
            B.release();
            Z.release();
            pscan_scan(bed, kp, pj, sp, ns, min_maf, prec, freq, n_obs, pass, beta, se, p, log10p, false);
            const std::vector<int64_t> rows = pscan_select(log10p, pass, ns, T, thr, false, st);
            int it_exact = 0;
            if (!rows.empty())
                pscan_exact_solve(K, bed, kp, pj, sp, rows, sigma_g, sigma_e, min_maf, rel_tol, max_iter, beta, se, p,
                                  log10p, &it_exact, st, blas);
            if (info) {
                info->gamma = pj.gamma;
                info->n_calibration = pj.ncal;
                info->pcg_iterations = iters;
                double mx = 0.0;
                for (double r : res) mx = std::max(mx, r);
                info->max_rel_residual = mx;
                info->n_exact = (int64_t)rows.size();
                info->exact_pcg_iterations = it_exact;
                info->calibration_rel_sd = pj.rel_sd;
                info->exact_log10p = thr;
                info->dlog10p_bound_below = pscan_bound_below(thr, pj.rel_sd);
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
