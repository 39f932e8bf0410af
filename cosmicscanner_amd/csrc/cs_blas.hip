// FP64 helpers on rocBLAS: products with a long inner dimension (split over it) and the
// covariate projection A^T G A by Householder reflectors.
#include "cs_blas.h"

#include <algorithm>
#include <string>

#include "cs_common.h"
#include "cs_ops.h"

namespace csamd {

namespace {

// C[e] = sum over the nb partial products (fixed order).
__global__ void k_sum_parts(const double* __restrict__ part, int64_t nb, int64_t len, double* __restrict__ C) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= len) return;
    double acc = 0.0;
    for (int64_t b = 0; b < nb; ++b) acc += part[b * len + e];
    C[e] = acc;
}

void check_blas(rocblas_status s, const char* what) {
    if (s != rocblas_status_success)
        throw RuntimeError(std::string("rocBLAS error ") + rocblas_status_to_string(s) + " in " + what);
}

}  // namespace

void gemm_tn_tall(rocblas_handle blas, hipStream_t st, const double* A, const double* B, int64_t n, int64_t s1,
                  int64_t s2, double* C) {
    if (n <= 0 || s1 <= 0 || s2 <= 0) return;
    const double one = 1.0, zero = 0.0;
    const int64_t chunk = std::max<int64_t>(1024, (n + 255) / 256);
    const int64_t full = n / chunk, nb = (n + chunk - 1) / chunk;
    DevBuf part((size_t)(nb * s1 * s2) * 8);
    if (full > 0)
        check_blas(rocblas_dgemm_strided_batched(blas, rocblas_operation_transpose, rocblas_operation_none, (int)s1,
                                                 (int)s2, (int)chunk, &one, A, (int)n, chunk, B, (int)n, chunk, &zero,
                                                 part.as<double>(), (int)s1, s1 * s2, (int)full),
                   "gemm_tn_tall (chunks)");
    if (nb > full)
        check_blas(rocblas_dgemm(blas, rocblas_operation_transpose, rocblas_operation_none, (int)s1, (int)s2,
                                 (int)(n - full * chunk), &one, A + full * chunk, (int)n, B + full * chunk, (int)n,
                                 &zero, part.as<double>() + full * s1 * s2, (int)s1),
                   "gemm_tn_tall (tail)");
    k_sum_parts<<<grid_for(s1 * s2), 256, 0, st>>>(part.as<double>(), nb, s1 * s2, C);
    CS_HIP(hipGetLastError());
    CS_HIP(hipStreamSynchronize(st));   // part is freed on return
}

void householder_project_sym(rocblas_handle blas, hipStream_t st, const double* V, const double* tau_h, int64_t n,
                             int64_t c, double* G) {
    if (c <= 0) return;
    const double one = 1.0, zero = 0.0, mone = -1.0;
    DevBuf v((size_t)n * 8), w((size_t)n * 8);
    check_blas(rocblas_set_pointer_mode(blas, rocblas_pointer_mode_host), "householder_project_sym");
    for (int64_t i = 0; i < c; ++i) {
        const int64_t m = n - i;
        const double tau = tau_h[i];
        if (tau == 0.0) continue;   // H_i = I
        // v = (1, V[i+1:, i])
        if (m > 1)
            CS_HIP(hipMemcpyAsync(v.as<double>() + 1, V + i * n + i + 1, (size_t)(m - 1) * 8, hipMemcpyDeviceToDevice,
                                  st));
        CS_HIP(hipMemcpyAsync(v.p, &one, 8, hipMemcpyHostToDevice, st));
        double* Gi = G + i * n + i;
        check_blas(rocblas_dsymv_64(blas, rocblas_fill_lower, m, &tau, Gi, n, v.as<double>(), 1, &zero,
                                    w.as<double>(), 1),
                   "householder_project_sym (symv)");
        double wv = 0.0;
        check_blas(rocblas_ddot_64(blas, m, w.as<double>(), 1, v.as<double>(), 1, &wv), "householder_project_sym (dot)");
        const double alpha = -0.5 * tau * wv;
        check_blas(rocblas_daxpy_64(blas, m, &alpha, v.as<double>(), 1, w.as<double>(), 1),
                   "householder_project_sym (axpy)");
        check_blas(rocblas_dsyr2_64(blas, rocblas_fill_lower, m, &mone, v.as<double>(), 1, w.as<double>(), 1, Gi, n),
                   "householder_project_sym (syr2)");
    }
    CS_HIP(hipStreamSynchronize(st));   // v, w are freed on return
}

}  // namespace csamd
