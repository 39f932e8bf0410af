// FP64 products with a long inner dimension (split over the inner dimension).
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

}  // namespace csamd
