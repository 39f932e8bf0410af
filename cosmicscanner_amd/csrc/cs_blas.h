// FP64 products with a long inner dimension on rocBLAS.
#pragma once
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstdint>

namespace csamd {

// C (s1 x s2, ldc = s1) = A^T B for tall A (n x s1) and B (n x s2), both column-major with leading
// dimension n.  rocBLAS tiles only the small s1 x s2 output (a handful of workgroups, each over
// the whole n), so n is split into chunks: one strided-batched GEMM of the chunk products, then a
// fixed-order sum of the partials (deterministic).  Synchronises the stream.
void gemm_tn_tall(rocblas_handle blas, hipStream_t st, const double* A, const double* B, int64_t n, int64_t s1,
                  int64_t s2, double* C);

}  // namespace csamd
