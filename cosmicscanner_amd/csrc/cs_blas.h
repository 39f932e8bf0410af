// FP64 helpers on rocBLAS: products with a long inner dimension, covariate projection.
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

// A^T G A for A = Q[:, c:], Q = H_1 ... H_c the Householder reflectors of an n x c matrix (geqrf
// output V, ld n, scalars tau on the host): the trailing (n - c) block of Q^T G Q, formed in place
// in the lower triangle of G (n x n, ld n).  The block [i:, i:] after H_i depends only on the
// block [i:, i:] before it, so each reflector updates the trailing block alone (LAPACK's
// two-sided update w = tau G v, w -= (tau / 2)(w . v) v, G -= v w^T + w v^T): O(c n^2)
// memory-bound work instead of the O(n^3) products with an explicit Q.  Leaves the handle in
// host pointer mode; synchronises the stream.
void householder_project_sym(rocblas_handle blas, hipStream_t st, const double* V, const double* tau_h, int64_t n,
                             int64_t c, double* G);

}  // namespace csamd
