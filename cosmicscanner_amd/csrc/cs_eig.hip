// Dense symmetric eigendecomposition for matrices beyond rocSOLVER's 32-bit interfaces.
//
// rocSOLVER's dsyevd / dorgqr / dormqr take rocblas_int sizes and index their n x n work arrays in
// 32 bits, which fails beyond about n = 32766 (EigenVce::prepare and MvEigenVce::prepare need the
// eigendecomposition of A^T G A, eigen_vce.cpp:19-63, 496-549).  This file computes it with 64-bit
// clean pieces:
//   1. blocked Householder tridiagonal reduction (the panel scheme of LAPACK's dsytrd / dlatrd,
//      lower storage) on rocBLAS *_64 level-2/3 routines and rocSOLVER dlarfg_64;
//   2. one divide-and-conquer level on the tridiagonal T (Cuppen's split T = diag(T1, T2) +
//      rho v v^T): rocSOLVER dstedc on the two halves (each within its 32-bit limits), then the
//      rank-one merge here -- deflation of small z components and of close eigenvalue pairs by
//      Givens rotations, the secular equation solved per root by bisection relative to the nearer
//      pole, eigenvectors from the Loewner-recomputed z (Gu and Eisenstat, SIAM J. Matrix Anal.
//      Appl. 16:172-191, 1995), and one FP64 GEMM with the halves' eigenvectors;
//   3. back-transformation by the reflectors in compact WY blocks (three GEMMs per block).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <numeric>
#include <vector>

#include "cs_common.h"
#include "cs_ops.h"

namespace csamd {

namespace {

void eig_chk(rocblas_status s, const char* what) {
    if (s != rocblas_status_success)
        throw RuntimeError(std::string("rocBLAS/rocSOLVER error ") + rocblas_status_to_string(s) + " in " + what);
}

__global__ void k_set_scalar(double* p, double v) { *p = v; }
__global__ void k_copy_scalar(const double* s, double* d) { *d = *s; }
__global__ void k_half_tau_dot(const double* tau, const double* dot, double* out) { *out = -0.5 * tau[0] * dot[0]; }

// diag[j] = A(j, j)
__global__ void k_get_diag(const double* __restrict__ A, int64_t lda, int64_t n, double* __restrict__ diag) {
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j < n) diag[j] = A[j + j * lda];
}

// A(j + 1, j) = off[j] for j in [j0, j1)
__global__ void k_put_subdiag(double* __restrict__ A, int64_t lda, const double* __restrict__ off, int64_t j0,
                              int64_t j1) {
    const int64_t j = j0 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j < j1) A[(j + 1) + j * lda] = off[j];
}

// y = A v for the symmetric m x m matrix A held in its lower triangle (column-major, ld lda), in one
// pass over the lower triangle.  Column strip bj (kSyW = 32 columns) is its 32 x 32 diagonal block
// plus the rows below it in tiles of 64 rows; work item (bj, ch) of a workgroup (4 waves) takes the
// diagonal block (ch = 0) and up to kSyChunk tiles.  A wave holds one tile row per lane: per column
// it adds A(r, c) v(c) to the lane's row sum -- complete over the strip, stored to P[bj][row] --
// and A(r, c) v(r) to a per-lane column accumulator, reduced over lanes and waves once per item
// into Tp[bj][ch][col].  k_symv_sum adds the partial sums of each row in a fixed order, so the
// result is deterministic.
constexpr int kSyW = 32;
constexpr int kSyR = 64;
constexpr int kSyChunk = 16;

// number of row tiles below the diagonal block of strip bj, and of work items holding them
__host__ __device__ inline int64_t symv_tiles(int64_t m, int64_t bj) {
    const int64_t r0 = (bj + 1) * kSyW;
    return r0 < m ? (m - r0 + kSyR - 1) / kSyR : 0;
}

__global__ __launch_bounds__(256) void k_symv_lower(const double* __restrict__ A, int64_t lda, int64_t m,
                                                    const double* __restrict__ v, int64_t nch,
                                                    double* __restrict__ P, double* __restrict__ Tp) {
    __shared__ double red[4][kSyW];
    __shared__ double vjs[kSyW];
    const int64_t bj = blockIdx.y, ch = blockIdx.x;
    const int64_t ntile = symv_tiles(m, bj);
    const int64_t nit = (ntile + kSyChunk - 1) / kSyChunk;
    if (ch >= (nit > 0 ? nit : 1)) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t c0 = bj * kSyW;
    if (threadIdx.x < kSyW) vjs[threadIdx.x] = c0 + threadIdx.x < m ? v[c0 + threadIdx.x] : 0.0;
    __syncthreads();
    if (ch == 0 && wave == 0 && lane < kSyW) {
        // diagonal block, both triangles from the lower one
        const int64_t gr = c0 + lane;
        if (gr < m) {
            double y = 0.0;
            for (int c = 0; c < kSyW; ++c) {
                const int64_t gc = c0 + c;
                if (gc >= m) break;
                const double a = c <= lane ? A[gr + gc * lda] : A[gc + gr * lda];
                y = fma(a, vjs[c], y);
            }
            P[bj * m + gr] = y;
        }
    }
    if (nit == 0) return;
    double acc[kSyW];
#pragma unroll
    for (int c = 0; c < kSyW; ++c) acc[c] = 0.0;
    const int64_t t0 = ch * kSyChunk;
    const int64_t t1 = t0 + kSyChunk < ntile ? t0 + kSyChunk : ntile;
    for (int64_t t = t0 + wave; t < t1; t += 4) {
        const int64_t gr = c0 + kSyW + t * kSyR + lane;
        const bool ok = gr < m;
        const double vr = ok ? v[gr] : 0.0;
        // a lane past the end reads row c0 + kSyW of the strip (in range) and drops it
        const double* col = A + (ok ? gr : c0 + kSyW) + c0 * lda;
        double y = 0.0;
#pragma unroll
        for (int cg = 0; cg < kSyW; cg += 16) {
            double a[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) a[u] = col[(cg + u) * lda];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                y = fma(a[u], vjs[cg + u], y);
                acc[cg + u] = fma(a[u], vr, acc[cg + u]);
            }
        }
        if (ok) P[bj * m + gr] = y;
    }
    // transpose-reduce the 32 column accumulators over the 64 lanes: lanes l and l + 32 end with
    // column l (the last step adds the two halves)
#pragma unroll
    for (int o = 16, cnt = kSyW; o >= 1; o >>= 1, cnt >>= 1) {
        const bool hi = (lane & o) != 0;
#pragma unroll
        for (int k = 0; k < cnt / 2; ++k) {
            const double keep = hi ? acc[k + cnt / 2] : acc[k];
            const double send = hi ? acc[k] : acc[k + cnt / 2];
            acc[k] = keep + __shfl_xor(send, o);
        }
    }
    acc[0] += __shfl_xor(acc[0], 32);
    if (lane < kSyW) red[wave][lane] = acc[0];
    __syncthreads();
    if (wave == 0 && lane < kSyW)
        Tp[(bj * nch + ch) * kSyW + lane] = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
}

// One workgroup per 32-row block b: 8 thread groups sum interleaved strips of P for each row (four
// independent accumulators each), then the partials and the column sums Tp combine in a fixed order.
__global__ __launch_bounds__(256) void k_symv_sum(int64_t m, int64_t nch, const double* __restrict__ P,
                                                  const double* __restrict__ Tp, double* __restrict__ y) {
    __shared__ double part[8][kSyW];
    const int rho = threadIdx.x & (kSyW - 1), q = threadIdx.x >> 5;
    const int64_t b = blockIdx.x, r = b * kSyW + rho;
    double s = 0.0;
    if (r < m) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        int64_t bj = q;
        for (; bj + 24 <= b; bj += 32) {
            s0 += P[bj * m + r];
            s1 += P[(bj + 8) * m + r];
            s2 += P[(bj + 16) * m + r];
            s3 += P[(bj + 24) * m + r];
        }
        for (; bj <= b; bj += 8) s0 += P[bj * m + r];
        s = (s0 + s1) + (s2 + s3);
    }
    part[q][rho] = s;
    __syncthreads();
    if (q == 0 && r < m) {
        double t = part[0][rho];
        for (int k = 1; k < 8; ++k) t += part[k][rho];
        const int64_t nit = (symv_tiles(m, b) + kSyChunk - 1) / kSyChunk;
        for (int64_t ch = 0; ch < nit; ++ch) t += Tp[(b * nch + ch) * kSyW + rho];
        y[r] = t;
    }
}

inline int64_t symv_strips(int64_t m) { return (m + kSyW - 1) / kSyW; }
inline int64_t symv_nch(int64_t m) { return std::max<int64_t>(1, (symv_tiles(m, 0) + kSyChunk - 1) / kSyChunk); }

// Workspace of symv_lower for orders up to n.
struct SymvWork {
    DevBuf P, Tp;
    explicit SymvWork(int64_t n) {
        P.alloc((size_t)(symv_strips(n) * n) * 8);
        Tp.alloc((size_t)(symv_strips(n) * symv_nch(n) * kSyW) * 8);
    }
};

void symv_lower(hipStream_t st, int64_t m, const double* A, int64_t lda, const double* v, double* y, SymvWork& ws) {
    if (m <= 0) return;
    const int64_t nb = symv_strips(m), nch = symv_nch(m);
    if ((size_t)(nb * m) * 8 > ws.P.bytes || (size_t)(nb * nch * kSyW) * 8 > ws.Tp.bytes)
        throw RuntimeError("symv_lower: workspace too small");
    k_symv_lower<<<dim3((unsigned)nch, (unsigned)nb), 256, 0, st>>>(A, lda, m, v, nch, ws.P.as<double>(),
                                                                    ws.Tp.as<double>());
    k_symv_sum<<<(unsigned)nb, 256, 0, st>>>(m, nch, ws.P.as<double>(), ws.Tp.as<double>(), y);
    CS_HIP(hipGetLastError());
}

// Blocked tridiagonal reduction of the symmetric n x n matrix A (lower triangle, leading
// dimension lda): A = Q T Q^T with Q = H(0) ... H(n-2), H(j) = I - tau_j v_j v_j^T, v_j zero in
// rows 0..j, 1 in row j + 1, stored in A(j + 2 :, j).  diag (n), off (n - 1), tau (n - 1) are device
// arrays.  All scalars stay on the device (no host synchronisation inside the loop).
void sytrd_lower(rocblas_handle h, hipStream_t st, int64_t n, double* A, int64_t lda, double* diag, double* off,
                 double* tau) {
    constexpr int64_t nb = 64;
    const int64_t ldw = n;
    DevBuf Wb((size_t)(n * nb) * 8), tmpb((size_t)nb * 8), scb(8 * 8);
    SymvWork sw(n);
    double* W = Wb.as<double>();
    double* tmp = tmpb.as<double>();
    double* sc = scb.as<double>();   // [0] -1, [1] 0, [2] 1, [3] dot, [4] alpha
    k_set_scalar<<<1, 1, 0, st>>>(sc + 0, -1.0);
    k_set_scalar<<<1, 1, 0, st>>>(sc + 1, 0.0);
    k_set_scalar<<<1, 1, 0, st>>>(sc + 2, 1.0);
    CS_HIP(hipGetLastError());
    const double *mone = sc, *zero = sc + 1, *one = sc + 2;
    eig_chk(rocblas_set_pointer_mode(h, rocblas_pointer_mode_device), "pointer mode");
    for (int64_t p = 0; p < n - 1; p += nb) {
        const int64_t w = std::min<int64_t>(nb, n - 1 - p);
        for (int64_t i = 0; i < w; ++i) {
            const int64_t j = p + i;
            double* colj = A + j + j * lda;
            if (i > 0) {
                // A(j:, j) -= A(j:, p:j) W(j, 0:i)^T + W(j:, 0:i) A(j, p:j)^T
                eig_chk(rocblas_dgemv_64(h, rocblas_operation_none, n - j, i, mone, A + j + p * lda, lda, W + j, ldw,
                                         one, colj, 1), "sytrd gemv");
                eig_chk(rocblas_dgemv_64(h, rocblas_operation_none, n - j, i, mone, W + j, ldw, A + j + p * lda, lda,
                                         one, colj, 1), "sytrd gemv");
            }
            const int64_t len = n - j - 1;
            double* alpha = A + (j + 1) + j * lda;
            eig_chk(rocsolver_dlarfg_64(h, len, alpha, alpha + 1, 1, tau + j), "sytrd larfg");
            k_copy_scalar<<<1, 1, 0, st>>>(alpha, off + j);
            k_set_scalar<<<1, 1, 0, st>>>(alpha, 1.0);
            CS_HIP(hipGetLastError());
            const double* v = alpha;
            double* wi = W + (j + 1) + i * ldw;
            symv_lower(st, len, A + (j + 1) + (j + 1) * lda, lda, v, wi, sw);
            if (i > 0) {
                eig_chk(rocblas_dgemv_64(h, rocblas_operation_transpose, len, i, one, W + (j + 1), ldw, v, 1, zero,
                                         tmp, 1), "sytrd gemv");
                eig_chk(rocblas_dgemv_64(h, rocblas_operation_none, len, i, mone, A + (j + 1) + p * lda, lda, tmp, 1,
                                         one, wi, 1), "sytrd gemv");
                eig_chk(rocblas_dgemv_64(h, rocblas_operation_transpose, len, i, one, A + (j + 1) + p * lda, lda, v,
                                         1, zero, tmp, 1), "sytrd gemv");
                eig_chk(rocblas_dgemv_64(h, rocblas_operation_none, len, i, mone, W + (j + 1), ldw, tmp, 1, one, wi,
                                         1), "sytrd gemv");
            }
            eig_chk(rocblas_dscal_64(h, len, tau + j, wi, 1), "sytrd scal");
            eig_chk(rocblas_ddot_64(h, len, wi, 1, v, 1, sc + 3), "sytrd dot");
            k_half_tau_dot<<<1, 1, 0, st>>>(tau + j, sc + 3, sc + 4);
            CS_HIP(hipGetLastError());
            eig_chk(rocblas_daxpy_64(h, len, sc + 4, v, 1, wi, 1), "sytrd axpy");
        }
        // trailing update A(t0:, t0:) -= V W^T + W V^T
        const int64_t t0 = p + w, tn = n - t0;
        if (tn > 0)
            eig_chk(rocblas_dsyr2k_64(h, rocblas_fill_lower, rocblas_operation_none, tn, w, mone, A + t0 + p * lda,
                                      lda, W + t0, ldw, one, A + t0 + t0 * lda, lda), "sytrd syr2k");
        k_put_subdiag<<<grid_for(w), 256, 0, st>>>(A, lda, off, p, p + w);
        CS_HIP(hipGetLastError());
    }
    eig_chk(rocblas_set_pointer_mode(h, rocblas_pointer_mode_host), "pointer mode");
    k_get_diag<<<grid_for(n), 256, 0, st>>>(A, lda, n, diag);
    CS_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------------
// Rank-one merge.  Non-deflated poles d[0..K) ascending (distinct), weights z (|z| = 1 over all n),
// rho > 0.  Root k lies in (d[k], d[k+1]) (k < K - 1) or (d[K-1], d[K-1] + rho).  Each root is
// found by bisection on tau = lambda - d[o] for the nearer pole o, with the secular function
// 1 + rho sum_j z_j^2 / ((d_j - d_o) - tau) increasing in tau.
__global__ void k_secular_roots(const double* __restrict__ d, const double* __restrict__ z, int64_t K, double rho,
                                int32_t* __restrict__ org, double* __restrict__ tau) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= K) return;
    auto f = [&](int64_t o, double t) {
        const double dov = d[o];
        double s = 0.0;
        for (int64_t j = 0; j < K; ++j) {
            const double den = (d[j] - dov) - t;
            s += z[j] * (z[j] / den);
        }
        return 1.0 + rho * s;
    };
    int64_t o;
    double lo, hi;   // bracket in tau
    if (k == K - 1) {
        o = k;
        lo = 0.0;
        hi = rho;
    } else {
        const double gap = d[k + 1] - d[k];
        if (f(k, 0.5 * gap) >= 0.0) {
            o = k;
            lo = 0.0;
            hi = 0.5 * gap;
        } else {
            o = k + 1;
            lo = -0.5 * gap;
            hi = 0.0;
        }
    }
    for (int it = 0; it < 400; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        if (f(o, mid) > 0.0) hi = mid;
        else lo = mid;
    }
    org[k] = (int32_t)o;
    tau[k] = 0.5 * (lo + hi);
}

// zhat_i^2 = (lambda_{K-1} - d_i) / rho * prod_{j<i} (lambda_j - d_i) / (d_j - d_i)
//            * prod_{i<=j<K-1} (lambda_j - d_i) / (d_{j+1} - d_i),  lambda_j - d_i = (d_{o_j} - d_i) + tau_j
__global__ void k_zhat(const double* __restrict__ d, const double* __restrict__ z, int64_t K, double rho,
                       const int32_t* __restrict__ org, const double* __restrict__ tau, double* __restrict__ zh) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= K) return;
    const double di = d[i];
    double v = ((d[org[K - 1]] - di) + tau[K - 1]) / rho;
    for (int64_t j = 0; j < i; ++j) v *= ((d[org[j]] - di) + tau[j]) / (d[j] - di);
    for (int64_t j = i; j < K - 1; ++j) v *= ((d[org[j]] - di) + tau[j]) / (d[j + 1] - di);
    const double a = sqrt(fabs(v));
    zh[i] = z[i] >= 0.0 ? a : -a;
}

// Column col[k] of Y (ld n): the eigenvector of root k on the non-deflated rows rows[0..K),
// u_i = zhat_i / ((d_i - d_{o_k}) - tau_k), normalised.  One workgroup per root.
__global__ __launch_bounds__(256) void k_merge_vectors(const double* __restrict__ d, const double* __restrict__ zh,
                                                       int64_t K, const int32_t* __restrict__ org,
                                                       const double* __restrict__ tau, const int64_t* __restrict__ rows,
                                                       const int64_t* __restrict__ col, int64_t n,
                                                       double* __restrict__ Y) {
    __shared__ double sh[256];
    const int64_t k = blockIdx.x;
    const double dk = d[org[k]], tk = tau[k];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < K; i += 256) {
        const double u = zh[i] / ((d[i] - dk) - tk);
        s += u * u;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    const double inv = 1.0 / sqrt(sh[0]);
    double* y = Y + col[k] * n;
    for (int64_t i = threadIdx.x; i < K; i += 256) y[rows[i]] = zh[i] / ((d[i] - dk) - tk) * inv;
}

// Qfull column t = the half-eigenvector perm[t] (Q1 columns 0..k-1 on rows 0..k-1, Q2 columns on
// rows k..n-1), zero elsewhere.
__global__ void k_fill_qfull(const double* __restrict__ Q1, int64_t k, const double* __restrict__ Q2, int64_t n,
                             const int64_t* __restrict__ perm, double* __restrict__ Qf) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * n; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = e / n, r = e - t * n;
        const int64_t q = perm[t];
        double v = 0.0;
        if (q < k) {
            if (r < k) v = Q1[r + q * k];
        } else if (r >= k) {
            v = Q2[(r - k) + (q - k) * (n - k)];
        }
        Qf[e] = v;
    }
}

// The deflation rotations in order on the columns of Qf: column p <- c q_p - s q_q,
// column q <- s q_p + c q_q.  One thread per row.
__global__ void k_apply_rotations(double* __restrict__ Qf, int64_t n, const int64_t* __restrict__ rp,
                                  const int64_t* __restrict__ rq, const double* __restrict__ rc,
                                  const double* __restrict__ rs, int64_t nrot) {
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= n) return;
    for (int64_t t = 0; t < nrot; ++t) {
        double* a = Qf + r + rp[t] * n;
        double* b = Qf + r + rq[t] * n;
        const double x = *a, y = *b, c = rc[t], s = rs[t];
        *a = c * x - s * y;
        *b = s * x + c * y;
    }
}

__global__ void k_set_entries(double* __restrict__ Y, const int64_t* __restrict__ idx, int64_t m, double v) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < m) Y[idx[i]] = v;
}

// Eigenpairs of the symmetric tridiagonal (d, e) of size n (host arrays): eigenvalues ascending into
// lam (host), eigenvectors into E (n x n, ld n, device).  Blocks of at most `leaf` rows go to
// rocSOLVER's dstedc; larger ones are split in halves and merged.
void tridiag_eig_dc(rocblas_handle h, hipStream_t st, int64_t n, const std::vector<double>& d,
                    const std::vector<double>& e, std::vector<double>& lam, double* E, int64_t leaf) {
    lam.assign((size_t)n, 0.0);
    DevBuf info(sizeof(rocblas_int));
    if (n <= leaf) {
        DevBuf dd((size_t)n * 8), de((size_t)std::max<int64_t>(1, n - 1) * 8);
        CS_HIP(hipMemcpyAsync(dd.p, d.data(), (size_t)n * 8, hipMemcpyHostToDevice, st));
        if (n > 1) CS_HIP(hipMemcpyAsync(de.p, e.data(), (size_t)(n - 1) * 8, hipMemcpyHostToDevice, st));
        eig_chk(rocsolver_dstedc(h, rocblas_evect_tridiagonal, (rocblas_int)n, dd.as<double>(), de.as<double>(), E,
                                 (rocblas_int)n, info.as<rocblas_int>()), "dstedc");
        rocblas_int inf = 0;
        CS_HIP(hipMemcpyAsync(&inf, info.p, sizeof(inf), hipMemcpyDeviceToHost, st));
        CS_HIP(hipMemcpyAsync(lam.data(), dd.p, (size_t)n * 8, hipMemcpyDeviceToHost, st));
        CS_HIP(hipStreamSynchronize(st));
        if (inf != 0) throw RuntimeError("tridiagonal eigensolver did not converge");
        return;
    }
    const int64_t k = n / 2, n2 = n - k;
    const double beta = e[(size_t)(k - 1)], rho = std::fabs(beta), sg = beta >= 0.0 ? 1.0 : -1.0;
    // halves T1 - rho e_last e_last^T and T2 - rho e_0 e_0^T
    std::vector<double> d1(d.begin(), d.begin() + k), e1(e.begin(), e.begin() + (k - 1));
    std::vector<double> d2(d.begin() + k, d.end()), e2(e.begin() + k, e.end());
    d1[(size_t)(k - 1)] -= rho;
    d2[0] -= rho;
    std::vector<double> l1, l2;
    DevBuf Q1((size_t)(k * k) * 8), Q2((size_t)(n2 * n2) * 8);
    tridiag_eig_dc(h, st, k, d1, e1, l1, Q1.as<double>(), leaf);
    tridiag_eig_dc(h, st, n2, d2, e2, l2, Q2.as<double>(), leaf);
    // z = [last row of Q1; sign(beta) first row of Q2] / sqrt(2), rho -> 2 rho (|z| = 1)
    std::vector<double> z((size_t)n), dv((size_t)n);
    CS_HIP(hipMemcpy2DAsync(z.data(), 8, Q1.as<double>() + (k - 1), (size_t)k * 8, 8, (size_t)k,
                            hipMemcpyDeviceToHost, st));
    CS_HIP(hipMemcpy2DAsync(z.data() + k, 8, Q2.as<double>(), (size_t)n2 * 8, 8, (size_t)n2, hipMemcpyDeviceToHost,
                            st));
    CS_HIP(hipStreamSynchronize(st));
    const double rs2 = 1.0 / std::sqrt(2.0);
    for (int64_t i = 0; i < n; ++i) {
        dv[(size_t)i] = i < k ? l1[(size_t)i] : l2[(size_t)(i - k)];
        z[(size_t)i] *= (i < k ? 1.0 : sg) * rs2;
    }
    const double r2 = 2.0 * rho;
    // sorted order of the poles
    std::vector<int64_t> perm((size_t)n);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int64_t a, int64_t b) { return dv[(size_t)a] < dv[(size_t)b]; });
    std::vector<double> ds((size_t)n), zs((size_t)n);
    double dmax = 0.0, zmax = 0.0;
    for (int64_t t = 0; t < n; ++t) {
        ds[(size_t)t] = dv[(size_t)perm[(size_t)t]];
        zs[(size_t)t] = z[(size_t)perm[(size_t)t]];
        dmax = std::max(dmax, std::fabs(ds[(size_t)t]));
        zmax = std::max(zmax, std::fabs(zs[(size_t)t]));
    }
    const double eps = std::ldexp(1.0, -53);
    const double tol = 8.0 * eps * std::max(dmax, r2 * zmax);
    // deflation: small weights, then close pole pairs by rotations (the earlier pole deflates)
    std::vector<uint8_t> defl((size_t)n, 0);
    std::vector<int64_t> rp, rq;
    std::vector<double> rc, rsn;
    int64_t pj = -1;
    for (int64_t t = 0; t < n; ++t) {
        if (r2 * std::fabs(zs[(size_t)t]) <= tol) {
            defl[(size_t)t] = 1;
            continue;
        }
        if (pj >= 0) {
            const double a = zs[(size_t)pj], b = zs[(size_t)t];
            const double r = std::hypot(a, b), c = b / r, s = a / r;
            if (std::fabs((ds[(size_t)t] - ds[(size_t)pj]) * c * s) <= tol) {
                const double dp = ds[(size_t)pj], dq = ds[(size_t)t];
                ds[(size_t)pj] = c * c * dp + s * s * dq;
                ds[(size_t)t] = s * s * dp + c * c * dq;
                zs[(size_t)pj] = 0.0;
                zs[(size_t)t] = r;
                defl[(size_t)pj] = 1;
                rp.push_back(pj);
                rq.push_back(t);
                rc.push_back(c);
                rsn.push_back(s);
            }
        }
        pj = t;
    }
    std::vector<int64_t> nd;
    for (int64_t t = 0; t < n; ++t)
        if (!defl[(size_t)t]) nd.push_back(t);
    const int64_t K = (int64_t)nd.size();
    // secular roots on the device
    std::vector<double> dK((size_t)K), zK((size_t)K), tau((size_t)K);
    std::vector<int32_t> org((size_t)K);
    for (int64_t i = 0; i < K; ++i) {
        dK[(size_t)i] = ds[(size_t)nd[(size_t)i]];
        zK[(size_t)i] = zs[(size_t)nd[(size_t)i]];
    }
    DevBuf ddK((size_t)std::max<int64_t>(1, K) * 8), dzK((size_t)std::max<int64_t>(1, K) * 8),
        dorg((size_t)std::max<int64_t>(1, K) * 4), dtau((size_t)std::max<int64_t>(1, K) * 8),
        dzh((size_t)std::max<int64_t>(1, K) * 8);
    if (K > 0) {
        CS_HIP(hipMemcpyAsync(ddK.p, dK.data(), (size_t)K * 8, hipMemcpyHostToDevice, st));
        CS_HIP(hipMemcpyAsync(dzK.p, zK.data(), (size_t)K * 8, hipMemcpyHostToDevice, st));
        k_secular_roots<<<grid_for(K, 64), 64, 0, st>>>(ddK.as<double>(), dzK.as<double>(), K, r2, dorg.as<int32_t>(),
                                                         dtau.as<double>());
        k_zhat<<<grid_for(K, 64), 64, 0, st>>>(ddK.as<double>(), dzK.as<double>(), K, r2, dorg.as<int32_t>(),
                                                dtau.as<double>(), dzh.as<double>());
        CS_HIP(hipGetLastError());
        CS_HIP(hipMemcpyAsync(tau.data(), dtau.p, (size_t)K * 8, hipMemcpyDeviceToHost, st));
        CS_HIP(hipMemcpyAsync(org.data(), dorg.p, (size_t)K * 4, hipMemcpyDeviceToHost, st));
        CS_HIP(hipStreamSynchronize(st));
    }
    // all eigenvalues (roots and deflated poles) in ascending order -> output columns
    struct Ev {
        double v;
        int64_t src;   // >= 0: root index; < 0: -(sorted pole index) - 1
    };
    std::vector<Ev> ev;
    ev.reserve((size_t)n);
    for (int64_t i = 0; i < K; ++i) ev.push_back({dK[(size_t)org[(size_t)i]] + tau[(size_t)i], i});
    for (int64_t t = 0; t < n; ++t)
        if (defl[(size_t)t]) ev.push_back({ds[(size_t)t], -t - 1});
    std::stable_sort(ev.begin(), ev.end(), [](const Ev& a, const Ev& b) { return a.v < b.v; });
    std::vector<int64_t> colK((size_t)std::max<int64_t>(1, K)), unit;
    for (int64_t c = 0; c < n; ++c) {
        lam[(size_t)c] = ev[(size_t)c].v;
        if (ev[(size_t)c].src >= 0) colK[(size_t)ev[(size_t)c].src] = c;
        else unit.push_back((-ev[(size_t)c].src - 1) + c * n);   // Y(t, c) = 1
    }
    // Y (sorted-pole basis x output columns), Qfull (rows x sorted poles), E = Qfull Y
    DevBuf Y((size_t)(n * n) * 8), Qf((size_t)(n * n) * 8);
    CS_HIP(hipMemsetAsync(Y.p, 0, Y.bytes, st));
    if (K > 0) {
        DevBuf drows((size_t)K * 8), dcol((size_t)K * 8);
        CS_HIP(hipMemcpyAsync(drows.p, nd.data(), (size_t)K * 8, hipMemcpyHostToDevice, st));
        CS_HIP(hipMemcpyAsync(dcol.p, colK.data(), (size_t)K * 8, hipMemcpyHostToDevice, st));
        k_merge_vectors<<<(unsigned)K, 256, 0, st>>>(ddK.as<double>(), dzh.as<double>(), K, dorg.as<int32_t>(),
                                                     dtau.as<double>(), drows.as<int64_t>(), dcol.as<int64_t>(), n,
                                                     Y.as<double>());
        CS_HIP(hipGetLastError());
        CS_HIP(hipStreamSynchronize(st));
    }
    if (!unit.empty()) {
        DevBuf du(unit.size() * 8);
        CS_HIP(hipMemcpyAsync(du.p, unit.data(), unit.size() * 8, hipMemcpyHostToDevice, st));
        k_set_entries<<<grid_for((int64_t)unit.size()), 256, 0, st>>>(Y.as<double>(), du.as<int64_t>(),
                                                                      (int64_t)unit.size(), 1.0);
        CS_HIP(hipGetLastError());
        CS_HIP(hipStreamSynchronize(st));
    }
    {
        DevBuf dperm((size_t)n * 8);
        CS_HIP(hipMemcpyAsync(dperm.p, perm.data(), (size_t)n * 8, hipMemcpyHostToDevice, st));
        k_fill_qfull<<<grid_for(n * n), 256, 0, st>>>(Q1.as<double>(), k, Q2.as<double>(), n, dperm.as<int64_t>(),
                                                     Qf.as<double>());
        CS_HIP(hipGetLastError());
        if (!rp.empty()) {
            const int64_t nr = (int64_t)rp.size();
            DevBuf a(nr * 8), b(nr * 8), c(nr * 8), s(nr * 8);
            CS_HIP(hipMemcpyAsync(a.p, rp.data(), nr * 8, hipMemcpyHostToDevice, st));
            CS_HIP(hipMemcpyAsync(b.p, rq.data(), nr * 8, hipMemcpyHostToDevice, st));
            CS_HIP(hipMemcpyAsync(c.p, rc.data(), nr * 8, hipMemcpyHostToDevice, st));
            CS_HIP(hipMemcpyAsync(s.p, rsn.data(), nr * 8, hipMemcpyHostToDevice, st));
            k_apply_rotations<<<grid_for(n), 256, 0, st>>>(Qf.as<double>(), n, a.as<int64_t>(), b.as<int64_t>(),
                                                          c.as<double>(), s.as<double>(), nr);
            CS_HIP(hipGetLastError());
            CS_HIP(hipStreamSynchronize(st));
        }
        CS_HIP(hipStreamSynchronize(st));
    }
    Q1.release();
    Q2.release();
    const double one = 1.0, zero = 0.0;
    eig_chk(rocblas_dgemm_64(h, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, Qf.as<double>(), n,
                             Y.as<double>(), n, &zero, E, n), "merge gemm");
    CS_HIP(hipStreamSynchronize(st));
}

// Vb (mrow x w, ld mrow): reflectors p .. p+w-1 of a Householder set (reflector i: zero above
// row i, 1 on it, V(r, i) below) restricted to rows p .. p+mrow-1.
__global__ void k_build_v(const double* __restrict__ V, int64_t ldv, int64_t p, int64_t w, int64_t mrow,
                          double* __restrict__ Vb) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < mrow * w; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / mrow, r = e - i * mrow;
        double v = 0.0;
        if (r == i) v = 1.0;
        else if (r > i) v = V[(p + r) + (p + i) * ldv];
        Vb[e] = v;
    }
}

}  // namespace

// E (m x ncols, ld lde) <- Q E (transpose false) or Q^T E (true) for Q = H(0) ... H(c-1),
// H(i) = I - tau_i v_i v_i^T with v_i as in k_build_v.  Blocks of 256 reflectors (with 64 the
// GEMMs' inner and outer widths left them skinny, about 20 TFLOP/s), each applied as
// E(p:, :) -= V T' (V^T E(p:, :)) with T the compact-WY factor of LAPACK's dlarft (forward,
// column-wise) and T' = T or T^T: three FP64 GEMMs per block on the 64-bit rocBLAS interface.
void apply_reflectors(rocblas_handle h, hipStream_t st, const double* V, int64_t ldv, const double* tau_h, int64_t m,
                      int64_t c, bool transpose, double* E, int64_t lde, int64_t ncols) {
    constexpr int64_t nb = 256;
    if (c <= 0 || ncols <= 0) return;
    DevBuf Vb((size_t)(m * std::min(nb, c)) * 8), Gb((size_t)(nb * nb) * 8), Tb((size_t)(nb * nb) * 8),
        Wb((size_t)(nb * ncols) * 8), W2b((size_t)(nb * ncols) * 8);
    const double one = 1.0, zero = 0.0, mone = -1.0;
    eig_chk(rocblas_set_pointer_mode(h, rocblas_pointer_mode_host), "pointer mode");
    const int64_t nblk = (c + nb - 1) / nb;
    for (int64_t bi = 0; bi < nblk; ++bi) {
        const int64_t p = (transpose ? bi : nblk - 1 - bi) * nb;
        const int64_t w = std::min<int64_t>(nb, c - p);
        const int64_t mrow = m - p;
        double* Vp = Vb.as<double>();
        k_build_v<<<grid_for(mrow * w), 256, 0, st>>>(V, ldv, p, w, mrow, Vp);
        CS_HIP(hipGetLastError());
        // T from the Gram matrix V^T V: T(i, i) = tau_i, T(0:i, i) = -tau_i T(0:i, 0:i) (V^T v_i)(0:i)
        eig_chk(rocblas_dgemm_64(h, rocblas_operation_transpose, rocblas_operation_none, w, w, mrow, &one, Vp, mrow,
                                 Vp, mrow, &zero, Gb.as<double>(), w), "reflectors gram");
        std::vector<double> G((size_t)(w * w)), T((size_t)(w * w), 0.0);
        CS_HIP(hipMemcpyAsync(G.data(), Gb.p, G.size() * 8, hipMemcpyDeviceToHost, st));
        CS_HIP(hipStreamSynchronize(st));
        for (int64_t i = 0; i < w; ++i) {
            const double ti = tau_h[p + i];
            T[(size_t)(i + i * w)] = ti;
            for (int64_t a = 0; a < i; ++a) {
                double s = 0.0;
                for (int64_t b = a; b < i; ++b) s += T[(size_t)(a + b * w)] * G[(size_t)(b + i * w)];
                T[(size_t)(a + i * w)] = -ti * s;
            }
        }
        CS_HIP(hipMemcpyAsync(Tb.p, T.data(), T.size() * 8, hipMemcpyHostToDevice, st));
        double* Es = E + p;
        eig_chk(rocblas_dgemm_64(h, rocblas_operation_transpose, rocblas_operation_none, w, ncols, mrow, &one, Vp,
                                 mrow, Es, lde, &zero, Wb.as<double>(), w), "reflectors V^T E");
        eig_chk(rocblas_dgemm_64(h, transpose ? rocblas_operation_transpose : rocblas_operation_none,
                                 rocblas_operation_none, w, ncols, w, &one, Tb.as<double>(), w, Wb.as<double>(), w,
                                 &zero, W2b.as<double>(), w), "reflectors T");
        eig_chk(rocblas_dgemm_64(h, rocblas_operation_none, rocblas_operation_none, mrow, ncols, w, &mone, Vp, mrow,
                                 W2b.as<double>(), w, &one, Es, lde), "reflectors update");
        CS_HIP(hipStreamSynchronize(st));   // Vb and T are rebuilt by the next block
    }
}

namespace {

__global__ void k_copy_cols(const double* __restrict__ src, int64_t lds, double* __restrict__ dst, int64_t ldd,
                            int64_t m, int64_t ncol) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < m * ncol; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = e / m, r = e - c * m;
        dst[r + c * ldd] = src[r + c * lds];
    }
}

}  // namespace

int64_t eig_large_leaf() {
    const char* e = std::getenv("CSAMD_EIG_LARGE");
    const long long v = e ? std::atoll(e) : 0;
    return (int64_t)(v > 0 ? std::min<long long>(v, kDenseEigenMaxN) : 0);
}

void sym_eig_large(rocblas_handle h, hipStream_t st, int64_t n, double* A, int64_t lda, double* D, int64_t leaf) {
    if (n < 1) return;
    if (leaf < 1 || leaf > kDenseEigenMaxN) leaf = kDenseEigenMaxN;
    if (n == 1) {
        CS_HIP(hipMemcpyAsync(D, A, 8, hipMemcpyDeviceToDevice, st));
        k_set_scalar<<<1, 1, 0, st>>>(A, 1.0);
        CS_HIP(hipGetLastError());
        CS_HIP(hipStreamSynchronize(st));
        return;
    }
    DevBuf diag((size_t)n * 8), off((size_t)n * 8), tau((size_t)n * 8);
    sytrd_lower(h, st, n, A, lda, diag.as<double>(), off.as<double>(), tau.as<double>());
    std::vector<double> d((size_t)n), e((size_t)(n - 1)), tau_h((size_t)(n - 1)), lam;
    CS_HIP(hipMemcpyAsync(d.data(), diag.p, (size_t)n * 8, hipMemcpyDeviceToHost, st));
    CS_HIP(hipMemcpyAsync(e.data(), off.p, (size_t)(n - 1) * 8, hipMemcpyDeviceToHost, st));
    CS_HIP(hipMemcpyAsync(tau_h.data(), tau.p, (size_t)(n - 1) * 8, hipMemcpyDeviceToHost, st));
    CS_HIP(hipStreamSynchronize(st));
    DevBuf E((size_t)(n * n) * 8);
    tridiag_eig_dc(h, st, n, d, e, lam, E.as<double>(), leaf);
    // Q = H(0) ... H(n-2) with H(j) acting on rows j+1..: the reflector set of rows 1.. of A
    apply_reflectors(h, st, A + 1, lda, tau_h.data(), n - 1, n - 1, false, E.as<double>() + 1, n, n);
    k_copy_cols<<<grid_for(n * n), 256, 0, st>>>(E.as<double>(), n, A, lda, n, n);
    CS_HIP(hipGetLastError());
    CS_HIP(hipMemcpyAsync(D, lam.data(), (size_t)n * 8, hipMemcpyHostToDevice, st));
    CS_HIP(hipStreamSynchronize(st));
}

}  // namespace csamd

using namespace csamd;

extern "C" int cs_sym_eigen_device(double* d_A, int64_t n, int64_t lda, double* d_D, int64_t leaf) {
    return guarded([&] {
        if (n < 0 || (n > 0 && (!d_A || !d_D))) throw InvalidArgument("cs_sym_eigen_device: null argument");
        if (lda < n) throw InvalidArgument("cs_sym_eigen_device: lda < n");
        if (leaf < 0 || leaf > kDenseEigenMaxN)
            throw InvalidArgument("cs_sym_eigen_device: leaf must be 0 (rocSOLVER dsyevd) or 1..32766");
        if (n == 0) return;
        hipStream_t st;
        CS_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        rocblas_handle h = nullptr;
        try {
            eig_chk(rocblas_create_handle(&h), "handle");
            eig_chk(rocblas_set_stream(h, st), "stream");
            if (leaf > 0) {
                sym_eig_large(h, st, n, d_A, lda, d_D, leaf);
            } else {
                if (n > kDenseEigenMaxN) throw InvalidArgument("cs_sym_eigen_device: leaf 0 (rocSOLVER dsyevd) supports n <= 32766");
                DevBuf E((size_t)n * 8), info(sizeof(rocblas_int));
                eig_chk(rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_lower, (rocblas_int)n, d_A,
                                         (rocblas_int)lda, d_D, E.as<double>(), info.as<rocblas_int>()),
                        "dsyevd");
                rocblas_int inf = 0;
                CS_HIP(hipMemcpyAsync(&inf, info.p, sizeof(inf), hipMemcpyDeviceToHost, st));
                CS_HIP(hipStreamSynchronize(st));
                if (inf != 0) throw RuntimeError("cs_sym_eigen_device: eigendecomposition did not converge");
            }
        } catch (...) {
            if (h) (void)rocblas_destroy_handle(h);
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
            throw;
        }
        (void)rocblas_destroy_handle(h);
        CS_HIP(hipStreamDestroy(st));
    });
}
