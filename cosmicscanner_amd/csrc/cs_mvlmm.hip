// Multivariate LMM on the device: the null model of MvEigenVce (libCosmicCore/src/eigen_vce.cpp:
// 496-910, 1031-1066) and the fixed-null Wald scan of MvLmmEngine::scan(PlinkReader&)
// (src/mvlmm_engine.cpp:210-394, 756-931), with covariate projection (the default).
//
// prepare: Householder QR of X, A = Q[:, c:], G_reml = A^T G A (from the c reflectors, no
//   explicit Q), eigendecomposition (D clamped at 0), W_proj = A U, Uty = U^T A^T Y (ne x d);
//   start values Vg = Ve = Cov(Y) / 2.
// Null model: EM (updateEM) or EM warm-up + AI-REML with step halving (runAireml).  Every O(ne)
//   sum -- the EM statistics, the AI matrix and score, the REML log-likelihood -- is one thread
//   per eigen-sample on the GPU followed by a fixed-order column reduction; the d(d+1) x d(d+1)
//   solve and the iteration control run on the host.
// Scan: per SNP the centred marker g (keep-subset mean of observed genotypes, missing -> 0),
//   s = W_proj^T g, H = sum_c s_c^2 Sigma_c^-1, r = sum_c s_c Sigma_c^-1 y_c (Sigma_c = D_c Vg + Ve),
//   beta = H^-1 r through the symmetric eigendecomposition of H (invertInformation), Wald =
//   beta^T H beta ~ chi^2_d, Var(beta) = H^-1 conditioned (conditionCovariance).  The rotation
//   s = W^T g for all SNPs is the dominant cost (2 n ne m): it runs as FP16 hi/lo split products
//   on MFMA (z = g split per SNP through a 4-entry table, W split once), FP32 accumulation over
//   the samples, with H and r fused into the epilogue as per-column-block partial sums.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "cs_blas.h"
#include "cs_common.h"
#include "cs_ops.h"
#include "cs_stats.h"

namespace csamd {

#define CS_MVBLAS(call)                                                                 \
    do {                                                                                \
        rocblas_status s_ = (call);                                                     \
        if (s_ != rocblas_status_success)                                               \
            throw ::csamd::RuntimeError(std::string("rocBLAS/rocSOLVER error ") +      \
                                        rocblas_status_to_string(s_) + " at " #call);   \
    } while (0)

constexpr int kMvMaxD = 6;

// ------------------------------------------------------------------ small SPD algebra (device)
template <int D>
__device__ inline bool chol_d(const double (&A)[D][D], double (&L)[D][D]) {
#pragma unroll
    for (int j = 0; j < D; ++j) {
        double s = A[j][j];
#pragma unroll
        for (int k = 0; k < j; ++k) s -= L[j][k] * L[j][k];
        if (!(s > 0.0)) return false;
        L[j][j] = sqrt(s);
#pragma unroll
        for (int i = j + 1; i < D; ++i) {
            double t = A[i][j];
#pragma unroll
            for (int k = 0; k < j; ++k) t -= L[i][k] * L[j][k];
            L[i][j] = t / L[j][j];
        }
#pragma unroll
        for (int i = 0; i < j; ++i) L[i][j] = 0.0;
    }
    return true;
}

// x = A^-1 b from the Cholesky factor.
template <int D>
__device__ inline void chol_solve_d(const double (&L)[D][D], const double (&b)[D], double (&x)[D]) {
    double y[D];
#pragma unroll
    for (int i = 0; i < D; ++i) {
        double t = b[i];
#pragma unroll
        for (int k = 0; k < i; ++k) t -= L[i][k] * y[k];
        y[i] = t / L[i][i];
    }
#pragma unroll
    for (int i = D - 1; i >= 0; --i) {
        double t = y[i];
#pragma unroll
        for (int k = i + 1; k < D; ++k) t -= L[k][i] * x[k];
        x[i] = t / L[i][i];
    }
}

template <int D>
__device__ inline void chol_inverse_d(const double (&L)[D][D], double (&Ai)[D][D]) {
#pragma unroll
    for (int c = 0; c < D; ++c) {
        double e[D], x[D];
#pragma unroll
        for (int i = 0; i < D; ++i) e[i] = i == c ? 1.0 : 0.0;
        chol_solve_d<D>(L, e, x);
#pragma unroll
        for (int i = 0; i < D; ++i) Ai[i][c] = x[i];
    }
}

// Cyclic Jacobi eigendecomposition of a symmetric D x D matrix: A = V diag(w) V^T.
template <int D>
__device__ inline void jacobi_d(const double (&S)[D][D], double (&w)[D], double (&V)[D][D]) {
    double A[D][D];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
            A[i][j] = S[i][j];
            V[i][j] = i == j ? 1.0 : 0.0;
        }
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0, tot = 0.0;
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = 0; j < D; ++j) {
                tot += A[i][j] * A[i][j];
                if (i != j) off += A[i][j] * A[i][j];
            }
        if (off <= 1e-30 * tot || off == 0.0) break;
#pragma unroll
        for (int p = 0; p < D; ++p)
#pragma unroll
            for (int q = p + 1; q < D; ++q) {
                if (A[p][q] == 0.0) continue;
                const double theta = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    const double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - s * akq;
                    A[k][q] = s * akp + c * akq;
                }
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    const double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - s * aqk;
                    A[q][k] = s * apk + c * aqk;
                }
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    const double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
    }
#pragma unroll
    for (int i = 0; i < D; ++i) w[i] = A[i][i];
}

// ------------------------------------------------------------------ null-model per-sample sums
// Row i of `out` (row-major, K columns) gets sample i's contribution; k_colsum_rows reduces.
template <int D>
__device__ inline void load_vgve(const double* __restrict__ P, double (&Vg)[D][D], double (&Ve)[D][D]) {
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
            Vg[i][j] = P[i * D + j];
            Ve[i][j] = P[D * D + i * D + j];
        }
}

// updateEM (eigen_vce.cpp:551-608) after projection (no covariates): K = 2 D^2.
template <int D>
__global__ void k_mv_em(const double* __restrict__ Dv, const double* __restrict__ Uty, int64_t ne,
                        const double* __restrict__ P, double* __restrict__ out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= ne) return;
    double Vg[D][D], Ve[D][D], V[D][D], L[D][D], Vi[D][D], y[D];
    load_vgve<D>(P, Vg, Ve);
    const double di = Dv[i];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        y[a] = Uty[a * ne + i];
#pragma unroll
        for (int b = 0; b < D; ++b) V[a][b] = Vg[a][b] * di + Ve[a][b];
    }
    double* o = out + i * (2 * D * D);
    if (!chol_d<D>(V, L)) {
        for (int k = 0; k < 2 * D * D; ++k) o[k] = __builtin_nan("");
        return;
    }
    chol_inverse_d<D>(L, Vi);
    // W_g = Vg D_i P_i, u = W_g y; W_e = Ve P_i, e = W_e y
    double Wg[D][D], We[D][D], u[D], e[D];
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) {
            double sg = 0.0, se = 0.0;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                sg += Vg[a][k] * di * Vi[k][b];
                se += Ve[a][k] * Vi[k][b];
            }
            Wg[a][b] = sg;
            We[a][b] = se;
        }
#pragma unroll
    for (int a = 0; a < D; ++a) {
        double su = 0.0, sv = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            su += Wg[a][k] * y[k];
            sv += We[a][k] * y[k];
        }
        u[a] = su;
        e[a] = sv;
    }
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) {
            double wgvg = 0.0, weve = 0.0;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                wgvg += Wg[a][k] * Vg[k][b];
                weve += We[a][k] * Ve[k][b];
            }
            o[a * D + b] = di > 1e-12 ? u[a] * u[b] / di + Vg[a][b] - wgvg : Vg[a][b];
            o[D * D + a * D + b] = e[a] * e[b] + Ve[a][b] - weve;
        }
}

// REML log-likelihood term (calcLogLikelihood, eigen_vce.cpp:617-632): -0.5 (log|V_i| + y_i^T V_i^-1 y_i)
// with V_i = D_i Vg + Ve + 1e-8 I; NaN when V_i is not positive definite.
template <int D>
__global__ void k_mv_logl(const double* __restrict__ Dv, const double* __restrict__ Uty, int64_t ne,
                          const double* __restrict__ P, double* __restrict__ out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= ne) return;
    double Vg[D][D], Ve[D][D], V[D][D], L[D][D], y[D], x[D];
    load_vgve<D>(P, Vg, Ve);
    const double di = Dv[i];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        y[a] = Uty[a * ne + i];
#pragma unroll
        for (int b = 0; b < D; ++b) V[a][b] = di * Vg[a][b] + Ve[a][b] + (a == b ? 1e-8 : 0.0);
    }
    if (!chol_d<D>(V, L)) {
        out[i] = __builtin_nan("");
        return;
    }
    chol_solve_d<D>(L, y, x);
    double quad = 0.0, ldet = 0.0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        quad += y[a] * x[a];
        ldet += 2.0 * log(L[a][a]);
    }
    out[i] = -0.5 * (ldet + quad);
}

// AI-REML score and average-information contributions (runAireml, eigen_vce.cpp:688-737):
// K = 2 vs + 3 vs (vs + 1) / 2 with vs = D (D + 1) / 2: [d_i s_a | s_a | (d_i^2 q_ab, d_i q_ab, q_ab)_{a<=b}].
template <int D>
__global__ void k_mv_ai(const double* __restrict__ Dv, const double* __restrict__ Uty, int64_t ne,
                        const double* __restrict__ P, double* __restrict__ out) {
    constexpr int VS = D * (D + 1) / 2;
    constexpr int K = 2 * VS + 3 * VS * (VS + 1) / 2;
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= ne) return;
    double Vg[D][D], Ve[D][D], V[D][D], L[D][D], Vi[D][D], y[D], v[D];
    load_vgve<D>(P, Vg, Ve);
    const double di = Dv[i], d2 = di * di;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        y[a] = Uty[a * ne + i];
#pragma unroll
        for (int b = 0; b < D; ++b) V[a][b] = di * Vg[a][b] + Ve[a][b] + (a == b ? 1e-8 : 0.0);
    }
    double* o = out + i * K;
    if (!chol_d<D>(V, L)) {
        for (int k = 0; k < K; ++k) o[k] = __builtin_nan("");
        return;
    }
    chol_inverse_d<D>(L, Vi);
#pragma unroll
    for (int a = 0; a < D; ++a) {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) t += Vi[a][k] * y[k];
        v[a] = t;
    }
    // w_b(x) = (V^-1 E_pq v)(x): p == q: v_p V^-1(x, p); else v_q V^-1(x, p) + v_p V^-1(x, q)
    int ia = 0;
#pragma unroll
    for (int j = 0; j < D; ++j)
#pragma unroll
        for (int k = j; k < D; ++k) {
            const double vjvk = v[j] * v[k];
            const double tr = j == k ? Vi[j][j] : 2.0 * Vi[j][k];
            const double quad = j == k ? vjvk : 2.0 * vjvk;
            o[ia] = 0.5 * di * (quad - tr);
            o[VS + ia] = 0.5 * (quad - tr);
            ++ia;
        }
    int pairs = 0;
    int a = 0;
#pragma unroll
    for (int j = 0; j < D; ++j)
#pragma unroll
        for (int k = j; k < D; ++k) {
            int b = 0;
#pragma unroll
            for (int p = 0; p < D; ++p)
#pragma unroll
                for (int q = p; q < D; ++q) {
                    if (b >= a) {
                        auto wb = [&](int x) {
                            return p == q ? v[p] * Vi[x][p] : v[q] * Vi[x][p] + v[p] * Vi[x][q];
                        };
                        double qv = j == k ? v[j] * wb(j) : v[j] * wb(k) + v[k] * wb(j);
                        qv *= 0.5;
                        double* oq = o + 2 * VS + 3 * pairs;
                        oq[0] = d2 * qv;
                        oq[1] = di * qv;
                        oq[2] = qv;
                        ++pairs;
                    }
                    ++b;
                }
            ++a;
        }
}

// Fixed-order column sums of a row-major rows x K matrix: one block per column.
__global__ __launch_bounds__(256) void k_colsum_rows(const double* __restrict__ A, int64_t rows, int64_t K,
                                                     double* __restrict__ out) {
    __shared__ double red[256];
    const int64_t c = blockIdx.x;
    double s = 0.0;
    for (int64_t r = threadIdx.x; r < rows; r += 256) s += A[r * K + c];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = red[0];
}

// ------------------------------------------------------------------ scan
// Per eigen-sample weights of the scan: Sigma_c^-1 entries (upper triangle, pair order) and
// Sigma_c^-1 y_c (MvLmmEngine::scan's Sigma_inv = (D_c Vg + Ve).llt().solve(I), mvlmm_engine.cpp:
// 774-783).  Row c of wt: [nq pairs | D].  Columns >= ne are zero.
template <int D>
__global__ void k_mv_weights(const double* __restrict__ Dv, const double* __restrict__ Uty, int64_t ne,
                             int64_t ne_pad, const double* __restrict__ P, double* __restrict__ wt,
                             int* __restrict__ bad) {
    constexpr int NQ = D * (D + 1) / 2;
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= ne_pad) return;
    double* o = wt + c * (NQ + D);
    if (c >= ne) {
        for (int k = 0; k < NQ + D; ++k) o[k] = 0.0;
        return;
    }
    double Vg[D][D], Ve[D][D], V[D][D], L[D][D], Vi[D][D], y[D];
    load_vgve<D>(P, Vg, Ve);
    const double dc = Dv[c];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        y[a] = Uty[a * ne + c];
#pragma unroll
        for (int b = 0; b < D; ++b) V[a][b] = dc * Vg[a][b] + Ve[a][b];
    }
    if (!chol_d<D>(V, L)) {
        atomicOr(bad, 1);
        for (int k = 0; k < NQ + D; ++k) o[k] = 0.0;
        return;
    }
    chol_inverse_d<D>(L, Vi);
    int ia = 0;
#pragma unroll
    for (int p = 0; p < D; ++p)
#pragma unroll
        for (int q = p; q < D; ++q) o[ia++] = Vi[p][q];
#pragma unroll
    for (int t = 0; t < D; ++t) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) s += Vi[t][k] * y[k];
        o[NQ + t] = s;
    }
}

// W_proj over the kept rows -> FP16 hi/lo over the full sample space, in the rotation kernel's
// B layout: for 32-sample step s and column c, 128 bytes at (s ne_pad + c) 128: hi[32] | lo[32].
// Sample S gets the sum of the W rows of its keep entries (keep multiplicities).
__global__ void k_mv_wsplit(const double* __restrict__ W, int64_t n, int64_t ne, int64_t ne_pad,
                            const int32_t* __restrict__ inv_off, const int32_t* __restrict__ inv_idx,
                            int64_t N, int64_t Npad, _Float16* __restrict__ Bw) {
    const int64_t total = Npad * ne_pad;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = e % Npad, c = e / Npad;
        double v = 0.0;
        if (s < N && c < ne)
            for (int32_t p = inv_off[s]; p < inv_off[s + 1]; ++p) v += W[c * n + inv_idx[p]];
        const _Float16 hi = (_Float16)v;
        const _Float16 lo = (_Float16)(v - (double)hi);
        _Float16* o = Bw + ((s >> 5) * ne_pad + c) * 64;
        o[s & 31] = hi;
        o[32 + (s & 31)] = lo;
    }
}

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int kMvRow = 144;              // LDS bytes per row: 32 hi + 32 lo FP16 + 16 pad
constexpr int kMvBuf = 512 * kMvRow;

struct MvRotArgs {
    const uint8_t* snp;      // SNP-major tiled genotypes
    int64_t KBs;             // 256-sample blocks per SNP row tile
    int64_t row0;            // first SNP row of this launch (multiple of 256)
    int64_t steps;           // 32-sample steps (Npad / 32)
    const uint4* tab;        // per SNP row of the launch: FP16 hi/lo table of the centred marker
    const _Float16* Bw;      // W hi/lo (k_mv_wsplit layout)
    int64_t ne_pad;
    const double* wt;        // ne_pad x (NQ + D) weights
    int nq, nw;              // NQ quadratic-form weights, NQ + D weights in all
    int64_t ncb;             // column blocks of 128 (ne_pad / 128)
    double* part;            // [row - row0][cb][nw]
};

// One 256-SNP x 256-column block per workgroup of 8 waves; per 32-sample step the workgroup
// decodes its 256 marker rows into LDS (one v_perm per FP16 pair from the SNP's 8-byte table),
// copies its 256 W columns, and wave (wr, wc) multiplies SNP rows 64 wr .. +64 by columns
// 128 wc .. +128 (Z_hi W_hi + Z_hi W_lo + Z_lo W_hi).  The epilogue reduces s^2 w and s w over the
// wave's 128 columns into per-(SNP, column block) partial sums.
__global__ __launch_bounds__(512) void k_mv_rot(MvRotArgs a) {
    __shared__ __align__(16) uint8_t lds[2 * kMvBuf];
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int64_t rb = blockIdx.x, cbk = blockIdx.y;
    // decode role: marker row dr (local), 16 samples (half dh) per step
    const int dr = t >> 1, dh = t & 1;
    const uint4 tb = a.tab[rb * 256 + dr];
    const int64_t grow = a.row0 + rb * 256 + dr;
    const uint8_t* arow = a.snp + (((grow >> 6) * a.KBs) << 12) +
                          ((((grow & 63) >> 4) * 64 + (grow & 15)) << 4) + 4 * dh;
    // copy role: W column cc (local), half ch (hi or lo)
    const int cc = t >> 1, ch = t & 1;
    const _Float16* wcol = a.Bw + (cbk * 256 + cc) * 64 + ch * 32;
    auto stage = [&](int64_t s, uint8_t* buf) {
        // marker: 16 samples from the SNP-major tile (step s: kb = s >> 3, byte 8 (s & 7))
        const uint32_t w = *reinterpret_cast<const uint32_t*>(arow + ((s >> 3) << 12) +
                                                               ((((s & 7) >> 1) * 16) << 4) + (s & 1) * 8);
        const uint32_t d0 = w & 0x03030303u, d1 = (w >> 2) & 0x03030303u;
        const uint32_t d2 = (w >> 4) & 0x03030303u, d3 = (w >> 6) & 0x03030303u;
        uint32_t hi[8], lo[8];
#pragma unroll
        for (int bb = 0; bb < 4; ++bb) {
            const uint32_t pick = (uint32_t)bb * 0x0101u + (uint32_t)(4 + bb) * 0x01010000u;
            const uint32_t s01 = (__builtin_amdgcn_perm(d1, d0, pick) << 1) | 0x01000100u;
            const uint32_t s23 = (__builtin_amdgcn_perm(d3, d2, pick) << 1) | 0x01000100u;
            hi[2 * bb] = __builtin_amdgcn_perm(tb.y, tb.x, s01);
            hi[2 * bb + 1] = __builtin_amdgcn_perm(tb.y, tb.x, s23);
            lo[2 * bb] = __builtin_amdgcn_perm(tb.w, tb.z, s01);
            lo[2 * bb + 1] = __builtin_amdgcn_perm(tb.w, tb.z, s23);
        }
        uint8_t* row = buf + dr * kMvRow + 32 * dh;
        *reinterpret_cast<uint4*>(row) = uint4{hi[0], hi[1], hi[2], hi[3]};
        *reinterpret_cast<uint4*>(row + 16) = uint4{hi[4], hi[5], hi[6], hi[7]};
        *reinterpret_cast<uint4*>(row + 64) = uint4{lo[0], lo[1], lo[2], lo[3]};
        *reinterpret_cast<uint4*>(row + 80) = uint4{lo[4], lo[5], lo[6], lo[7]};
        // W column: 64 bytes (hi or lo half of the step)
        const uint4* src = reinterpret_cast<const uint4*>(wcol + s * a.ne_pad * 64);
        uint8_t* dst = buf + (256 + cc) * kMvRow + 64 * ch;
#pragma unroll
        for (int q = 0; q < 4; ++q) reinterpret_cast<uint4*>(dst)[q] = src[q];
    };
    const int wr = wave >> 1, wc = wave & 1, g = lane >> 4, r = lane & 15;
    f32x4 acc[4][8];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int n = 0; n < 8; ++n) acc[m][n] = f32x4{0.f, 0.f, 0.f, 0.f};
    stage(0, lds);
    __syncthreads();
    const int aoff = (64 * wr + r) * kMvRow + 16 * g;
    const int boff = (256 + 128 * wc + r) * kMvRow + 16 * g;
    for (int64_t s = 0; s < a.steps; ++s) {
        const int cur = (int)(s & 1);
        const uint8_t* buf = lds + cur * kMvBuf;
        f16x8 ah[4], al[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            ah[m] = *reinterpret_cast<const f16x8*>(buf + aoff + 16 * m * kMvRow);
            al[m] = *reinterpret_cast<const f16x8*>(buf + aoff + 16 * m * kMvRow + 64);
        }
#pragma unroll
        for (int n = 0; n < 8; ++n) {
            const f16x8 bh = *reinterpret_cast<const f16x8*>(buf + boff + 16 * n * kMvRow);
            const f16x8 bl = *reinterpret_cast<const f16x8*>(buf + boff + 16 * n * kMvRow + 64);
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[m], bh, acc[m][n], 0, 0, 0);
                acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[m], bl, acc[m][n], 0, 0, 0);
                acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[m], bh, acc[m][n], 0, 0, 0);
            }
        }
        if (s + 1 < a.steps) stage(s + 1, lds + (cur ^ 1) * kMvBuf);
        __syncthreads();
    }
    // epilogue: lane holds s[row = 64 wr + 16 m + 4 g + v][col = 128 wc + 16 n + r].  The block's
    // 256 columns of weights go to LDS (free after the loop); per row the wave forms
    // sum_c s_c^2 w_c and sum_c s_c w_c over its 128 columns (16 lanes x 8 tiles).
    float* wl = reinterpret_cast<float*>(lds);
    for (int e = t; e < 256 * a.nw; e += 512) {
        const int col = e / a.nw, k = e - col * a.nw;
        wl[e] = (float)a.wt[(cbk * 256 + col) * a.nw + k];
    }
    __syncthreads();
    const int64_t cb = cbk * 2 + wc;
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            float sv[8];
#pragma unroll
            for (int n = 0; n < 8; ++n) sv[n] = acc[m][n][v];
            const int64_t row = rb * 256 + 64 * wr + 16 * m + 4 * g + v;
            for (int k = 0; k < a.nw; ++k) {
                const bool is_q = k < a.nq;   // s^2 Sigma^-1 entries, then s (Sigma^-1 y)
                float x = 0.f;
#pragma unroll
                for (int n = 0; n < 8; ++n) {
                    const float w = wl[(128 * wc + 16 * n + r) * a.nw + k];
                    x += (is_q ? sv[n] * sv[n] : sv[n]) * w;
                }
                x += __shfl_xor(x, 1);
                x += __shfl_xor(x, 2);
                x += __shfl_xor(x, 4);
                x += __shfl_xor(x, 8);
                if (r == 0) a.part[(row * a.ncb + cb) * a.nw + k] = (double)x;
            }
        }
}

// Per SNP: sum the column-block partials in order, then invertInformation / Wald / chi^2_D /
// conditionCovariance (mvlmm_engine.cpp:289-352, src/mvlmm_numerics.h:27-66).
// status: 0 tested (OK), 1 tested (NUMERICAL_ROUNDOFF), 2 tested (BOUNDARY_SOLUTION),
// -1 not tested.
template <int D>
__global__ void k_mv_final(const double* __restrict__ part, int64_t rows, int64_t ncb, int64_t row0,
                           const int32_t* __restrict__ sel, const int64_t* __restrict__ counts,
                           int64_t M, double* __restrict__ freq, double* __restrict__ beta,
                           double* __restrict__ vbeta, double* __restrict__ wald, double* __restrict__ p,
                           int32_t* __restrict__ status) {
    constexpr int NQ = D * (D + 1) / 2;
    constexpr int NW = NQ + D;
    const int64_t lr = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (lr >= rows) return;
    const int64_t j = row0 + lr;
    if (j >= M) return;
    const int64_t o = sel[j];
    if (o < 0) return;
    const int64_t* c = counts + j * 4;
    const int64_t obs = c[0] + c[2] + c[3];
    const double mean = obs > 0 ? (2.0 * (double)c[0] + (double)c[2]) / (double)obs : 0.0;
    freq[o] = mean / 2.0;
    double h[NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) h[k] = 0.0;
    for (int64_t cb = 0; cb < ncb; ++cb) {
        const double* pp = part + (lr * ncb + cb) * NW;
#pragma unroll
        for (int k = 0; k < NW; ++k) h[k] += pp[k];
    }
    double H[D][D], rhs[D];
    int ia = 0;
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = a; b < D; ++b) {
            H[a][b] = h[ia];
            H[b][a] = h[ia];
            ++ia;
        }
#pragma unroll
    for (int t = 0; t < D; ++t) rhs[t] = h[NQ + t];
    status[o] = -1;
    wald[o] = p[o] = __builtin_nan("");
#pragma unroll
    for (int a = 0; a < D; ++a) {
        beta[o * D + a] = __builtin_nan("");
#pragma unroll
        for (int b = 0; b < D; ++b) vbeta[(o * D + a) * D + b] = __builtin_nan("");
    }
    bool finite = true;
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) finite = finite && isfinite(H[a][b]);
    if (!finite) return;
    double w[D], V[D][D];
    jacobi_d<D>(H, w, V);
    double wmin = w[0], wabs = fabs(w[0]);
#pragma unroll
    for (int a = 1; a < D; ++a) {
        wmin = fmin(wmin, w[a]);
        wabs = fmax(wabs, fabs(w[a]));
    }
    const double eps = 2.220446049250313e-16;
    const double tol = 128.0 * eps * fmax(1.0, wabs) * D;
    if (!(wmin > tol)) return;
    double C[D][D];
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < D; ++k) s += V[a][k] * V[b][k] / w[k];
            C[a][b] = s;
        }
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = a + 1; b < D; ++b) {
            const double s = 0.5 * (C[a][b] + C[b][a]);
            C[a][b] = C[b][a] = s;
        }
    double bt[D];
    bool ok = true;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) s += C[a][k] * rhs[k];
        bt[a] = s;
        ok = ok && isfinite(s);
    }
    if (!ok) return;
    double wd = 0.0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) s += H[a][k] * bt[k];
        wd += bt[a] * s;
    }
    if (!isfinite(wd) || wd < 0.0) return;
    const double pv = d_chisq_prob(wd, D);
    // conditionCovariance of Var(beta) = C
    double cw[D], CV[D][D];
    jacobi_d<D>(C, cw, CV);
    double cmin = cw[0], cabs = fabs(cw[0]);
#pragma unroll
    for (int a = 1; a < D; ++a) {
        cmin = fmin(cmin, cw[a]);
        cabs = fmax(cabs, fabs(cw[a]));
    }
    const double ctol = 128.0 * eps * fmax(1.0, cabs) * D;
    int st = 0;
    if (cmin < -ctol) return;
    if (cmin < 0.0) {
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
            for (int b = 0; b < D; ++b) {
                double s = 0.0;
#pragma unroll
                for (int k = 0; k < D; ++k) s += CV[a][k] * CV[b][k] * fmax(cw[k], 0.0);
                C[a][b] = s;
            }
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
            for (int b = a + 1; b < D; ++b) {
                const double s = 0.5 * (C[a][b] + C[b][a]);
                C[a][b] = C[b][a] = s;
            }
        st = 1;
    } else if (cmin <= ctol) {
        st = 2;
    }
#pragma unroll
    for (int a = 0; a < D; ++a) {
        beta[o * D + a] = bt[a];
#pragma unroll
        for (int b = 0; b < D; ++b) vbeta[(o * D + a) * D + b] = C[a][b];
    }
    wald[o] = wd;
    p[o] = pv;
    status[o] = st;
}

// ------------------------------------------------------------------ host small algebra
// Symmetric eigendecomposition (cyclic Jacobi) of an n x n row-major matrix.
static void host_jacobi(std::vector<double> A, int n, std::vector<double>& w, std::vector<double>& V) {
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                const double x = A[(size_t)i * n + j];
                tot += x * x;
                if (i != j) off += x * x;
            }
        if (off == 0.0 || off <= 1e-30 * tot) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[(size_t)p * n + q];
                if (apq == 0.0) continue;
                const double theta = (A[(size_t)q * n + q] - A[(size_t)p * n + p]) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    const double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
                    A[(size_t)k * n + p] = c * akp - s * akq;
                    A[(size_t)k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
                    A[(size_t)p * n + k] = c * apk - s * aqk;
                    A[(size_t)q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
                    V[(size_t)k * n + p] = c * vkp - s * vkq;
                    V[(size_t)k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    w.resize((size_t)n);
    for (int i = 0; i < n; ++i) w[(size_t)i] = A[(size_t)i * n + i];
}

// nearPD of runAireml (eigen_vce.cpp:662-669): clip eigenvalues at 1e-10 when the smallest is <= 1e-10.
static void near_pd(std::vector<double>& M, int d) {
    std::vector<double> w, V;
    host_jacobi(M, d, w, V);
    const double mn = *std::min_element(w.begin(), w.end());
    if (mn > 1e-10) return;
    for (double& x : w) x = std::max(x, 1e-10);
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j) {
            double s = 0.0;
            for (int k = 0; k < d; ++k) s += V[(size_t)i * d + k] * w[(size_t)k] * V[(size_t)j * d + k];
            M[(size_t)i * d + j] = s;
        }
}

// LDL^T with symmetric diagonal pivoting (largest remaining |diagonal| first); solves A X = B
// for the columns of B (row-major n x nb, overwritten).  False when a pivot is zero or not finite.
static bool ldlt_solve(std::vector<double> A, int n, std::vector<double>& B, int nb) {
    std::vector<int> perm((size_t)n);
    for (int i = 0; i < n; ++i) perm[(size_t)i] = i;
    auto at = [&](int i, int j) -> double& { return A[(size_t)i * n + j]; };
    for (int k = 0; k < n; ++k) {
        int piv = k;
        double best = std::fabs(at(k, k));
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(at(i, i)) > best) {
                best = std::fabs(at(i, i));
                piv = i;
            }
        if (piv != k) {
            for (int j = 0; j < n; ++j) std::swap(at(k, j), at(piv, j));
            for (int i = 0; i < n; ++i) std::swap(at(i, k), at(i, piv));
            std::swap(perm[(size_t)k], perm[(size_t)piv]);
        }
        const double dk = at(k, k);
        if (!(dk != 0.0) || !std::isfinite(dk)) return false;
        // trailing update A22 -= a21 a21^T / dk with the unscaled column, then l21 = a21 / dk
        std::vector<double> col((size_t)n, 0.0);
        for (int i = k + 1; i < n; ++i) col[(size_t)i] = at(i, k);
        for (int i = k + 1; i < n; ++i) {
            for (int j = k + 1; j <= i; ++j) at(i, j) -= col[(size_t)i] * col[(size_t)j] / dk;
            at(i, k) = col[(size_t)i] / dk;
        }
        for (int i = k + 1; i < n; ++i)
            for (int j = i + 1; j < n; ++j) at(i, j) = at(j, i);
    }
    std::vector<double> x((size_t)n);
    for (int c = 0; c < nb; ++c) {
        for (int i = 0; i < n; ++i) x[(size_t)i] = B[(size_t)perm[(size_t)i] * nb + c];
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < i; ++k) x[(size_t)i] -= at(i, k) * x[(size_t)k];
        for (int i = 0; i < n; ++i) x[(size_t)i] /= at(i, i);
        for (int i = n - 1; i >= 0; --i)
            for (int k = i + 1; k < n; ++k) x[(size_t)i] -= at(k, i) * x[(size_t)k];
        for (int i = 0; i < n; ++i) B[(size_t)perm[(size_t)i] * nb + c] = x[(size_t)i];
    }
    return true;
}

}  // namespace csamd

using namespace csamd;

struct cs_mvlmm {
    int64_t n = 0, c = 0, ne = 0, d = 0;
    rocblas_handle blas = nullptr;
    hipStream_t st = nullptr;
    DevBuf Q, tau, Y, G, D, E, t, Uty, W, info, P, work, sums;   // Q: Householder reflectors of X (n x c)
    std::vector<double> tau_h;              // their scalars on the host
    std::vector<double> Yh;                 // n x d host copy (Cov(Y) start values)
    std::vector<double> Vg, Ve, Vg_se, Ve_se;   // d x d row-major
    bool prepared = false, have_null = false, have_se = false;
    int em_iter = 0, ai_iter = 0;
    double logl = NAN;
    ~cs_mvlmm() {
        if (blas) (void)rocblas_destroy_handle(blas);
        if (st) (void)hipStreamDestroy(st);
    }
};

namespace {

template <class F>
void dispatch_d(int64_t d, F&& f) {
    switch (d) {
        case 1: f(std::integral_constant<int, 1>{}); break;
        case 2: f(std::integral_constant<int, 2>{}); break;
        case 3: f(std::integral_constant<int, 3>{}); break;
        case 4: f(std::integral_constant<int, 4>{}); break;
        case 5: f(std::integral_constant<int, 5>{}); break;
        case 6: f(std::integral_constant<int, 6>{}); break;
        default: throw InvalidArgument("multivariate LMM: 1 to 6 traits are supported");
    }
}

// Column sums over the ne eigen-samples of a per-sample kernel's output (K columns).
std::vector<double> mv_sums(cs_mvlmm* h, int64_t K) {
    h->sums.ensure((size_t)K * 8);
    k_colsum_rows<<<(unsigned)K, 256, 0, h->st>>>(h->work.as<double>(), h->ne, K, h->sums.as<double>());
    CS_HIP(hipGetLastError());
    std::vector<double> out((size_t)K);
    CS_HIP(hipMemcpyAsync(out.data(), h->sums.p, (size_t)K * 8, hipMemcpyDeviceToHost, h->st));
    CS_HIP(hipStreamSynchronize(h->st));
    return out;
}

void upload_params(cs_mvlmm* h, const std::vector<double>& Vg, const std::vector<double>& Ve) {
    const size_t dd = (size_t)(h->d * h->d);
    std::vector<double> P(2 * dd);
    std::copy(Vg.begin(), Vg.end(), P.begin());
    std::copy(Ve.begin(), Ve.end(), P.begin() + dd);
    h->P.ensure(P.size() * 8);
    CS_HIP(hipMemcpyAsync(h->P.p, P.data(), P.size() * 8, hipMemcpyHostToDevice, h->st));
}

void em_update(cs_mvlmm* h, std::vector<double>& Vg_new, std::vector<double>& Ve_new) {
    const int64_t d = h->d, ne = h->ne, K = 2 * d * d;
    upload_params(h, h->Vg, h->Ve);
    h->work.ensure((size_t)(ne * K) * 8);
    dispatch_d(d, [&](auto DC) {
        constexpr int DD = decltype(DC)::value;
        k_mv_em<DD><<<(unsigned)((ne + 255) / 256), 256, 0, h->st>>>(h->D.as<double>(), h->Uty.as<double>(), ne,
                                                                   h->P.as<double>(), h->work.as<double>());
    });
    CS_HIP(hipGetLastError());
    const std::vector<double> s = mv_sums(h, K);
    Vg_new.assign(s.begin(), s.begin() + d * d);
    Ve_new.assign(s.begin() + d * d, s.end());
    for (double& x : Vg_new) x /= (double)ne;
    for (double& x : Ve_new) x /= (double)ne;
}

double logl_at(cs_mvlmm* h, const std::vector<double>& Vg, const std::vector<double>& Ve) {
    upload_params(h, Vg, Ve);
    h->work.ensure((size_t)h->ne * 8);
    dispatch_d(h->d, [&](auto DC) {
        constexpr int DD = decltype(DC)::value;
        k_mv_logl<DD><<<(unsigned)((h->ne + 255) / 256), 256, 0, h->st>>>(
            h->D.as<double>(), h->Uty.as<double>(), h->ne, h->P.as<double>(), h->work.as<double>());
    });
    CS_HIP(hipGetLastError());
    const double s = mv_sums(h, 1)[0];
    return std::isfinite(s) ? s : -INFINITY;
}

// AI matrix (n_params x n_params, row-major) and score at (Vg, Ve); false when a V_i is not PD.
bool ai_build(cs_mvlmm* h, std::vector<double>& AI, std::vector<double>& score) {
    const int d = (int)h->d, vs = d * (d + 1) / 2, np = 2 * vs;
    const int64_t K = 2 * vs + 3 * vs * (vs + 1) / 2;
    upload_params(h, h->Vg, h->Ve);
    h->work.ensure((size_t)(h->ne * K) * 8);
    dispatch_d(d, [&](auto DC) {
        constexpr int DD = decltype(DC)::value;
        k_mv_ai<DD><<<(unsigned)((h->ne + 255) / 256), 256, 0, h->st>>>(
            h->D.as<double>(), h->Uty.as<double>(), h->ne, h->P.as<double>(), h->work.as<double>());
    });
    CS_HIP(hipGetLastError());
    const std::vector<double> s = mv_sums(h, K);
    for (double x : s)
        if (std::isnan(x)) return false;
    score.assign(s.begin(), s.begin() + np);
    AI.assign((size_t)np * np, 0.0);
    int pidx = 0;
    for (int a = 0; a < vs; ++a)
        for (int b = a; b < vs; ++b) {
            const double* q = &s[(size_t)(2 * vs + 3 * pidx)];
            AI[(size_t)a * np + b] += q[0];
            AI[(size_t)a * np + vs + b] += q[1];
            AI[(size_t)(vs + a) * np + b] += q[1];
            AI[(size_t)(vs + a) * np + vs + b] += q[2];
            ++pidx;
        }
    std::vector<double> sym(AI.size());
    for (int i = 0; i < np; ++i)
        for (int j = 0; j < np; ++j)
            sym[(size_t)i * np + j] = 0.5 * (AI[(size_t)i * np + j] + AI[(size_t)j * np + i]);
    AI = sym;
    return true;
}

bool run_em(cs_mvlmm* h, int max_iter, double tol) {
    const int64_t d = h->d;
    std::vector<double> Vg_new, Ve_new;
    for (int iter = 0; iter < max_iter; ++iter) {
        em_update(h, Vg_new, Ve_new);
        double err = 0.0;
        for (size_t k = 0; k < Vg_new.size(); ++k) {
            err += (Vg_new[k] - h->Vg[k]) * (Vg_new[k] - h->Vg[k]);
            err += (Ve_new[k] - h->Ve[k]) * (Ve_new[k] - h->Ve[k]);
        }
        err /= std::max(1.0, 2.0 * (double)(d * d));
        h->Vg = Vg_new;
        h->Ve = Ve_new;
        h->em_iter = iter + 1;
        if (err < tol) return true;
    }
    return false;
}

bool run_aireml(cs_mvlmm* h, int max_iter, double tol, int em_min_iter, int ai_max_iter, double ai_tol) {
    const int d = (int)h->d, vs = d * (d + 1) / 2, np = 2 * vs;
    std::vector<std::pair<int, int>> pairs;
    for (int j = 0; j < d; ++j)
        for (int k = j; k < d; ++k) pairs.emplace_back(j, k);
    // phase 1: EM
    std::vector<double> Vg_new, Ve_new;
    int em_iter = 0;
    for (; em_iter < max_iter; ++em_iter) {
        em_update(h, Vg_new, Ve_new);
        double err = 0.0;
        for (size_t k = 0; k < Vg_new.size(); ++k) {
            err += (Vg_new[k] - h->Vg[k]) * (Vg_new[k] - h->Vg[k]);
            err += (Ve_new[k] - h->Ve[k]) * (Ve_new[k] - h->Ve[k]);
        }
        err /= std::max(1.0, 2.0 * (double)(d * d));
        h->Vg = Vg_new;
        h->Ve = Ve_new;
        if (err < tol && em_iter >= em_min_iter) break;
    }
    h->em_iter = std::min(em_iter + 1, max_iter);
    near_pd(h->Vg, d);
    near_pd(h->Ve, d);
    // phase 2: AI-REML with step halving
    double logl_old = logl_at(h, h->Vg, h->Ve);
    const double logl_initial = logl_old;
    const std::vector<double> Vg_initial = h->Vg, Ve_initial = h->Ve;
    h->ai_iter = 0;
    if (!std::isfinite(logl_old)) return true;
    const double reg = 1e-8;
    std::vector<double> AI, score;
    for (int iter = 0; iter < ai_max_iter; ++iter) {
        if (!ai_build(h, AI, score)) return false;
        for (int i = 0; i < np; ++i) AI[(size_t)i * np + i] *= (1.0 + reg);
        std::vector<double> delta = score;
        if (!ldlt_solve(AI, np, delta, 1)) break;
        double step = 1.0;
        bool accepted = false;
        std::vector<double> Vg_t, Ve_t;
        for (int hh = 0; hh < 12; ++hh) {
            Vg_t = h->Vg;
            Ve_t = h->Ve;
            for (int a = 0; a < vs; ++a) {
                const int j = pairs[(size_t)a].first, k = pairs[(size_t)a].second;
                Vg_t[(size_t)j * d + k] += step * delta[(size_t)a];
                Vg_t[(size_t)k * d + j] = Vg_t[(size_t)j * d + k];
                Ve_t[(size_t)j * d + k] += step * delta[(size_t)(a + vs)];
                Ve_t[(size_t)k * d + j] = Ve_t[(size_t)j * d + k];
            }
            near_pd(Vg_t, d);
            near_pd(Ve_t, d);
            const double ln = logl_at(h, Vg_t, Ve_t);
            if (std::isfinite(ln) && ln >= logl_old - 1e-10) {
                accepted = true;
                break;
            }
            step *= 0.5;
        }
        if (!accepted) break;
        h->Vg = Vg_t;
        h->Ve = Ve_t;
        h->ai_iter = iter + 1;
        const double logl_new = logl_at(h, h->Vg, h->Ve);
        double nrm = 0.0;
        for (double x : delta) nrm += x * x;
        const double err = std::sqrt(nrm) / std::sqrt((double)np);
        logl_old = logl_new;
        if (err < ai_tol) return true;
    }
    if (logl_old < logl_initial + 1e-8) {
        h->Vg = Vg_initial;
        h->Ve = Ve_initial;
    }
    // standard errors from the AI matrix at the final parameters
    h->Vg_se.assign((size_t)(d * d), 0.0);
    h->Ve_se.assign((size_t)(d * d), 0.0);
    if (ai_build(h, AI, score)) {
        for (int i = 0; i < np; ++i) AI[(size_t)i * np + i] *= (1.0 + 1e-8);
        std::vector<double> inv((size_t)np * np, 0.0);
        for (int i = 0; i < np; ++i) inv[(size_t)i * np + i] = 1.0;
        if (ldlt_solve(AI, np, inv, np)) {
            for (int a = 0; a < vs; ++a) {
                const int j = pairs[(size_t)a].first, k = pairs[(size_t)a].second;
                const double sg = std::sqrt(std::fabs(inv[(size_t)a * np + a]));
                const double se = std::sqrt(std::fabs(inv[(size_t)(vs + a) * np + vs + a]));
                h->Vg_se[(size_t)j * d + k] = h->Vg_se[(size_t)k * d + j] = sg;
                h->Ve_se[(size_t)j * d + k] = h->Ve_se[(size_t)k * d + j] = se;
            }
            h->have_se = true;
        }
    }
    return true;
}

}  // namespace

extern "C" {

int cs_mvlmm_create(const double* Y, int64_t ldy, int64_t n, int64_t d, const float* X, int64_t ldx, int64_t c,
                    cs_mvlmm** out) {
    return guarded([&] {
        if (!Y || !out || (c > 0 && !X)) throw InvalidArgument("cs_mvlmm_create: null argument");
        if (d < 1 || d > kMvMaxD) throw InvalidArgument("multivariate LMM: 1 to 6 traits are supported");
        if (n < 2 || c < 0 || c >= n || ldy < n || (c > 0 && ldx < n))
            throw RuntimeError("MvEigenVce::prepare: Dimension mismatch");
        auto h = std::make_unique<cs_mvlmm>();
        h->n = n;
        h->c = c;
        h->ne = n - c;
        h->d = d;
        CS_HIP(hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking));
        CS_MVBLAS(rocblas_create_handle(&h->blas));
        CS_MVBLAS(rocblas_set_stream(h->blas, h->st));
        h->info.alloc(sizeof(rocblas_int));
        h->Yh.resize((size_t)(n * d));
        for (int64_t t = 0; t < d; ++t)
            for (int64_t i = 0; i < n; ++i) h->Yh[(size_t)(t * n + i)] = Y[t * ldy + i];
        h->Y.alloc((size_t)(n * d) * 8);
        CS_HIP(hipMemcpyAsync(h->Y.p, h->Yh.data(), (size_t)(n * d) * 8, hipMemcpyHostToDevice, h->st));
        if (c > 0) {
            std::vector<double> Xd((size_t)(n * c));
            for (int64_t j = 0; j < c; ++j)
                for (int64_t i = 0; i < n; ++i) Xd[(size_t)(j * n + i)] = (double)X[j * ldx + i];
            h->Q.alloc((size_t)(n * c) * 8);
            h->tau.alloc((size_t)c * 8);
            CS_HIP(hipMemcpyAsync(h->Q.p, Xd.data(), Xd.size() * 8, hipMemcpyHostToDevice, h->st));
            CS_MVBLAS(rocsolver_dgeqrf(h->blas, (int)n, (int)c, h->Q.as<double>(), (int)n, h->tau.as<double>()));
            h->tau_h.resize((size_t)c);
            CS_HIP(hipMemcpyAsync(h->tau_h.data(), h->tau.p, (size_t)c * 8, hipMemcpyDeviceToHost, h->st));
        }
        CS_HIP(hipStreamSynchronize(h->st));
        *out = h.release();
    });
}

int cs_mvlmm_destroy(cs_mvlmm* h) {
    return guarded([&] { delete h; });
}

int cs_mvlmm_prepare_dense(cs_mvlmm* h, const double* G, int64_t ldg, int g_on_device) {
    return guarded([&] {
        if (!h || !G) throw InvalidArgument("cs_mvlmm_prepare_dense: null argument");
        const int64_t n = h->n, ne = h->ne, c = h->c, d = h->d;
        if (ldg < n) throw RuntimeError("MvEigenVce::prepare: Dimension mismatch");
        const double one = 1.0, zero = 0.0;
        // working copy of G (n x n, ld n): projected in place, then overwritten by the eigenvectors
        h->G.ensure((size_t)(n * n) * 8);
        double* Gw = h->G.as<double>();
        CS_HIP(hipMemcpy2DAsync(Gw, (size_t)n * 8, G, (size_t)ldg * 8, (size_t)n * 8, (size_t)n,
                                g_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, h->st));
        h->D.ensure((size_t)ne * 8);
        h->E.ensure((size_t)ne * 8);
        h->t.ensure((size_t)(n * d) * 8);
        h->Uty.ensure((size_t)(ne * d) * 8);
        h->W.ensure((size_t)(n * ne) * 8);
        CS_HIP(hipMemcpyAsync(h->t.p, h->Y.p, (size_t)(n * d) * 8, hipMemcpyDeviceToDevice, h->st));
        if (c > 0) {
            // A = Q[:, c:] for Q = H_1 ... H_c, the Householder reflectors of X: A^T G A in place
            householder_project_sym(h->blas, h->st, h->Q.as<double>(), h->tau_h.data(), n, c, Gw);
            // t = Q^T Y (rows c: are A^T Y)
            CS_MVBLAS(rocsolver_dormqr(h->blas, rocblas_side_left, rocblas_operation_transpose, (int)n, (int)d,
                                       (int)c, h->Q.as<double>(), (int)n, h->tau.as<double>(), h->t.as<double>(),
                                       (int)n));
        }
        // eigendecomposition of A^T G A (ne x ne at offset (c, c), ld n); U overwrites it
        double* U = Gw + c * n + c;
        const bool large = eig_use_large(ne);
        if (large) {
            sym_eig_large(h->blas, h->st, ne, U, n, h->D.as<double>(), eig_large_leaf());
            CS_HIP(hipMemsetAsync(h->info.p, 0, sizeof(rocblas_int), h->st));
        } else {
            CS_MVBLAS(rocsolver_dsyevd(h->blas, rocblas_evect_original, rocblas_fill_lower, (int)ne, U, (int)n,
                                       h->D.as<double>(), h->E.as<double>(), h->info.as<rocblas_int>()));
        }
        std::vector<double> Dh((size_t)ne);
        CS_HIP(hipMemcpyAsync(Dh.data(), h->D.p, (size_t)ne * 8, hipMemcpyDeviceToHost, h->st));
        CS_HIP(hipStreamSynchronize(h->st));
        for (double& x : Dh)
            if (x < 0) x = 0;
        CS_HIP(hipMemcpyAsync(h->D.p, Dh.data(), (size_t)ne * 8, hipMemcpyHostToDevice, h->st));
        // Uty = U^T (A^T Y) (ne x d)
        CS_MVBLAS(rocblas_dgemm_64(h->blas, rocblas_operation_transpose, rocblas_operation_none, ne, d, ne, &one, U,
                                   n, h->t.as<double>() + c, n, &zero, h->Uty.as<double>(), ne));
        // W_proj = A U = Q [0; U] (n x ne)
        double* W = h->W.as<double>();
        if (c > 0) CS_HIP(hipMemset2DAsync(W, (size_t)n * 8, 0, (size_t)c * 8, (size_t)ne, h->st));
        CS_HIP(hipMemcpy2DAsync(W + c, (size_t)n * 8, U, (size_t)n * 8, (size_t)ne * 8, (size_t)ne,
                                hipMemcpyDeviceToDevice, h->st));
        if (c > 0 && large)
            apply_reflectors(h->blas, h->st, h->Q.as<double>(), n, h->tau_h.data(), n, c, false, W, n, ne);
        else if (c > 0)
            CS_MVBLAS(rocsolver_dormqr(h->blas, rocblas_side_left, rocblas_operation_none, (int)n, (int)ne, (int)c,
                                       h->Q.as<double>(), (int)n, h->tau.as<double>(), W, (int)n));
        rocblas_int info = 0;
        CS_HIP(hipMemcpyAsync(&info, h->info.p, sizeof(info), hipMemcpyDeviceToHost, h->st));
        CS_HIP(hipStreamSynchronize(h->st));
        h->G.release();
        if (info != 0) throw RuntimeError("MvEigenVce::prepare: eigendecomposition did not converge");
        // start values (eigen_vce.cpp:541-545): Cov(Y) with the divisor of the projected size
        const double nn = (double)(c > 0 ? ne : n);
        std::vector<double> mu((size_t)d, 0.0);
        for (int64_t a = 0; a < d; ++a) {
            for (int64_t i = 0; i < n; ++i) mu[(size_t)a] += h->Yh[(size_t)(a * n + i)];
            mu[(size_t)a] /= (double)n;
        }
        h->Vg.assign((size_t)(d * d), 0.0);
        for (int64_t a = 0; a < d; ++a)
            for (int64_t b = 0; b < d; ++b) {
                double s = 0.0;
                for (int64_t i = 0; i < n; ++i)
                    s += (h->Yh[(size_t)(a * n + i)] - mu[(size_t)a]) * (h->Yh[(size_t)(b * n + i)] - mu[(size_t)b]);
                h->Vg[(size_t)(a * d + b)] = s / (nn - 1.0) * 0.5;
            }
        h->Ve = h->Vg;
        h->Vg_se.clear();
        h->Ve_se.clear();
        h->prepared = true;
        h->have_null = false;
        h->have_se = false;
    });
}

int cs_mvlmm_null(cs_mvlmm* h, int solver, int max_iter, double tol, int em_min_iter, int ai_max_iter,
                  double ai_tol, int* converged, int* em_iterations, int* ai_iterations) {
    return guarded([&] {
        if (!h) throw InvalidArgument("cs_mvlmm_null: null handle");
        if (!h->prepared) throw RuntimeError("null model: prepare() has not run");
        h->have_se = false;
        const bool ok = solver == 1 ? run_aireml(h, max_iter, tol, em_min_iter, ai_max_iter, ai_tol)
                                    : run_em(h, max_iter, tol);
        if (solver != 1) h->ai_iter = 0;
        h->logl = logl_at(h, h->Vg, h->Ve);
        h->have_null = true;
        if (converged) *converged = ok ? 1 : 0;
        if (em_iterations) *em_iterations = h->em_iter;
        if (ai_iterations) *ai_iterations = h->ai_iter;
    });
}

int cs_mvlmm_get(cs_mvlmm* h, double* Vg, double* Ve, double* Vg_se, double* Ve_se, double* logl) {
    return guarded([&] {
        if (!h) throw InvalidArgument("cs_mvlmm_get: null handle");
        if (!h->prepared) throw RuntimeError("null model: prepare() has not run");
        const size_t dd = (size_t)(h->d * h->d);
        if (Vg) std::copy(h->Vg.begin(), h->Vg.end(), Vg);
        if (Ve) std::copy(h->Ve.begin(), h->Ve.end(), Ve);
        if (Vg_se) {
            if (h->have_se) std::copy(h->Vg_se.begin(), h->Vg_se.end(), Vg_se);
            else std::fill(Vg_se, Vg_se + dd, NAN);
        }
        if (Ve_se) {
            if (h->have_se) std::copy(h->Ve_se.begin(), h->Ve_se.end(), Ve_se);
            else std::fill(Ve_se, Ve_se + dd, NAN);
        }
        if (logl) *logl = h->logl;
    });
}

int cs_mvlmm_set_variance(cs_mvlmm* h, const double* Vg, const double* Ve) {
    return guarded([&] {
        if (!h || !Vg || !Ve) throw InvalidArgument("cs_mvlmm_set_variance: null argument");
        if (!h->prepared) throw RuntimeError("null model: prepare() has not run");
        const size_t dd = (size_t)(h->d * h->d);
        h->Vg.assign(Vg, Vg + dd);
        h->Ve.assign(Ve, Ve + dd);
        h->have_null = true;
        h->have_se = false;
    });
}

int cs_mvlmm_get_eigen(cs_mvlmm* h, double* D, double* Uty, double* W_proj, int64_t ldw) {
    return guarded([&] {
        if (!h) throw InvalidArgument("cs_mvlmm_get_eigen: null handle");
        if (!h->prepared) throw RuntimeError("null model: prepare() has not run");
        if (D) CS_HIP(hipMemcpyAsync(D, h->D.p, (size_t)h->ne * 8, hipMemcpyDeviceToHost, h->st));
        if (Uty) CS_HIP(hipMemcpyAsync(Uty, h->Uty.p, (size_t)(h->ne * h->d) * 8, hipMemcpyDeviceToHost, h->st));
        if (W_proj) {
            if (ldw < h->n) throw RuntimeError("cs_mvlmm_get_eigen: dimension mismatch");
            CS_HIP(hipMemcpy2DAsync(W_proj, (size_t)ldw * 8, h->W.p, (size_t)h->n * 8, (size_t)h->n * 8,
                                    (size_t)h->ne, hipMemcpyDeviceToHost, h->st));
        }
        CS_HIP(hipStreamSynchronize(h->st));
    });
}

int cs_mvlmm_scan(cs_mvlmm* h, cs_bed* bed, const int32_t* keep, int64_t n_keep, const int32_t* snps,
                  int64_t n_snps, double* freq, double* beta, double* vbeta, double* wald, double* p,
                  int32_t* status) {
    return guarded([&] {
        if (!h || !bed) throw InvalidArgument("cs_mvlmm_scan: null argument");
        if (!h->have_null) throw RuntimeError("scan: the null model has not been fitted");
        const int64_t N = bed->N, M = bed->M, n = h->n, ne = h->ne, d = h->d;
        if ((keep && n_keep > 0 ? n_keep : N) != n) throw RuntimeError("MvLmmEngine::scan: Sample size mismatch");
        if (keep && n_keep > 0)
            for (int64_t r = 0; r < n; ++r)
                if (keep[r] < 0 || keep[r] >= N) throw RuntimeError("scan: sample index out of range");
        std::vector<int32_t> all;
        if (!snps || n_snps <= 0) {
            all.resize((size_t)M);
            for (int64_t j = 0; j < M; ++j) all[(size_t)j] = (int32_t)j;
            snps = all.data();
            n_snps = M;
        }
        std::vector<int32_t> sel((size_t)M, -1);
        int64_t lo = M, hi = -1;
        for (int64_t i = 0; i < n_snps; ++i) {
            if (snps[i] < 0 || snps[i] >= M) throw RuntimeError("scan: variant index out of range");
            sel[(size_t)snps[i]] = (int32_t)i;
            lo = std::min<int64_t>(lo, snps[i]);
            hi = std::max<int64_t>(hi, snps[i]);
        }
        ensure_stats(bed);
        hipStream_t st = h->st;
        const int nq = (int)(d * (d + 1) / 2), nw = nq + (int)d;
        const int64_t ne_pad = round_up(ne, 256), ncb = ne_pad / 128;
        // keep multiplicities, inverse lists, class counts over the keep list
        std::vector<int32_t> w((size_t)bed->Npad, 0), off((size_t)N + 1, 0), idx((size_t)n);
        for (int64_t r = 0; r < n; ++r) w[(size_t)(keep && n_keep > 0 ? keep[r] : r)] += 1;
        for (int64_t s = 0; s < N; ++s) off[(size_t)s + 1] = off[(size_t)s] + w[(size_t)s];
        {
            std::vector<int32_t> fill(off.begin(), off.end() - 1);
            for (int64_t r = 0; r < n; ++r) idx[(size_t)fill[(size_t)(keep && n_keep > 0 ? keep[r] : r)]++] = (int32_t)r;
        }
        DevBuf counts((size_t)M * 32);
        weighted_counts(bed, w.data(), (int64_t)w.size(), counts.as<int64_t>(), st);
        std::vector<int64_t> cnt((size_t)M * 4);
        CS_HIP(hipMemcpyAsync(cnt.data(), counts.p, cnt.size() * 8, hipMemcpyDeviceToHost, st));
        DevBuf doff(off.size() * 4), didx(idx.size() * 4);
        CS_HIP(hipMemcpyAsync(doff.p, off.data(), off.size() * 4, hipMemcpyHostToDevice, st));
        CS_HIP(hipMemcpyAsync(didx.p, idx.data(), idx.size() * 4, hipMemcpyHostToDevice, st));
        // W hi/lo over the full sample space and the per-eigen-sample weights
        DevBuf Bw((size_t)(bed->Npad * ne_pad) * 4);
        k_mv_wsplit<<<grid_for(bed->Npad * ne_pad), 256, 0, st>>>(h->W.as<double>(), n, ne, ne_pad,
                                                                 doff.as<int32_t>(), didx.as<int32_t>(), N,
                                                                 bed->Npad, Bw.as<_Float16>());
        CS_HIP(hipGetLastError());
        upload_params(h, h->Vg, h->Ve);
        DevBuf wt((size_t)(ne_pad * nw) * 8), dbad(sizeof(int));
        CS_HIP(hipMemsetAsync(dbad.p, 0, sizeof(int), st));
        dispatch_d(d, [&](auto DC) {
            constexpr int DD = decltype(DC)::value;
            k_mv_weights<DD><<<(unsigned)((ne_pad + 255) / 256), 256, 0, st>>>(h->D.as<double>(), h->Uty.as<double>(), ne, ne_pad,
                                                              h->P.as<double>(), wt.as<double>(), dbad.as<int>());
        });
        CS_HIP(hipGetLastError());
        int bad = 0;
        CS_HIP(hipMemcpyAsync(&bad, dbad.p, sizeof(int), hipMemcpyDeviceToHost, st));
        CS_HIP(hipStreamSynchronize(st));
        if (bad) throw RuntimeError("MvLmmEngine::scan: D_c Vg + Ve is not positive definite");
        // outputs
        DevBuf dsel((size_t)M * 4), dfreq((size_t)n_snps * 8), dbeta((size_t)(n_snps * d) * 8),
            dvb((size_t)(n_snps * d * d) * 8), dwald((size_t)n_snps * 8), dp((size_t)n_snps * 8),
            dst((size_t)n_snps * 4);
        CS_HIP(hipMemcpyAsync(dsel.p, sel.data(), (size_t)M * 4, hipMemcpyHostToDevice, st));
        // marker tables per SNP row: z = x - mean (keep subset, observed), missing -> 0
        const int64_t row_lo = lo / 256 * 256, row_hi = round_up(hi + 1, 256);
        const int64_t chunk = std::min<int64_t>(row_hi - row_lo, 8192);
        DevBuf dtab((size_t)chunk * 16), part((size_t)(chunk * ncb * nw) * 8);
        std::vector<uint4> tab((size_t)chunk);
        auto h16 = [](double v) -> uint32_t {
            const _Float16 x = (_Float16)v;
            uint16_t u;
            std::memcpy(&u, &x, 2);
            return u;
        };
        for (int64_t r0 = row_lo; r0 < row_hi; r0 += chunk) {
            const int64_t rows = std::min<int64_t>(chunk, row_hi - r0);
            for (int64_t lr = 0; lr < rows; ++lr) {
                const int64_t j = r0 + lr;
                double mean = 0.0;
                if (j < M) {
                    const int64_t* c = &cnt[(size_t)j * 4];
                    const int64_t obs = c[0] + c[2] + c[3];
                    mean = obs > 0 ? (2.0 * (double)c[0] + (double)c[2]) / (double)obs : 0.0;
                }
                const double z[4] = {2.0 - mean, 0.0, 1.0 - mean, -mean};
                uint32_t zh[4], zl[4];
                for (int q = 0; q < 4; ++q) {
                    zh[q] = h16(z[q]);
                    _Float16 hh;
                    const uint16_t u = (uint16_t)zh[q];
                    std::memcpy(&hh, &u, 2);
                    zl[q] = h16(z[q] - (double)hh);
                }
                tab[(size_t)lr] = uint4{zh[0] | zh[1] << 16, zh[2] | zh[3] << 16, zl[0] | zl[1] << 16,
                                        zl[2] | zl[3] << 16};
            }
            CS_HIP(hipMemcpyAsync(dtab.p, tab.data(), (size_t)rows * 16, hipMemcpyHostToDevice, st));
            MvRotArgs a{};
            a.snp = bed->snp.as<uint8_t>();
            a.KBs = bed->KBs;
            a.row0 = r0;
            a.steps = bed->Npad / 32;
            a.tab = dtab.as<uint4>();
            a.Bw = Bw.as<_Float16>();
            a.ne_pad = ne_pad;
            a.wt = wt.as<double>();
            a.nq = nq;
            a.nw = nw;
            a.ncb = ncb;
            a.part = part.as<double>();
            k_mv_rot<<<dim3((unsigned)(rows / 256), (unsigned)(ne_pad / 256)), 512, 0, st>>>(a);
            CS_HIP(hipGetLastError());
            dispatch_d(d, [&](auto DC) {
                constexpr int DD = decltype(DC)::value;
                k_mv_final<DD><<<(unsigned)((rows + 255) / 256), 256, 0, st>>>(
                    part.as<double>(), rows, ncb, r0, dsel.as<int32_t>(), counts.as<int64_t>(), M,
                    dfreq.as<double>(), dbeta.as<double>(), dvb.as<double>(), dwald.as<double>(), dp.as<double>(),
                    dst.as<int32_t>());
            });
            CS_HIP(hipGetLastError());
            CS_HIP(hipStreamSynchronize(st));
        }
        auto d2h = [&](void* dst_, const DevBuf& src, size_t bytes) {
            if (dst_) CS_HIP(hipMemcpyAsync(dst_, src.p, bytes, hipMemcpyDeviceToHost, st));
        };
        d2h(freq, dfreq, (size_t)n_snps * 8);
        d2h(beta, dbeta, (size_t)(n_snps * d) * 8);
        d2h(vbeta, dvb, (size_t)(n_snps * d * d) * 8);
        d2h(wald, dwald, (size_t)n_snps * 8);
        d2h(p, dp, (size_t)n_snps * 8);
        d2h(status, dst, (size_t)n_snps * 4);
        CS_HIP(hipStreamSynchronize(st));
    });
}

}  // extern "C"
