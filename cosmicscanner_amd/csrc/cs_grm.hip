// Kernel (b), GRM build: the dense centred GRM of GrmBuilder::compute, algorithm 1 (VanRaden,
// libCosmicCore/src/grm_builder.cpp:63-399), computed exactly in integers.
//
// The reference centres each valid SNP by its float32-rounded mean mu_j = fl32(2 p_j)
// (grm_builder.cpp:240) and accumulates G += Z Z^T in FP64 (:345), then divides by the block-order
// FP64 sum of float32 2pq (:266-272, :392).  With SB the smallest shift >= 29 that makes every
// valid A_j = mu_j 2^SB an integer (29 whenever every mean is at least 2^-6; at most 40),
// z_ij = o_ij (x_ij - mu_j) is exactly U_ij 2^-SB with U_ij = o_ij (2^SB x~_ij - A_j), so
// G_raw = 2^-2SB I with the exact 128-bit integer
//   I_ik = 2^2SB T1 - 2^SB (R_i + R_k) + S + 2^SB (C_ik + C_ki) - Q_i - Q_k + D_ik
//   T1 = X~ X~^T,  R_i = sum_j A_j x~_ij,  S = sum_j A_j^2,  Q_i = sum_j A_j^2 m_ij,
//   C_ik = sum_j x~_ij A_j m_kj,  D_ik = sum_j A_j^2 m_ij m_kj   (m = missing indicator).
// T1 is one int8 MFMA product of packed genotypes with themselves.  The C terms come from
// products of one operand with itself: with A_j = sum_l 256^l d_lj in balanced signed 8-bit
// digits, E_l holds x~_ij where genotype (i, j) is observed and d_lj where it is missing, so
//   E_l E_l^T = T1 + C_l + C_l^T + D_l,   D_l,ik = sum_j d_lj^2 m_ij m_kj,
// and  sum_l 256^l E_l E_l^T = W T1 + (C + C^T) + sum_j w_j m_ij m_kj  with W = sum_l 256^l,
// w_j = sum_l 256^l d_lj^2.  Hence, with c = sum_l 256^l E_l E_l^T,
//   I_ik = (2^2SB - 2^SB W) T1 - 2^SB (R_i + R_k) + S + 2^SB c - Q_i - Q_k + D^v_ik,
//   D^v_ik = sum_j v_j m_ij m_kj,   v_j = A_j^2 - 2^SB w_j,
// where D^v, nonzero only for pairs of samples missing at the same SNP, is summed from the
// missing lists.  Without missing genotypes c and D^v vanish.  The output is the correctly
// rounded FP64 of G_raw and the correctly rounded FP32 of G_raw / sum_2pq (the GCTA .grm.bin
// value, grm_builder.cpp:1977-2026).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <thread>

#include "cs_common.h"
#include "cs_geno_mfma.h"
#include "cs_hwe.h"
#include "cs_ops.h"

namespace csamd {

typedef __int128 i128;
typedef unsigned __int128 u128;


constexpr uint32_t kLutMF = 0x0000FF00u;   // code 01 -> 0xFF

struct GrmArgs {
    const uint8_t* smp;    // masked tiled sample-major genotypes
    int64_t KB;            // 256-SNP blocks per row tile
    int64_t kb0, kb1;      // SNP-block range of this launch
    const int2* blocks;    // (bi, bk) 128-row block pairs, bk <= bi
    int64_t I0, P, N, Npad;
    const uint8_t* digA;   // digit arrays [l][kb][t][g][16] of the E passes (l = digit, stride dstride)
    int64_t dstride;
    int shift;             // 8 * (first digit of the pass)
    int32_t* t1;           // panel accumulators, [P][Npad]
    i128* c;
};


// Products of the packed sample-major genotypes with themselves on int8 MFMA, one 128 x 128
// lower-triangle block per workgroup; wave (wi, wk) owns the 64 x 64 quadrant rows
// i0 = 128 bi + 64 wi, columns k0 = 128 bk + 64 wk (4 x 4 MFMA tiles per step).  The block needs
// four 64-row operand tiles (A rows of bi, B rows of bk); wave w decodes tile w once per 64-k step
// into LDS and every wave reads its A and B operands from there.
//   DIG = false: T1 = X~ X~^T (one form, x~), stored as int32.
//   DIG = true: ND digits of c = sum_l 256^l E_l E_l^T, E_l = x~ where observed and the digit d_lj
//   where missing (the code extraction and the two perms shared by the digits, one and-or per
//   digit), added to the int128 panel.  Every term of E_l E_l^T is bounded by 4 (both observed),
//   256 (one missing) or 16384 (both missing), so a launch covers a SNP range whose bound stays in
//   int32 (chosen on the host from the largest per-sample missing count).
// LDS is double-buffered (one barrier per step).  Quadrants above the diagonal decode their tile
// but skip the MFMAs and the epilogue.  (Decoding step s + 1 beside the MFMAs of step s, in one
// basic block, was slower: 3.45 against 3.20 s at 50k x 500k.)
template <int ND, bool DIG>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DIG ? 2 : 3, DIG ? 2 : 3))) void
k_grm_xpass(GrmArgs a) {
    __shared__ v4i lds[2][ND][4][4][64];   // [buffer][form][tile][m-subtile][lane]
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = lane >> 4, r = lane & 15;
    const int2 blk = a.blocks[blockIdx.x];
    const int wi = wave >> 1, wk = wave & 1;
    const int64_t i0 = (int64_t)blk.x * 128 + 64 * wi;
    const int64_t k0 = (int64_t)blk.y * 128 + 64 * wk;
    const bool active = k0 < i0 + 64;
    const bool sideA = wave < 2;
    const int64_t trow = sideA ? (int64_t)blk.x * 128 + 64 * wave : (int64_t)blk.y * 128 + 64 * (wave - 2);
    const uint8_t* tbase = a.smp + (((trow >> 6) * a.KB) << 12) + lane * 16;
    v4i acc[ND][4][4];
#pragma unroll
    for (int l = 0; l < ND; ++l)
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int n = 0; n < 4; ++n) acc[l][m][n] = v4i{0, 0, 0, 0};
    // decode step (kb, t) of the packed words cw into LDS buffer bw
    auto decode = [&](const uint4 (&cw)[4], int64_t kb, int t, int bw) {
        uint4 d[ND];
        if (DIG) {
#pragma unroll
            for (int l = 0; l < ND; ++l)
                d[l] = *reinterpret_cast<const uint4*>(a.digA + l * a.dstride + (((kb * 4 + t) * 4 + g) << 4));
        }
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const uint32_t w = t == 0 ? cw[m].x : t == 1 ? cw[m].y : t == 2 ? cw[m].z : cw[m].w;
            const uint32_t s0 = w & 0x03030303u, s1 = (w >> 2) & 0x03030303u;
            const uint32_t s2 = (w >> 4) & 0x03030303u, s3 = (w >> 6) & 0x03030303u;
            const uint32_t x0 = __builtin_amdgcn_perm(0u, kLutX, s0), x1 = __builtin_amdgcn_perm(0u, kLutX, s1);
            const uint32_t x2 = __builtin_amdgcn_perm(0u, kLutX, s2), x3 = __builtin_amdgcn_perm(0u, kLutX, s3);
            if (!DIG) {
                lds[bw][0][wave][m][lane] = v4i{(int)x0, (int)x1, (int)x2, (int)x3};
            } else {
                const uint32_t f0 = __builtin_amdgcn_perm(0u, kLutMF, s0), f1 = __builtin_amdgcn_perm(0u, kLutMF, s1);
                const uint32_t f2 = __builtin_amdgcn_perm(0u, kLutMF, s2), f3 = __builtin_amdgcn_perm(0u, kLutMF, s3);
#pragma unroll
                for (int l = 0; l < ND; ++l)
                    lds[bw][l][wave][m][lane] = v4i{(int)((f0 & d[l].x) | x0), (int)((f1 & d[l].y) | x1),
                                                    (int)((f2 & d[l].z) | x2), (int)((f3 & d[l].w) | x3)};
            }
        }
    };
    uint4 cv[4], cn[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) cv[m] = *reinterpret_cast<const uint4*>(tbase + (a.kb0 << 12) + m * 1024);
    int buf = 0;
    for (int64_t kb = a.kb0; kb < a.kb1; ++kb) {
        const int64_t kn = kb + 1 < a.kb1 ? kb + 1 : kb;
#pragma unroll
        for (int m = 0; m < 4; ++m) cn[m] = *reinterpret_cast<const uint4*>(tbase + (kn << 12) + m * 1024);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            decode(cv, kb, t, buf);
            __syncthreads();
            if (active) {
#pragma unroll
                for (int l = 0; l < ND; ++l) {
                    v4i B[4];
#pragma unroll
                    for (int n = 0; n < 4; ++n) B[n] = lds[buf][l][2 + wk][n][lane];
#pragma unroll
                    for (int m = 0; m < 4; ++m) {
                        const v4i A = lds[buf][l][wi][m][lane];
#pragma unroll
                        for (int n = 0; n < 4; ++n)
                            acc[l][m][n] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A, B[n], acc[l][m][n], 0, 0, 0);
                    }
                }
            }
            buf ^= 1;
        }
#pragma unroll
        for (int m = 0; m < 4; ++m) cv[m] = cn[m];
    }
    if (!active) return;
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int n = 0; n < 4; ++n)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t i = i0 + 16 * m + 4 * g + v;
                const int64_t k = k0 + 16 * n + r;
                if (i >= a.N || k >= a.N || k > i || i < a.I0 || i >= a.I0 + a.P) continue;
                if (!DIG) {
                    a.t1[(i - a.I0) * a.Npad + k] = acc[0][m][n][v];
                } else {
                    i128 add = 0;
#pragma unroll
                    for (int l = 0; l < ND; ++l) add += (i128)acc[l][m][n][v] << (a.shift + 8 * l);
                    a.c[(i - a.I0) * a.Npad + k] += add;
                }
            }
}

// Invalid SNPs -> code 01 in a copy of the tiled sample-major matrix (keep mask: 2 bits per SNP,
// 11 = valid).
__global__ void k_mask_invalid(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n16,
                               int64_t KB, const uint32_t* __restrict__ keep) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n16;
         e += (int64_t)gridDim.x * blockDim.x) {
        // tile (rt, kb): 256 chunks of 16 B; chunk = (m*64 + g*16 + r) -> SNPs 256 kb + 64 g ..
        const int64_t tile = e >> 8;
        const int chunk = (int)(e & 255);
        const int64_t kb = tile % KB;
        const int g = (chunk >> 4) & 3;
        const int64_t q0 = (kb * 256 + g * 64) / 16;   // keep-mask dword of the chunk's first SNPs
        uint4 v = src[e];
        const uint32_t k0 = keep[q0], k1 = keep[q0 + 1], k2 = keep[q0 + 2], k3 = keep[q0 + 3];
        v.x = (v.x & k0) | (0x55555555u & ~k0);
        v.y = (v.y & k1) | (0x55555555u & ~k1);
        v.z = (v.z & k2) | (0x55555555u & ~k2);
        v.w = (v.w & k3) | (0x55555555u & ~k3);
        dst[e] = v;
    }
}

// R_i = sum_j A_j x~_ij (int64) and Q_i = sum_j A_j^2 m_ij (int128) per sample.
__global__ __launch_bounds__(256) void k_grm_rowsums(const uint8_t* __restrict__ smp, int64_t KB,
                                                     int64_t N, const long long* __restrict__ A,
                                                     long long* __restrict__ R, i128* __restrict__ Q) {
    const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
    const int64_t tr = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (tr * 64 >= N) return;
    const uint8_t* base = smp + ((tr * KB) << 12) + lane * 16;
    long long rs[4] = {0, 0, 0, 0};
    i128 qs[4] = {0, 0, 0, 0};
    for (int64_t kb = 0; kb < KB; ++kb) {
        const int64_t j0 = kb * 256 + g * 64;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const uint4 w4 = *reinterpret_cast<const uint4*>(base + (kb << 12) + m * 1024);
            const uint32_t ws[4] = {w4.x, w4.y, w4.z, w4.w};
            for (int q = 0; q < 4; ++q)
                for (int u = 0; u < 16; ++u) {
                    const long long aj = A[j0 + 16 * q + u];
                    const uint32_t code = (ws[q] >> (2 * u)) & 3u;
                    if (code == 1) qs[m] += (i128)aj * aj;
                    else rs[m] += aj * (code == 0 ? 2 : (code == 2 ? 1 : 0));
                }
        }
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        for (int off = 16; off < 64; off <<= 1) {
            rs[m] += __shfl_xor(rs[m], off);
            const unsigned long long lo = (unsigned long long)qs[m];
            const long long hi = (long long)(qs[m] >> 64);
            const unsigned long long lo2 = __shfl_xor(lo, off);
            const long long hi2 = __shfl_xor(hi, off);
            qs[m] += ((i128)hi2 << 64) + (i128)lo2;
        }
        const int64_t i = tr * 64 + 16 * m + r;
        if (g == 0 && i < N) {
            R[i] = rs[m];
            Q[i] = qs[m];
        }
    }
}

// D^v_ik = sum_j v_j m_ij m_kj for the panel rows i in [row0, row1) and k <= i, from the missing
// lists: one workgroup per row i walks the row's missing SNPs j and, for each, the samples k
// missing at j, adding v_j (split into three 42-bit limbs, the top one signed) to an LDS
// accumulator of kDCols columns.  Chunks of columns are taken in order; each thread holds kDPer of
// the row's SNPs (value limbs and list position) in registers from one chunk to the next, so
// every pair is visited once, and reads a SNP's list four entries at a time.  A chunk is stored to
// the three output planes [row][Npad] (added to them for rows whose list is longer than one batch
// of 256 kDPer SNPs).  Each limb sum stays below 2^63 for up to 2^21 SNPs.
constexpr int kDLimbBits = 42;
constexpr int kDCols = 2048;
constexpr int kDPer = 8;
constexpr int kDBatch = 256 * kDPer;

__global__ __launch_bounds__(256) void k_grm_dlds(const int64_t* __restrict__ soff, const int32_t* __restrict__ sidx,
                                                  const int64_t* __restrict__ off, const int32_t* __restrict__ idx,
                                                  const unsigned long long* __restrict__ v3, int64_t row0,
                                                  int64_t row1, int64_t I0, int64_t Npad,
                                                  unsigned long long* __restrict__ d0,
                                                  unsigned long long* __restrict__ d1,
                                                  unsigned long long* __restrict__ d2) {
    __shared__ unsigned long long acc[3][kDCols];
    const int64_t i = row0 + blockIdx.x;
    if (i >= row1) return;   // whole workgroup, before any barrier
    const int tid = threadIdx.x;
    const int64_t s0 = soff[i], s1 = soff[i + 1];
    const int64_t nch = i / kDCols + 1;
    unsigned long long* o0 = d0 + (i - I0) * Npad;
    unsigned long long* o1 = d1 + (i - I0) * Npad;
    unsigned long long* o2 = d2 + (i - I0) * Npad;
    for (int64_t b0 = s0;; b0 += kDBatch) {
        const bool first = b0 == s0;
        int64_t p[kDPer], pe[kDPer];
        unsigned long long v0[kDPer], v1[kDPer], v2[kDPer];
#pragma unroll
        for (int u = 0; u < kDPer; ++u) {
            const int64_t q = b0 + tid + 256 * u;
            p[u] = pe[u] = 0;
            v0[u] = v1[u] = v2[u] = 0ull;
            if (q < s1) {
                const int64_t j = sidx[q];
                v0[u] = v3[3 * j];
                v1[u] = v3[3 * j + 1];
                v2[u] = v3[3 * j + 2];
                if ((v0[u] | v1[u] | v2[u]) != 0ull) {
                    p[u] = off[j];
                    pe[u] = off[j + 1];
                }
            }
        }
        for (int64_t ch = 0; ch < nch; ++ch) {
            const int64_t c0 = ch * kDCols;
            const int64_t c1 = min<int64_t>(c0 + kDCols, i + 1);
            for (int e = tid; e < kDCols; e += 256) acc[0][e] = acc[1][e] = acc[2][e] = 0ull;
            __syncthreads();
#pragma unroll
            for (int u = 0; u < kDPer; ++u) {
                while (p[u] < pe[u]) {
                    int64_t kk[4];
#pragma unroll
                    for (int w = 0; w < 4; ++w) kk[w] = p[u] + w < pe[u] ? (int64_t)idx[p[u] + w] : INT64_MAX;
                    int taken = 0;
#pragma unroll
                    for (int w = 0; w < 4; ++w) {
                        if (taken == w && kk[w] < c1) {   // the list is sorted: stop at the first k >= c1
                            const int e = (int)(kk[w] - c0);
                            atomicAdd(&acc[0][e], v0[u]);
                            atomicAdd(&acc[1][e], v1[u]);
                            atomicAdd(&acc[2][e], v2[u]);
                            ++taken;
                        }
                    }
                    p[u] += taken;
                    if (taken < 4) break;
                }
            }
            __syncthreads();
            for (int e = tid; e < (int)(c1 - c0); e += 256) {
                if (first) {
                    o0[c0 + e] = acc[0][e];
                    o1[c0 + e] = acc[1][e];
                    o2[c0 + e] = acc[2][e];
                } else {
                    o0[c0 + e] += acc[0][e];
                    o1[c0 + e] += acc[1][e];
                    o2[c0 + e] += acc[2][e];
                }
            }
            __syncthreads();
        }
        if (b0 + kDBatch >= s1) break;
    }
}

__device__ __forceinline__ int top_bit(u128 a) {
    const unsigned long long hi = (unsigned long long)(a >> 64), lo = (unsigned long long)a;
    return hi ? 127 - __builtin_clzll(hi) : 63 - __builtin_clzll(lo);
}

// Correctly rounded double of an int128.
__device__ double i128_to_f64_rn(i128 v) {
    if (v == 0) return 0.0;
    const bool neg = v < 0;
    const u128 a = neg ? (u128)0 - (u128)v : (u128)v;
    const int top = top_bit(a);
    double res;
    if (top <= 52) {
        res = (double)(unsigned long long)a;
    } else {
        const int sh = top - 52;
        u128 q = a >> sh;
        const u128 rem = a & ((((u128)1) << sh) - 1);
        const u128 half = ((u128)1) << (sh - 1);
        if (rem > half || (rem == half && (q & 1))) q += 1;
        res = ldexp((double)(unsigned long long)q, sh);
    }
    return neg ? -res : res;
}

// Correctly rounded float of (v 2^-sc) / S with S = D 2^(se - 53), D a 53-bit integer.
__device__ float i128_div_f32_rn(i128 v, unsigned long long D, int se, int sc) {
    if (v == 0) return 0.0f;
    const bool neg = v < 0;
    const u128 a = neg ? (u128)0 - (u128)v : (u128)v;
    int sh = 27 + 53 - top_bit(a);
    if (sh < 0) sh = 0;
    const u128 num = a << sh;
    const u128 Q = num / D;
    const bool sticky = (num % D) != 0;
    const int drop = top_bit(Q) - 23;
    u128 keepq = Q >> drop;
    const u128 rem = Q & ((((u128)1) << drop) - 1);
    const u128 half = ((u128)1) << (drop - 1);
    if (rem > half || (rem == half && (sticky || (keepq & 1)))) keepq += 1;
    const int e2 = drop - sh - sc - (se - 53);
    const float res = ldexpf((float)(unsigned long long)keepq, e2);
    return neg ? -res : res;
}

struct CombineArgs {
    int64_t I0, P, N, Npad;
    int sb;               // U = z 2^sb
    const int32_t* t1;
    unsigned long long t1c_lo;       // coefficient of T1: 2^2sb - 2^sb W
    long long t1c_hi;
    const i128* c;                   // sum_l 256^l E_l E_l^T (null: no missing genotypes)
    const unsigned long long* d0;    // D^v limbs (null: no missing genotypes)
    const unsigned long long* d1;
    const unsigned long long* d2;
    const long long* R;
    const i128* Q;
    unsigned long long S_lo;
    long long S_hi;
    unsigned long long D53;
    int se;
    float* tri;           // lower triangle, GCTA order (i (i+1)/2 + k), device
    double* raw;          // optional full square N x N (raw G, correctly rounded)
};

__global__ void k_grm_combine(CombineArgs a) {
    const int64_t total = a.P * a.Npad;
    const i128 S = ((i128)a.S_hi << 64) + (i128)a.S_lo;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t li = e / a.Npad, k = e - li * a.Npad;
        const int64_t i = a.I0 + li;
        if (i >= a.N || k > i) continue;
        const i128 one = 1;
        const i128 t1c = ((i128)a.t1c_hi << 64) + (i128)a.t1c_lo;
        i128 I = (i128)a.t1[e] * t1c - ((i128)a.R[i] + (i128)a.R[k]) * (one << a.sb) + S - a.Q[i] - a.Q[k];
        if (a.c) I += a.c[e] * (one << a.sb);
        if (a.d0)
            I += (i128)(long long)a.d2[e] * (one << (2 * kDLimbBits)) + (i128)a.d1[e] * (one << kDLimbBits) +
                 (i128)a.d0[e];
        if (a.tri) a.tri[i * (i + 1) / 2 + k] = i128_div_f32_rn(I, a.D53, a.se, 2 * a.sb);
        if (a.raw) {
            const double g = ldexp(i128_to_f64_rn(I), -2 * a.sb);
            a.raw[i * a.N + k] = g;
            a.raw[k * a.N + i] = g;
        }
    }
}

// ------------------------------------------------------------- standardised GRM (algorithms 0, 3)
// GrmBuilder::compute with --make-grm-alg 0/3 (grm_builder.cpp:241-243, 300-304, 386-387): per
// valid SNP z = (x - mu_j) s_j with the float32 mean mu_j = fl32(2p) and the float32 scale
// s_j = fl32(1/sqrt(2p(1-p))), missing -> 0, G = Z Z^T / M.  The three non-zero values of each
// SNP are exact in FP64 and split into FP16 hi + lo (z = hi + lo + O(2^-22 z)); the product is
// Z_hi Z_hi^T + Z_hi Z_lo^T + Z_lo Z_hi^T on FP16 MFMA with FP32 accumulation, flushed into an
// FP64 accumulator every chunk of SNPs.
//
// One 256 x 256 lower-triangle block per workgroup of 8 waves.  Per 32-SNP step the block
// decodes its 512 sample rows (256 of the row block, 256 of the column block) from the packed
// sample-major tiles into LDS through per-SNP lookup tables (v_perm on the 2-bit codes), double
// buffered; wave (wr, wc) multiplies rows 64 wr .. +64 by columns 128 wc .. +128.
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kStdRow = 144;                 // bytes per decoded row: 32 hi + 32 lo FP16 + 16 pad
constexpr int kStdBuf = 512 * kStdRow;       // one buffer: 512 rows
constexpr int64_t kStdChunkSteps = 512;      // 32-SNP steps per FP32 accumulation chunk

struct StdArgs {
    const uint8_t* smp;    // tiled sample-major genotypes
    int64_t KB;            // 256-SNP blocks per row tile
    int64_t s0, s1;        // 32-SNP step range of this launch
    const int2* blocks;    // (bi, bk) 256-row block pairs, bk <= bi
    const uint4* tab;      // per SNP: {hi(c0)|hi(c1)<<16, hi(c2)|hi(c3)<<16, lo.., lo..}
    int64_t I0, P, N, Npad;
    double* acc;           // panel accumulator [P][Npad]
};

__device__ __forceinline__ uint32_t std_sel(uint32_t c, bool high) {
    // v_perm selector taking FP16 entry c of an 8-byte table into the low (or high) half, zero
    // elsewhere (selector byte 0x0c gives 0x00)
    return high ? c * 0x02020000u + 0x01000c0cu : c * 0x0202u + 0x0c0c0100u;
}

__device__ __forceinline__ void std_decode(const uint8_t* __restrict__ src, const uint4* tq, uint8_t* dst) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint32_t w = src[(i >> 1) * 1024 + (i & 1) * 128];
        const uint32_t s0 = std_sel(w & 3u, false), s1 = std_sel((w >> 2) & 3u, true);
        const uint32_t s2 = std_sel((w >> 4) & 3u, false), s3 = std_sel(w >> 6, true);
        uint2 hi, lo;
        hi.x = __builtin_amdgcn_perm(tq[0].y, tq[0].x, s0) | __builtin_amdgcn_perm(tq[1].y, tq[1].x, s1);
        hi.y = __builtin_amdgcn_perm(tq[2].y, tq[2].x, s2) | __builtin_amdgcn_perm(tq[3].y, tq[3].x, s3);
        lo.x = __builtin_amdgcn_perm(tq[0].w, tq[0].z, s0) | __builtin_amdgcn_perm(tq[1].w, tq[1].z, s1);
        lo.y = __builtin_amdgcn_perm(tq[2].w, tq[2].z, s2) | __builtin_amdgcn_perm(tq[3].w, tq[3].z, s3);
        uint8_t* row = dst + (i * 8) * kStdRow;
        *reinterpret_cast<uint2*>(row) = hi;
        *reinterpret_cast<uint2*>(row + 64) = lo;
    }
}

__global__ __launch_bounds__(512) void k_grm_std(StdArgs a) {
    __shared__ __align__(16) uint8_t lds[2 * kStdBuf];
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int2 blk = a.blocks[blockIdx.x];
    // decode role: byte column db (4 SNPs) of local rows wave*64 + drl + 8 i, i = 0..7
    const int db = lane & 7, drl = lane >> 3;
    const int64_t rbase = wave < 4 ? (int64_t)blk.x * 256 + wave * 64 : (int64_t)blk.y * 256 + (wave - 4) * 64;
    const uint8_t* tile = a.smp + (((rbase >> 6) * a.KB) << 12) + (drl << 4) + db;
    const int dst_off = (wave * 64 + drl) * kStdRow + 8 * db;
    // MFMA role
    const int wr = wave >> 1, wc = wave & 1, g = lane >> 4, r = lane & 15;
    const bool active = !(blk.x == blk.y && 128 * wc >= 64 * wr + 64);
    f32x4 acc[4][8];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int n = 0; n < 8; ++n) acc[m][n] = f32x4{0.f, 0.f, 0.f, 0.f};
    auto step_src = [&](int64_t s) {
        return tile + (((s >> 3)) << 12) + ((((s & 7) >> 1) * 16) << 4) + (s & 1) * 8;
    };
    uint4 tq[4];
    {
        const uint4* tp = a.tab + a.s0 * 32 + 4 * db;
#pragma unroll
        for (int q = 0; q < 4; ++q) tq[q] = tp[q];
        std_decode(step_src(a.s0), tq, lds + dst_off);
    }
    __syncthreads();
    const int aoff = (64 * wr + r) * kStdRow + 16 * g;
    const int boff = (256 + 128 * wc + r) * kStdRow + 16 * g;
    for (int64_t s = a.s0; s < a.s1; ++s) {
        const int cur = (int)((s - a.s0) & 1);
        const bool more = s + 1 < a.s1;
        if (more) {
            const uint4* tp = a.tab + (s + 1) * 32 + 4 * db;
#pragma unroll
            for (int q = 0; q < 4; ++q) tq[q] = tp[q];
        }
        if (active) {
            const uint8_t* buf = lds + cur * kStdBuf;
            f16x8 ah[4], al[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                ah[m] = *reinterpret_cast<const f16x8*>(buf + aoff + 16 * m * kStdRow);
                al[m] = *reinterpret_cast<const f16x8*>(buf + aoff + 16 * m * kStdRow + 64);
            }
#pragma unroll
            for (int n = 0; n < 8; ++n) {
                const f16x8 bh = *reinterpret_cast<const f16x8*>(buf + boff + 16 * n * kStdRow);
                const f16x8 bl = *reinterpret_cast<const f16x8*>(buf + boff + 16 * n * kStdRow + 64);
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[m], bh, acc[m][n], 0, 0, 0);
                    acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[m], bl, acc[m][n], 0, 0, 0);
                    acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[m], bh, acc[m][n], 0, 0, 0);
                }
            }
        }
        if (more) std_decode(step_src(s + 1), tq, lds + (cur ^ 1) * kStdBuf + dst_off);
        __syncthreads();
    }
    if (!active) return;
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int n = 0; n < 8; ++n)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t i = (int64_t)blk.x * 256 + 64 * wr + 16 * m + 4 * g + v;
                const int64_t k = (int64_t)blk.y * 256 + 128 * wc + 16 * n + r;
                if (i >= a.N || k > i || i < a.I0 || i >= a.I0 + a.P) continue;
                a.acc[(i - a.I0) * a.Npad + k] += (double)acc[m][n][v];
            }
}

// Panel output of the standardised GRM: raw Z Z^T (FP64, both triangles) and fl32(raw / M).
__global__ void k_grm_std_final(const double* __restrict__ acc, int64_t I0, int64_t P, int64_t N, int64_t Npad,
                                double M, float* __restrict__ tri, double* __restrict__ raw) {
    const int64_t total = P * Npad;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t li = e / Npad, k = e - li * Npad;
        const int64_t i = I0 + li;
        if (i >= N || k > i) continue;
        const double g = acc[e];
        if (tri) tri[i * (i + 1) / 2 + k] = (float)(g / M);
        if (raw) {
            raw[i * N + k] = g;
            raw[k * N + i] = g;
        }
    }
}

// ------------------------------------------------------------------------------------------ host
struct GrmParams {
    std::vector<uint8_t> valid;
    std::vector<long long> A;      // mu_j 2^sb (0 for invalid SNPs), Mpad entries
    int sb = 29;                   // smallest shift >= 29 with every valid mu_j 2^sb an integer
    double sum_2pq = 0.0;          // float 2pq summed in 2048-SNP block order (compute, :266-272)
    double sum_2pq_seq = 0.0;      // ... summed SNP by SNP (computeSplit's per-chromosome sums)
    int64_t n_valid = 0;
    int64_t kb_lo = 0, kb_hi = 0;  // 256-SNP blocks holding the valid SNPs
    bool any_missing = false;
};

// The optional HWE filter of GrmBuilder::compute (grm_builder.cpp:220-223): keep[j] = 0 when the
// exact HWE p-value of SNP j's all-sample genotype counts (code 00 = hom. A1, 10 = het., 11 =
// hom. A2) is below min_hwe; all ones when min_hwe <= 0.  Threads over the SNPs.
static std::vector<uint8_t> hwe_keep(const std::vector<int64_t>& cnt, int64_t M, float min_hwe) {
    std::vector<uint8_t> keep((size_t)M, 1);
    if (!(min_hwe > 0.0f)) return keep;
    const int nt = (int)std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            for (int64_t j = t; j < M; j += nt) {
                const int64_t* c = &cnt[(size_t)j * 4];
                if (hwe_exact_p(c[0], c[2], c[3]) < (double)min_hwe) keep[(size_t)j] = 0;
            }
        });
    for (auto& x : th) x.join();
    return keep;
}

// Per-SNP filter and float32 parameters of GrmBuilder::compute (grm_builder.cpp:167-272) from
// the all-sample class counts; `snps` (optional) restricts the SNP set, as a chromosome does in
// GrmBuilder::computeSplit (grm_builder.cpp:506-600, 684-713).
static GrmParams grm_params(cs_bed* b, float min_maf, float min_hwe, const int32_t* snps, int64_t n_snps) {
    GrmParams p;
    std::vector<int64_t> cnt((size_t)b->M * 4);
    CS_HIP(hipMemcpy(cnt.data(), b->counts.p, cnt.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
    const std::vector<uint8_t> hwe_ok = hwe_keep(cnt, b->M, min_hwe);
    p.valid.assign((size_t)b->Mpad, 0);
    p.A.assign((size_t)b->Mpad, 0);
    std::vector<uint8_t> sel;
    if (snps) {
        sel.assign((size_t)b->M, 0);
        for (int64_t i = 0; i < n_snps; ++i) {
            if (snps[i] < 0 || snps[i] >= b->M) throw RuntimeError("GRM: variant index out of range");
            sel[(size_t)snps[i]] = 1;
        }
    }
    std::vector<float> pq2((size_t)b->M, 0.0f), mean((size_t)b->M, 0.0f);
    for (int64_t j = 0; j < b->M; ++j) {
        if (snps && !sel[(size_t)j]) continue;
        const int64_t* c = &cnt[(size_t)j * 4];
        const long long n_obs = c[0] + c[2] + c[3];
        const long long sum_x = 2 * c[0] + c[2];
        if (n_obs == 0) continue;
        const double freq = (double)sum_x / (2.0 * (double)n_obs);
        const double maf = std::min(freq, 1.0 - freq);
        if (maf < (double)min_maf || std::fabs(maf) < 1e-8) continue;
        if (!hwe_ok[(size_t)j]) continue;
        pq2[(size_t)j] = (float)(1.0f * 2.0 * freq * (1.0 - freq));
        const float mean_f = (float)(freq * 2.0);
        mean[(size_t)j] = mean_f;
        while (p.sb <= 40 && std::ldexp((double)mean_f, p.sb) != std::floor(std::ldexp((double)mean_f, p.sb))) ++p.sb;
        if (p.sb > 40) throw InvalidArgument("GRM: a SNP mean needs more than 40 fractional bits");
        p.valid[(size_t)j] = 1;
        if (c[1] > 0) p.any_missing = true;
    }
    // A_j at the common shift; R_i = sum A_j x~_ij stays below 2^62 in int64
    long double asum = 0.0L;
    for (int64_t j = 0; j < b->M; ++j)
        if (p.valid[(size_t)j]) {
            p.A[(size_t)j] = (long long)std::ldexp((double)mean[(size_t)j], p.sb);
            asum += (long double)p.A[(size_t)j];
        }
    if (2.0L * asum >= std::ldexp(1.0L, 62)) throw InvalidArgument("GRM: too many SNPs for the exact row sums");
    double total = 0.0, seq = 0.0;
    int64_t lo = b->M, hi = -1;
    for (int64_t blk = 0; blk < b->M; blk += 2048) {
        double block = 0.0;
        for (int64_t j = blk; j < std::min<int64_t>(blk + 2048, b->M); ++j)
            if (p.valid[(size_t)j]) {
                block += pq2[(size_t)j];
                seq += pq2[(size_t)j];
                p.n_valid++;
                lo = std::min(lo, j);
                hi = std::max(hi, j);
            }
        total += block;
    }
    p.sum_2pq = total;
    p.sum_2pq_seq = seq;
    p.kb_lo = hi >= 0 ? lo / 256 : 0;
    p.kb_hi = hi >= 0 ? hi / 256 + 1 : 0;
    return p;
}

// Balanced signed 8-bit digits of a signed 128-bit value (little-endian), `nd` digits.
static void digits128(i128 v, int nd, int8_t* out) {
    for (int l = 0; l < nd; ++l) {
        int d;
        if (l == nd - 1) {
            d = (int)v;
        } else {
            d = (int)(((long long)(v & 255) + 128) % 256) - 128;
            if (d < -128) d += 256;
            v = (v - d) >> 8;
        }
        out[l] = (int8_t)d;
    }
}

}  // namespace csamd

struct GrmTiming {
    double ms_passes = 0, ms_combine = 0, ms_total = 0, mfma_ops = 0;
    int passes = 0;
};
static thread_local GrmTiming g_grm_timing;

using namespace csamd;

struct GrmOut {
    float* tri_host = nullptr;     // n(n+1)/2 GCTA lower triangle
    float* tri_dev = nullptr;
    double* raw_host = nullptr;    // n x n raw Z Z^T, symmetric
    double* raw_dev = nullptr;
    double sum_2pq = 0.0, sum_2pq_seq = 0.0;
    int64_t n_valid = 0;
};

// Per-SNP FP16 hi/lo lookup tables of the standardised GRM (grm_builder.cpp:167-257, 283-305):
// filter and float32 mean/scale as GrmBuilder::compute computes them for algorithms 0 and 3.
static std::vector<uint4> grm_std_tables(cs_bed* b, float min_maf, float min_hwe, const int32_t* snps,
                                         int64_t n_snps, GrmOut& out, int64_t& s_lo, int64_t& s_hi) {
    std::vector<int64_t> cnt((size_t)b->M * 4);
    CS_HIP(hipMemcpy(cnt.data(), b->counts.p, cnt.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
    const std::vector<uint8_t> hwe_ok = hwe_keep(cnt, b->M, min_hwe);
    std::vector<uint8_t> sel;
    if (snps) {
        sel.assign((size_t)b->M, 0);
        for (int64_t i = 0; i < n_snps; ++i) {
            if (snps[i] < 0 || snps[i] >= b->M) throw RuntimeError("GRM: variant index out of range");
            sel[(size_t)snps[i]] = 1;
        }
    }
    std::vector<uint4> tab((size_t)b->Mpad, uint4{0, 0, 0, 0});
    auto h16 = [](double v) -> uint32_t {
        const _Float16 h = (_Float16)v;
        uint16_t u;
        std::memcpy(&u, &h, 2);
        return u;
    };
    double total = 0.0, block = 0.0, seq = 0.0;
    int64_t nv = 0, lo = b->M, hi = -1;
    for (int64_t j = 0; j < b->M; ++j) {
        if (j % 2048 == 0) {
            total += block;
            block = 0.0;
        }
        if (snps && !sel[(size_t)j]) continue;
        const int64_t* c = &cnt[(size_t)j * 4];
        const long long n_obs = c[0] + c[2] + c[3];
        const long long sum_x = 2 * c[0] + c[2];
        if (n_obs == 0) continue;
        const double freq = (double)sum_x / (2.0 * (double)n_obs);
        const double maf = std::min(freq, 1.0 - freq);
        if (maf < (double)min_maf || std::fabs(maf) < 1e-8) continue;
        if (!hwe_ok[(size_t)j]) continue;
        const float pq2f = (float)(1.0f * 2.0 * freq * (1.0 - freq));
        block += pq2f;
        seq += pq2f;
        const double mean = (double)(float)(freq * 2.0);
        const double pq2 = 2.0 * freq * (1.0 - freq);
        const double scale = (double)(float)(pq2 > 1e-9 ? 1.0 / std::sqrt(pq2) : 0.0);
        const double z[4] = {(2.0 - mean) * scale, 0.0, (1.0 - mean) * scale, (0.0 - mean) * scale};
        uint32_t zh[4], zl[4];
        for (int q = 0; q < 4; ++q) {
            zh[q] = h16(z[q]);
            _Float16 hh;
            uint16_t u = (uint16_t)zh[q];
            std::memcpy(&hh, &u, 2);
            zl[q] = h16(z[q] - (double)hh);
        }
        tab[(size_t)j] = uint4{zh[0] | zh[1] << 16, zh[2] | zh[3] << 16, zl[0] | zl[1] << 16, zl[2] | zl[3] << 16};
        ++nv;
        lo = std::min(lo, j);
        hi = std::max(hi, j);
    }
    total += block;
    out.sum_2pq = total;
    out.sum_2pq_seq = seq;
    out.n_valid = nv;
    s_lo = hi >= 0 ? lo / 32 : 0;
    s_hi = hi >= 0 ? hi / 32 + 1 : 0;
    return tab;
}

static void grm_std_run(cs_bed* b, float min_maf, float min_hwe, const int32_t* snps, int64_t n_snps, GrmOut& out) {
    ensure_stats(b);
    ensure_sample_major(b);
    int64_t s_lo = 0, s_hi = 0;
    std::vector<uint4> tab = grm_std_tables(b, min_maf, min_hwe, snps, n_snps, out, s_lo, s_hi);
    const int64_t N = b->N, Npad = b->Npad, KB = b->KBm;
    hipStream_t st;
    CS_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    hipEvent_t e0, e1, e2;
    CS_HIP(hipEventCreate(&e0));
    CS_HIP(hipEventCreate(&e1));
    CS_HIP(hipEventCreate(&e2));
    try {
        DevBuf dtab(tab.size() * sizeof(uint4));
        CS_HIP(hipMemcpyAsync(dtab.p, tab.data(), dtab.bytes, hipMemcpyHostToDevice, st));
        DevBuf tri_own;
        float* tri = out.tri_dev;
        if (!tri && out.tri_host) {
            tri_own.alloc((size_t)(N * (N + 1) / 2) * sizeof(float));
            tri = tri_own.as<float>();
        }
        DevBuf raw_own;
        double* raw = out.raw_dev;
        if (!raw && out.raw_host) {
            raw_own.alloc((size_t)(N * N) * sizeof(double));
            raw = raw_own.as<double>();
        }
        size_t panel_bytes = 8ull << 30;
        if (const char* env = std::getenv("CSAMD_GRM_PANEL_BYTES")) panel_bytes = std::strtoull(env, nullptr, 10);
        int64_t P = std::max<int64_t>(256, (int64_t)(panel_bytes / ((size_t)Npad * 8)) / 256 * 256);
        P = std::min<int64_t>(P, Npad);
        DevBuf acc((size_t)(P * Npad) * 8);
        float ms_pass = 0.f, ms_fin = 0.f;
        int launches = 0;
        double ops = 0;
        const double Mv = (double)out.n_valid;
        for (int64_t I0 = 0; I0 < Npad; I0 += P) {
            const int64_t rows = std::min<int64_t>(P, Npad - I0);
            std::vector<int2> blocks;
            int64_t wave_tiles = 0;   // 64 x 128 wave tiles issuing MFMAs
            for (int64_t bi = I0 / 256; bi < (I0 + rows) / 256; ++bi)
                for (int64_t bk = 0; bk <= bi; ++bk) {
                    blocks.push_back(int2{(int)bi, (int)bk});
                    wave_tiles += bi == bk ? 6 : 8;
                }
            DevBuf dblocks(blocks.size() * sizeof(int2));
            CS_HIP(hipMemcpyAsync(dblocks.p, blocks.data(), dblocks.bytes, hipMemcpyHostToDevice, st));
            CS_HIP(hipMemsetAsync(acc.p, 0, (size_t)(rows * Npad) * 8, st));
            CS_HIP(hipEventRecord(e0, st));
            StdArgs sa{};
            sa.smp = b->smp.as<uint8_t>();
            sa.KB = KB;
            sa.blocks = dblocks.as<int2>();
            sa.tab = dtab.as<uint4>();
            sa.I0 = I0;
            sa.P = rows;
            sa.N = N;
            sa.Npad = Npad;
            sa.acc = acc.as<double>();
            for (int64_t c0 = s_lo; c0 < s_hi; c0 += kStdChunkSteps) {
                sa.s0 = c0;
                sa.s1 = std::min<int64_t>(s_hi, c0 + kStdChunkSteps);
                k_grm_std<<<(unsigned)blocks.size(), 512, 0, st>>>(sa);
                ++launches;
            }
            CS_HIP(hipGetLastError());
            CS_HIP(hipEventRecord(e1, st));
            ops += (double)wave_tiles * 3.0 * 64.0 * 128.0 * 2.0 * (double)((s_hi - s_lo) * 32);
            k_grm_std_final<<<grid_for(rows * Npad), 256, 0, st>>>(acc.as<double>(), I0, rows, N, Npad,
                                                                  Mv > 0 ? Mv : 1.0, tri, raw);
            CS_HIP(hipGetLastError());
            CS_HIP(hipEventRecord(e2, st));
            CS_HIP(hipEventSynchronize(e2));
            float a1 = 0, a2 = 0;
            CS_HIP(hipEventElapsedTime(&a1, e0, e1));
            CS_HIP(hipEventElapsedTime(&a2, e1, e2));
            ms_pass += a1;
            ms_fin += a2;
        }
        if (out.tri_host)
            CS_HIP(hipMemcpyAsync(out.tri_host, tri, (size_t)(N * (N + 1) / 2) * sizeof(float),
                                  hipMemcpyDeviceToHost, st));
        if (out.raw_host)
            CS_HIP(hipMemcpyAsync(out.raw_host, raw, (size_t)(N * N) * sizeof(double), hipMemcpyDeviceToHost, st));
        CS_HIP(hipStreamSynchronize(st));
        g_grm_timing.ms_passes = ms_pass;
        g_grm_timing.ms_combine = ms_fin;
        g_grm_timing.passes = launches;
        g_grm_timing.mfma_ops = ops;
    } catch (...) {
        (void)hipStreamSynchronize(st);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipEventDestroy(e2);
        (void)hipStreamDestroy(st);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    CS_HIP(hipStreamDestroy(st));
}

static void grm_run(cs_bed* b, int alg, float min_maf, float min_hwe, const int32_t* snps, int64_t n_snps,
                    GrmOut& out) {
    if (!(min_hwe >= 0.0f && min_hwe <= 1.0f)) throw InvalidArgument("GRM: min_hwe must be in [0, 1]");
    if (alg == 0 || alg == 3) {
        grm_std_run(b, min_maf, min_hwe, snps, n_snps, out);
        return;
    }
    if (alg != 1)
        throw InvalidArgument("GRM: algorithms 0/3 (standardised, Yang) and 1 (VanRaden) are built here");
    ensure_stats(b);
    ensure_sample_major(b);
    GrmParams prm = grm_params(b, min_maf, min_hwe, snps, n_snps);
    out.sum_2pq = prm.sum_2pq;
    out.sum_2pq_seq = prm.sum_2pq_seq;
    out.n_valid = prm.n_valid;
    float* out_tri_host = out.tri_host;
    float* out_tri_dev = out.tri_dev;
    double* out_raw_host = out.raw_host;
    const int64_t N = b->N, Npad = b->Npad, Mpad = b->Mpad, KB = b->KBm;
    if (Mpad > (int64_t(1) << 29))   // x~ x~ products reach 4 M in int32
        throw InvalidArgument("GRM: more than 2^29 SNPs");
    hipStream_t st;
    CS_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    hipEvent_t e0, e1, e2;
    CS_HIP(hipEventCreate(&e0));
    CS_HIP(hipEventCreate(&e1));
    CS_HIP(hipEventCreate(&e2));
    try {
        // keep mask and masked copy
        std::vector<uint32_t> keep((size_t)(Mpad / 16), 0);
        for (int64_t j = 0; j < b->M; ++j)
            if (prm.valid[(size_t)j]) keep[(size_t)(j / 16)] |= 3u << (2 * (j % 16));
        DevBuf dkeep(keep.size() * 4), masked((size_t)(Npad * Mpad / 4));
        CS_HIP(hipMemcpyAsync(dkeep.p, keep.data(), keep.size() * 4, hipMemcpyHostToDevice, st));
        const int64_t n16 = Npad * Mpad / 64;
        k_mask_invalid<<<grid_for(n16), 256, 0, st>>>(b->smp.as<uint4>(), masked.as<uint4>(), n16, KB,
                                                     dkeep.as<uint32_t>());
        CS_HIP(hipGetLastError());
        // A_j, digits, row sums, S
        DevBuf dA((size_t)Mpad * 8), dR((size_t)Npad * 8), dQ((size_t)Npad * 16);
        CS_HIP(hipMemcpyAsync(dA.p, prm.A.data(), (size_t)Mpad * 8, hipMemcpyHostToDevice, st));
        CS_HIP(hipMemsetAsync(dR.p, 0, dR.bytes, st));
        CS_HIP(hipMemsetAsync(dQ.p, 0, dQ.bytes, st));
        k_grm_rowsums<<<(unsigned)(Npad / 256), 256, 0, st>>>(masked.as<uint8_t>(), KB, N, dA.as<long long>(),
                                                             dR.as<long long>(), dQ.as<i128>());
        CS_HIP(hipGetLastError());
        i128 S = 0;
        for (long long a : prm.A) S += (i128)a * a;
        // balanced signed 8-bit digits of A_j for the C products (enough for the largest A_j), and
        // the per-SNP missing lists for D
        long long amax = 0;
        for (long long a : prm.A) amax = std::max(amax, a);
        int nC = 0;
        int64_t mmax = 0;   // most missing genotypes of one sample
        if (prm.any_missing) {
            // the fewest digits whose top digit stays in [-128, 127] for every A_j
            for (nC = 1; nC < 16; ++nC) {
                bool ok = true;
                for (long long a : prm.A) {
                    if (!a) continue;
                    int8_t dd[16];
                    digits128((i128)a, nC, dd);
                    i128 v = 0;
                    for (int l = nC - 1; l >= 0; --l) v = v * 256 + dd[l];
                    if (v != (i128)a) {
                        ok = false;
                        break;
                    }
                }
                if (ok) break;
            }
            ensure_missing_lists(b);
            if (b->M > ((int64_t)1 << 21)) throw InvalidArgument("GRM: more than 2^21 SNPs with missing genotypes");
            std::vector<int64_t> soff((size_t)N + 1);
            CS_HIP(hipMemcpy(soff.data(), b->miss_smp_off.p, soff.size() * 8, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < N; ++i) mmax = std::max(mmax, soff[(size_t)i + 1] - soff[(size_t)i]);
        }
        const bool useD = prm.any_missing;
        // digit arrays in the operand order [kb][t][g][16]: element k = 256 kb + 64 g + 16 t + 4 q + s
        // at byte 4 s + q; per SNP the pair value v_j = A_j^2 - 2^SB w_j (w_j = sum_l 256^l d_lj^2)
        // as three 42-bit limbs, the top one signed
        std::vector<uint8_t> dig((size_t)nC * Mpad, 0);
        std::vector<unsigned long long> v3(useD ? (size_t)Mpad * 3 : 0, 0ull);
        const i128 lmask = (((i128)1) << kDLimbBits) - 1;
        for (int64_t j = 0; j < Mpad; ++j) {
            if (!prm.A[(size_t)j]) continue;
            int8_t dd[16];
            const int64_t kb = j >> 8, rr = j & 255, g = rr >> 6, t = (rr >> 4) & 3, q = (rr >> 2) & 3, s = rr & 3;
            const int64_t off = (((kb * 4 + t) * 4 + g) << 4) + 4 * s + q;
            digits128((i128)prm.A[(size_t)j], nC, dd);
            i128 w = 0;
            for (int l = nC - 1; l >= 0; --l) w = w * 256 + (i128)((int)dd[l] * (int)dd[l]);
            for (int l = 0; l < nC; ++l) dig[(size_t)l * Mpad + off] = (uint8_t)dd[l];
            if (useD) {
                const i128 v = (i128)prm.A[(size_t)j] * (i128)prm.A[(size_t)j] - (w << prm.sb);
                v3[(size_t)j * 3] = (unsigned long long)(v & lmask);
                v3[(size_t)j * 3 + 1] = (unsigned long long)((v >> kDLimbBits) & lmask);
                v3[(size_t)j * 3 + 2] = (unsigned long long)(long long)(v >> (2 * kDLimbBits));
            }
        }
        DevBuf ddig(std::max<size_t>(1, dig.size())), dv3(std::max<size_t>(8, v3.size() * 8));
        if (!dig.empty()) CS_HIP(hipMemcpyAsync(ddig.p, dig.data(), dig.size(), hipMemcpyHostToDevice, st));
        if (!v3.empty()) CS_HIP(hipMemcpyAsync(dv3.p, v3.data(), v3.size() * 8, hipMemcpyHostToDevice, st));
        // coefficient of T1: 2^2SB - 2^SB W, W = sum_{l < nC} 256^l
        i128 Wsum = 0;
        for (int l = nC - 1; l >= 0; --l) Wsum = Wsum * 256 + 1;
        const i128 t1c = (((i128)1) << (2 * prm.sb)) - (Wsum << prm.sb);
        // E-pass SNP ranges: |E_l E_l^T| <= 4 L + 256 (m_i + m_k) + 16384 min(m_i, m_k) over a range
        // of L SNPs in which samples i and k miss m_i and m_k genotypes, so ranges of L SNPs with
        // 4 L + 16896 min(L, mmax) < 2^31 keep every int32 accumulator exact
        int64_t chunk_kb = prm.kb_hi - prm.kb_lo;
        if (useD) {
            while (chunk_kb > 1 && 4.0 * 256.0 * (double)chunk_kb +
                                       16896.0 * (double)std::min<int64_t>(256 * chunk_kb, mmax) >= 2147483647.0)
                chunk_kb = (chunk_kb + 1) / 2;
            if (4.0 * 256.0 + 16896.0 * 256.0 >= 2147483647.0) throw RuntimeError("GRM: E-pass bound");
        }
        if (const char* env = std::getenv("CSAMD_GRM_EPASS_KB")) chunk_kb = std::min<int64_t>(chunk_kb, std::atoll(env));
        chunk_kb = std::max<int64_t>(1, chunk_kb);
        // S2pq as D * 2^(se-53)
        int se = 0;
        const double fm = std::frexp(prm.sum_2pq > 0 ? prm.sum_2pq : 1.0, &se);
        const unsigned long long D53 = (unsigned long long)std::ldexp(fm, 53);
        // output
        DevBuf tri_own;
        float* tri = out_tri_dev;
        if (!tri && out_tri_host) {
            tri_own.alloc((size_t)(N * (N + 1) / 2) * sizeof(float));
            tri = tri_own.as<float>();
        }
        DevBuf raw_own;
        double* raw = out.raw_dev;
        if (!raw && out_raw_host) {
            raw_own.alloc((size_t)(N * N) * sizeof(double));
            raw = raw_own.as<double>();
        }
        // panels of P rows (multiple of 128) bounded to ~8 GB of accumulators
        // (CSAMD_GRM_PANEL_BYTES overrides the bound)
        size_t panel_bytes = 8ull << 30;
        if (const char* env = std::getenv("CSAMD_GRM_PANEL_BYTES")) panel_bytes = std::strtoull(env, nullptr, 10);
        const size_t per_row = (size_t)Npad * (4 + (nC ? 16 : 0) + (useD ? 24 : 0));
        int64_t P = std::max<int64_t>(128, (int64_t)(panel_bytes / per_row) / 128 * 128);
        P = std::min<int64_t>(P, round_up(Npad, 128));
        DevBuf t1((size_t)(P * Npad) * 4), cacc(nC ? (size_t)(P * Npad) * 16 : 1),
            dacc(useD ? (size_t)(P * Npad) * 24 : 1);
        float ms_pass = 0.f, ms_comb = 0.f;
        int launches = 0;
        double mfma_ops = 0;
        for (int64_t I0 = 0; I0 < Npad; I0 += P) {
            const int64_t rows = std::min<int64_t>(P, Npad - I0);
            std::vector<int2> blocks;
            for (int64_t bi = I0 / 128; bi < (I0 + rows) / 128; ++bi)
                for (int64_t bk = 0; bk <= bi; ++bk) blocks.push_back(int2{(int)bi, (int)bk});
            DevBuf dblocks(blocks.size() * sizeof(int2));
            CS_HIP(hipMemcpyAsync(dblocks.p, blocks.data(), blocks.size() * sizeof(int2),
                                  hipMemcpyHostToDevice, st));
            if (nC) CS_HIP(hipMemsetAsync(cacc.p, 0, (size_t)(rows * Npad) * 16, st));
            unsigned long long* dp0 = dacc.as<unsigned long long>();
            unsigned long long* dp1 = dp0 + rows * Npad;
            unsigned long long* dp2 = dp1 + rows * Npad;
            if (useD && I0 < N) {
                // D^v from the missing lists; every entry (i, k <= i) of the panel's real rows is
                // written.  (On a second stream beside the MFMA passes it took 0.44 s instead of
                // 0.11 s alone and the GRM 3.16 s in all: the two compete for the CUs.)
                const int64_t r1 = std::min<int64_t>(N, I0 + rows);
                k_grm_dlds<<<(unsigned)(r1 - I0), 256, 0, st>>>(
                    b->miss_smp_off.as<int64_t>(), b->miss_smp_idx.as<int32_t>(), b->miss_snp_off.as<int64_t>(),
                    b->miss_snp_idx.as<int32_t>(), dv3.as<unsigned long long>(), I0, r1, I0, Npad, dp0, dp1, dp2);
                CS_HIP(hipGetLastError());
            }
            GrmArgs ga{};
            ga.smp = masked.as<uint8_t>();
            ga.KB = KB;
            ga.blocks = dblocks.as<int2>();
            ga.I0 = I0;
            ga.P = rows;
            ga.N = N;
            ga.Npad = Npad;
            ga.t1 = t1.as<int32_t>();
            ga.c = cacc.as<i128>();
            ga.dstride = Mpad;
            const unsigned nb = (unsigned)blocks.size();
            int64_t quadrants = 0;   // 64 x 64 wave tiles on or below the diagonal
            for (const int2& bl : blocks) quadrants += bl.x == bl.y ? 3 : 4;
            CS_HIP(hipEventRecord(e0, st));
            ga.kb0 = prm.kb_lo;
            ga.kb1 = prm.kb_hi;
            k_grm_xpass<1, false><<<nb, 256, 0, st>>>(ga);
            ++launches;
            // E passes, two digits per launch, over SNP ranges that keep int32 exact
            for (int64_t c0 = prm.kb_lo; c0 < prm.kb_hi; c0 += chunk_kb) {
                GrmArgs gc = ga;
                gc.kb0 = c0;
                gc.kb1 = std::min<int64_t>(prm.kb_hi, c0 + chunk_kb);
                for (int l = 0; l < nC; l += 2) {
                    GrmArgs g2 = gc;
                    g2.shift = 8 * l;
                    g2.digA = ddig.as<uint8_t>() + (size_t)l * Mpad;
                    if (l + 1 < nC) k_grm_xpass<2, true><<<nb, 256, 0, st>>>(g2);
                    else k_grm_xpass<1, true><<<nb, 256, 0, st>>>(g2);
                    ++launches;
                }
            }
            CS_HIP(hipGetLastError());
            CS_HIP(hipEventRecord(e1, st));
            // int8 MFMA ops issued: T1 and each digit's E product over all SNP blocks
            mfma_ops += (double)(1 + nC) * (double)quadrants * 64.0 * 64.0 *
                        (double)((prm.kb_hi - prm.kb_lo) * 256) * 2.0;
            CombineArgs ca{};
            ca.I0 = I0;
            ca.P = rows;
            ca.N = N;
            ca.Npad = Npad;
            ca.t1 = t1.as<int32_t>();
            ca.t1c_lo = (unsigned long long)t1c;
            ca.t1c_hi = (long long)(t1c >> 64);
            ca.sb = prm.sb;
            ca.c = nC ? cacc.as<i128>() : nullptr;
            ca.d0 = useD ? dp0 : nullptr;
            ca.d1 = useD ? dp1 : nullptr;
            ca.d2 = useD ? dp2 : nullptr;
            ca.R = dR.as<long long>();
            ca.Q = dQ.as<i128>();
            ca.S_lo = (unsigned long long)S;
            ca.S_hi = (long long)(S >> 64);
            ca.D53 = D53;
            ca.se = se;
            ca.tri = tri;
            ca.raw = raw;
            k_grm_combine<<<grid_for(rows * Npad), 256, 0, st>>>(ca);
            CS_HIP(hipGetLastError());
            CS_HIP(hipEventRecord(e2, st));
            CS_HIP(hipEventSynchronize(e2));
            float a1 = 0, a2 = 0;
            CS_HIP(hipEventElapsedTime(&a1, e0, e1));
            CS_HIP(hipEventElapsedTime(&a2, e1, e2));
            ms_pass += a1;
            ms_comb += a2;
        }
        if (out_tri_host)
            CS_HIP(hipMemcpyAsync(out_tri_host, tri, (size_t)(N * (N + 1) / 2) * sizeof(float),
                                  hipMemcpyDeviceToHost, st));
        if (out_raw_host)
            CS_HIP(hipMemcpyAsync(out_raw_host, raw, (size_t)(N * N) * sizeof(double),
                                  hipMemcpyDeviceToHost, st));
        CS_HIP(hipStreamSynchronize(st));
        g_grm_timing.ms_passes = ms_pass;
        g_grm_timing.ms_combine = ms_comb;
        g_grm_timing.passes = launches;
        g_grm_timing.mfma_ops = mfma_ops;
    } catch (...) {
        (void)hipStreamSynchronize(st);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipEventDestroy(e2);
        (void)hipStreamDestroy(st);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    CS_HIP(hipStreamDestroy(st));
}

extern "C" {

int cs_grm_build(cs_bed* bed, int alg, float min_maf, float min_hwe, float* grm_tri_f32, double* grm_raw_f64,
                 double* sum_2pq, int64_t* n_valid) {
    return guarded([&] {
        if (!bed) throw InvalidArgument("cs_grm_build: null handle");
        GrmOut o;
        o.tri_host = grm_tri_f32;
        o.raw_host = grm_raw_f64;
        grm_run(bed, alg, min_maf, min_hwe, nullptr, 0, o);
        if (sum_2pq) *sum_2pq = o.sum_2pq;
        if (n_valid) *n_valid = o.n_valid;
    });
}

int cs_grm_build_device(cs_bed* bed, int alg, float min_maf, float min_hwe, float* d_grm_tri_f32,
                        double* sum_2pq, int64_t* n_valid) {
    return guarded([&] {
        if (!bed || !d_grm_tri_f32) throw InvalidArgument("cs_grm_build_device: null argument");
        GrmOut o;
        o.tri_dev = d_grm_tri_f32;
        grm_run(bed, alg, min_maf, min_hwe, nullptr, 0, o);
        if (sum_2pq) *sum_2pq = o.sum_2pq;
        if (n_valid) *n_valid = o.n_valid;
    });
}

int cs_grm_build_raw(cs_bed* bed, int alg, float min_maf, float min_hwe, const int32_t* snps, int64_t n_snps,
                     double* raw_f64, int raw_on_device, double* sum_2pq_block,
                     double* sum_2pq_seq, int64_t* n_valid) {
    return guarded([&] {
        if (!bed || !raw_f64) throw InvalidArgument("cs_grm_build_raw: null argument");
        GrmOut o;
        (raw_on_device ? o.raw_dev : o.raw_host) = raw_f64;
        grm_run(bed, alg, min_maf, min_hwe, (snps && n_snps > 0) ? snps : nullptr, n_snps, o);
        if (sum_2pq_block) *sum_2pq_block = o.sum_2pq;
        if (sum_2pq_seq) *sum_2pq_seq = o.sum_2pq_seq;
        if (n_valid) *n_valid = o.n_valid;
    });
}

int cs_grm_last_timing(double* ms_passes, double* ms_combine, int* passes, double* mfma_ops) {
    return guarded([&] {
        if (ms_passes) *ms_passes = g_grm_timing.ms_passes;
        if (ms_combine) *ms_combine = g_grm_timing.ms_combine;
        if (passes) *passes = g_grm_timing.passes;
        if (mfma_ops) *mfma_ops = g_grm_timing.mfma_ops;
    });
}

}  // extern "C"
