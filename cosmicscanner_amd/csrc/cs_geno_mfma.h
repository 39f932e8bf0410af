// Kernel (b) core: the packed-genotype x int8-digit product on CDNA4 MFMA (16x16x64 i8).
//
// out[split][row][type][col] = sum_k  op_type(code(row, k)) * B_type[k][col]
// with op_0 = x~ (dosage 2/1/0, 0 for missing) and op_1 = o (1 if observed, else 0), over a packed
// 2-bit matrix whose rows are contiguous along k (SNP-major: rows = SNPs, k = samples;
// sample-major: rows = samples, k = SNPs).  All arithmetic is exact int32.
//
// Operand layouts.  One wave owns 64 rows (4 M tiles of 16) and NTB N tiles of 16 columns.  Each
// lane (group g = lane/16, r = lane%16) loads 16 packed bytes = 64 k-elements of its row per 256-k
// block, and uses dword t of them at MFMA step t.  Inside a step the 16 operand bytes are ordered
// byte(4s + q) <-> k = 64g + 16t + 4q + s, which makes the 2-bit -> int8 unpack a shift + mask
// (d_s = (w >> 2s) & 0x03030303) plus one v_perm byte lookup per 4 genotypes.  The B digits are
// stored pre-permuted the same way: [kblock][t][ntile][lane][16 bytes] (bperm_offset below), so
// the same k pairs up in A and B whatever the hardware's k order inside an operand is.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace csamd {

typedef int v4i __attribute__((ext_vector_type(4)));

// byte c of the LUT = value for code c: x~ = {2, 0, 1, 0}, o = {1, 0, 1, 1}
constexpr uint32_t kLutX = 0x00010002u;
constexpr uint32_t kLutO = 0x01010001u;

__host__ __device__ inline int64_t bperm_offset(int64_t kidx, int col, int NT) {
    const int64_t kb = kidx >> 8;
    const int r = (int)(kidx & 255);
    const int g = r >> 6, t = (r >> 4) & 3, q = (r >> 2) & 3, s = r & 3;
    const int n = col >> 4, c = col & 15;
    return ((((kb * 4 + t) * NT + n) * 64 + g * 16 + c) << 4) + 4 * s + q;
}

template <int NTB, bool TWO_B>
__global__ __launch_bounds__(256) void k_geno_mfma(const uint8_t* __restrict__ A, int64_t lda,
                                                   int64_t rows_pad, int64_t kb_per_split,
                                                   int64_t kb_total, int NT,
                                                   const int8_t* __restrict__ Bx,
                                                   const int8_t* __restrict__ Bo,
                                                   int32_t* __restrict__ out) {
    constexpr int MT = 4;
    constexpr int NB = TWO_B ? 2 : 1;
    constexpr int TILE = 4 * NTB * 1024 * NB;   // B bytes per 256-k block (all steps, tiles)
    constexpr int PER_T = TILE / 4096;          // uint4 per thread to stage one tile
    __shared__ __attribute__((aligned(16))) int8_t sB[2][TILE];

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = lane >> 4, r = lane & 15;
    const int64_t row0 = ((int64_t)blockIdx.x * 4 + wave) * (16 * MT);
    const int split = blockIdx.y;
    const int nt0 = blockIdx.z * NTB;
    const int64_t kb0 = (int64_t)split * kb_per_split;
    int64_t kb1 = kb0 + kb_per_split;
    if (kb1 > kb_total) kb1 = kb_total;

    v4i acc[2 * MT * NTB];
#pragma unroll
    for (int i = 0; i < 2 * MT * NTB; ++i) acc[i] = v4i{0, 0, 0, 0};

    // Global B of one block: [t][ntile (all NT)][lane][16]; the staged tile keeps only this
    // block's NTB tiles: sB[(t*NTB + n)*1024 + lane*16] (+ 4*NTB*1024 for the o-operand).
    auto stage_load = [&](int64_t kb, uint4 (&reg)[PER_T]) {
#pragma unroll
        for (int i = 0; i < PER_T; ++i) {
            const int e = (i * 256 + threadIdx.x) * 16;         // byte in the staged tile
            const int which = e / (4 * NTB * 1024);              // 0: x operand, 1: o operand
            const int rem = e - which * (4 * NTB * 1024);
            const int t = rem / (NTB * 1024);
            const int rem2 = rem - t * (NTB * 1024);
            const int n = rem2 >> 10, l16 = rem2 & 1023;
            const int8_t* src = (which ? Bo : Bx) + (((kb * 4 + t) * NT + nt0 + n) << 10) + l16;
            reg[i] = *reinterpret_cast<const uint4*>(src);
        }
    };
    auto stage_store = [&](int buf, const uint4 (&reg)[PER_T]) {
#pragma unroll
        for (int i = 0; i < PER_T; ++i)
            *reinterpret_cast<uint4*>(&sB[buf][(i * 256 + threadIdx.x) * 16]) = reg[i];
    };

    const uint8_t* abase = A + (row0 + r) * lda + g * 16;
    const int64_t astride = 16 * lda;
    uint4 a[MT], an[MT];
    uint4 breg[PER_T];
    if (kb0 < kb1) {
#pragma unroll
        for (int m = 0; m < MT; ++m) a[m] = *reinterpret_cast<const uint4*>(abase + m * astride + kb0 * 64);
        stage_load(kb0, breg);
        stage_store(0, breg);
    }
    __syncthreads();
    int buf = 0;
    for (int64_t kb = kb0; kb < kb1; ++kb) {
        const bool more = kb + 1 < kb1;
        if (more) {
#pragma unroll
            for (int m = 0; m < MT; ++m)
                an[m] = *reinterpret_cast<const uint4*>(abase + m * astride + (kb + 1) * 64);
            stage_load(kb + 1, breg);
        }
        const int8_t* tile = sB[buf];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            v4i bx[NTB], bo[NTB];
#pragma unroll
            for (int n = 0; n < NTB; ++n) {
                bx[n] = *reinterpret_cast<const v4i*>(tile + ((t * NTB + n) << 10) + lane * 16);
                if (TWO_B)
                    bo[n] = *reinterpret_cast<const v4i*>(tile + 4 * NTB * 1024 +
                                                          ((t * NTB + n) << 10) + lane * 16);
            }
#pragma unroll
            for (int m = 0; m < MT; ++m) {
                const uint32_t w = t == 0 ? a[m].x : t == 1 ? a[m].y : t == 2 ? a[m].z : a[m].w;
                const uint32_t d0 = w & 0x03030303u, d1 = (w >> 2) & 0x03030303u;
                const uint32_t d2 = (w >> 4) & 0x03030303u, d3 = (w >> 6) & 0x03030303u;
                const v4i X = v4i{(int)__builtin_amdgcn_perm(0u, kLutX, d0),
                                  (int)__builtin_amdgcn_perm(0u, kLutX, d1),
                                  (int)__builtin_amdgcn_perm(0u, kLutX, d2),
                                  (int)__builtin_amdgcn_perm(0u, kLutX, d3)};
                const v4i O = v4i{(int)__builtin_amdgcn_perm(0u, kLutO, d0),
                                  (int)__builtin_amdgcn_perm(0u, kLutO, d1),
                                  (int)__builtin_amdgcn_perm(0u, kLutO, d2),
                                  (int)__builtin_amdgcn_perm(0u, kLutO, d3)};
#pragma unroll
                for (int n = 0; n < NTB; ++n) {
                    v4i& cx = acc[(m * NTB + n) * 2];
                    v4i& co = acc[(m * NTB + n) * 2 + 1];
                    cx = __builtin_amdgcn_mfma_i32_16x16x64_i8(X, bx[n], cx, 0, 0, 0);
                    co = __builtin_amdgcn_mfma_i32_16x16x64_i8(O, TWO_B ? bo[n] : bx[n], co, 0, 0, 0);
                }
            }
        }
        if (more) {
            stage_store(buf ^ 1, breg);
#pragma unroll
            for (int m = 0; m < MT; ++m) a[m] = an[m];
        }
        __syncthreads();
        buf ^= 1;
    }
    // C layout of 16x16 MFMA: lane holds column (lane % 16), rows 4*(lane/16) + v, v = 0..3.
    const int NC = NT * 16;
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int n = 0; n < NTB; ++n)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t row = row0 + 16 * m + 4 * g + v;
                const int col = (nt0 + n) * 16 + r;
                int32_t* o = out + (((int64_t)split * rows_pad + row) * 2) * NC + col;
                o[0] = acc[(m * NTB + n) * 2][v];
                o[NC] = acc[(m * NTB + n) * 2 + 1][v];
            }
}

}  // namespace csamd
