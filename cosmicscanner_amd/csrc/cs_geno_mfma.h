// Kernel (b) core: packed-genotype x int8-digit products on CDNA4 MFMA (v_mfma_i32_16x16x64_i8).
//
// For a packed 2-bit matrix whose rows are contiguous along k (SNP-major: rows = SNPs, k =
// samples; sample-major: rows = samples, k = SNPs) the kernel forms, exactly in int32,
//     X[row][c] = sum_k x~(code(row,k)) Bx[k][c],   O[row][c] = sum_k o(code(row,k)) Bo[k][c]
// with x~ = dosage of A1 (2/1/0, 0 for missing) and o = 1 if observed, for the L limbs of each
// right-hand-side column, then combines the limbs in int64 and applies a fused epilogue.
//
// HBM layout (both resident copies): "MFMA tiles".  Rows are grouped in tiles of 64 and k in
// blocks of 256 (64 bytes); tile (rt, kb) is one contiguous 4 KB piece at (rt*KB + kb)*4096, and
// inside it the byte of (row = 64 rt + 16 m + r, kbyte = 64 kb + 16 g + b) sits at
// (m*64 + g*16 + r)*16 + b.  A wave owns one row tile and streams its tiles sequentially, each
// lane loading 16 contiguous bytes per m-subtile, so every load instruction reads 1 KB of
// contiguous memory.
//
// Operand order.  Lane (g = lane/16, r = lane%16) of m-subtile m supplies 16 packed bytes = 64
// k-elements of row 16m + r and uses dword t of them at MFMA step t; inside a step the operand
// bytes are ordered byte(4s + q) <-> k = 256 kb + 64 g + 16 t + 4 q + s, so unpacking is a shift
// and mask (d_s = (w >> 2s) & 0x03030303) plus one v_perm byte lookup per 4 genotypes.  B digits
// are stored pre-permuted the same way ([kb][t][tile][lane][16 bytes], bperm_offset), so A and B
// pair the same k whatever the hardware's k order inside an operand.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>

namespace csamd {

typedef int v4i __attribute__((ext_vector_type(4)));
__device__ __forceinline__ v4i as_v4i(const v4i& x) { return x; }
__device__ __forceinline__ v4i as_v4i(const uint4& x) { return v4i{(int)x.x, (int)x.y, (int)x.z, (int)x.w}; }

// byte c of the LUT = value for code c: x~ = {2, 0, 1, 0}, o = {1, 0, 1, 1}
constexpr uint32_t kLutX = 0x00010002u;
constexpr uint32_t kLutO = 0x01010001u;

__host__ __device__ inline int64_t tiled_offset(int64_t row, int64_t kbyte, int64_t KB) {
    const int64_t rt = row >> 6;
    const int rr = (int)(row & 63), m = rr >> 4, r = rr & 15;
    const int64_t kb = kbyte >> 6;
    const int kk = (int)(kbyte & 63), g = kk >> 4, b = kk & 15;
    return ((rt * KB + kb) << 12) + ((m * 64 + g * 16 + r) << 4) + b;
}

// B operand byte of k-element kidx, tile n (of NT), column c (0..15) within the tile.
__host__ __device__ inline int64_t bperm_offset(int64_t kidx, int n, int c, int NT) {
    const int64_t kb = kidx >> 8;
    const int rr = (int)(kidx & 255);
    const int g = rr >> 6, t = (rr >> 4) & 3, q = (rr >> 2) & 3, s = rr & 3;
    return ((((kb * 4 + t) * NT + n) * 64 + g * 16 + c) << 4) + 4 * s + q;
}

enum { EPI_PQ = 0, EPI_KV_GT = 1, EPI_KV_G = 2 };

struct GenoArgs {
    const uint8_t* A;
    int64_t KB;                // k blocks per row tile
    int64_t rows;              // real rows (epilogue guard)
    int64_t kb_per_split, kb_total;
    int G;                     // groups of 16 right-hand sides (= gridDim.z)
    // flat_tiles > 0: a 1-D grid in XCD-aware order (one group per workgroup).  Workgroup b runs
    // on XCD b % 8; its tile T = 8 (s / G) + b % 8 and group s % G (s = b / 8), so the G groups
    // of one (row block, split) tile run back to back on one XCD and share the A tiles in its
    // L2.  T = y * tiles_x + x; workgroups with T >= flat_tiles exit at once.
    int64_t flat_tiles, tiles_x;
    const int8_t* Bx;
    const int8_t* Bo;          // EPI_KV_G with two operands: the digits of -(mu W)
    int64_t k;                 // real number of right-hand sides
    long long* pq;             // EPI_PQ: [row][2][16 G]
    const double* scale;       // EPI_KV_GT: per-column scale of the quantised V
    const double* means;
    const double* inv_stds;
    const int32_t* mult;       // variant multiplicities
    double inv_M;
    double* W1;                // EPI_KV_GT: [row][k] forward operands w and mu * w
    double* W2;
    unsigned long long* mx;    // EPI_KV_GT: per-column max(|w|, |mu w|) as double bits
    long long* yint;           // EPI_KV_G: [row][k] int64, accumulated atomically over splits
    const long long* qext;     // ONLY_X: O^T B per [row][16 G] (exact, from the missing lists)
    int out_u;                 // EPI_KV_GT: write u = (P - mu Q) s scale to W1 for every row with
                               // inv_std != 0 (no multiplicity, W2 and maxima untouched)
};

// ONLY_X: only the genotype product X~^T B runs on MFMA.  The observed-indicator product O^T B
// (O = 1 - missing) is then supplied exactly as a.qext (column sums minus the missing entries)
// for EPI_PQ / EPI_KV_GT, and is added by the caller after the kernel for EPI_KV_G.
//
// MT x GZ: a wave covers MT 16-row subtiles and GZ groups of 16 right-hand sides; every decoded
// genotype fragment feeds GZ L (x2) MFMAs.  Grid: x = padded rows / (64 MT), y = k splits,
// z = ceil(G / GZ).
// LDSB (one group of 16 right-hand sides per workgroup, GZ = 1): the workgroup's four waves
// share the B fragments of each 256-k block through LDS (double-buffered, staged one block
// ahead), so a workgroup reads each B byte once instead of once per wave.
template <int L, bool TWO_B>
struct GenoLds {
    static constexpr int kChunksPerOp = 4 * L * 64;                     // 16-byte chunks per k block
    static constexpr int kChunks = (TWO_B ? 2 : 1) * kChunksPerOp;
    static constexpr int kPerThread = kChunks / 256;
};

template <int L, bool TWO_B, int EPI, bool ONLY_X = false, int MT = 4, int GZ = 1, bool LDSB = false, bool PK = false,
          bool DMA = false>
__device__ __forceinline__ void geno_body(const GenoArgs& a, void* lds_raw = nullptr) {
    // PK: the L limbs of up to 16 / L columns share one 16-column tile (column c, limb l in tile
    // column c L + l), so each decoded fragment feeds one MFMA per operand instead of L.
    constexpr int LT = PK ? 1 : L;   // limb tiles per group
    static_assert(!PK || GZ == 1, "packed limbs use one group");
    // B staging type: native vectors for the two-operand and packed forms (a HIP_vector_type
    // array there is copied through a stack object), the HIP vector type for the one-operand
    // form with register staging (fewer registers at three waves per SIMD).  With DMA there are
    // no staging registers and the LDS array must be a native vector type (see k_geno1s).
    using ST = typename std::conditional<TWO_B || PK || DMA, v4i, uint4>::type;
    ST* ldsb = static_cast<ST*>(lds_raw);
    // Z W needs only X~ W1 - O (mu W): with the second operand stored negated, both products
    // accumulate into one set of int32 registers (exact while |sum| < 2^31: 384 per k-element).
    constexpr bool FUSE = TWO_B && EPI == EPI_KV_G && !ONLY_X;
    int64_t bx = blockIdx.x, by = blockIdx.y;
    int bz = blockIdx.z;
    if (a.flat_tiles > 0) {
        const int64_t b = blockIdx.x, s = b >> 3;
        const int64_t q = s / a.G;
        bz = (int)(s - q * a.G);
        const int64_t T = q * 8 + (b & 7);
        if (T >= a.flat_tiles) return;   // whole workgroup, before any barrier
        by = T / a.tiles_x;
        bx = T - by * a.tiles_x;
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = lane >> 4, r = lane & 15;
    const int64_t rbase = (bx * 4 + wave) * (16 * MT);   // first row of the wave
    const int64_t tr = rbase >> 6;
    const int m0 = (int)((rbase & 63) >> 4);
    const int gz0 = bz * GZ;
    const int NT = a.G * LT;
    const int64_t kb0 = by * a.kb_per_split;
    int64_t kb1 = kb0 + a.kb_per_split;
    if (kb1 > a.kb_total) kb1 = a.kb_total;

    v4i accX[MT][GZ][LT], accO[MT][GZ][LT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int j = 0; j < GZ; ++j)
#pragma unroll
            for (int l = 0; l < LT; ++l) {
                accX[m][j][l] = v4i{0, 0, 0, 0};
                accO[m][j][l] = v4i{0, 0, 0, 0};
            }
    // Group validity.  The LDS kernels (one group per workgroup) are launched with z = G or the
    // flat mapping, so their group is always valid.  A compile-time true keeps the MFMA blocks
    // of the m-subtiles in one basic block, which the compiler interleaves; it is used where the
    // interleaved schedule still fits the register budget (the one-operand two-product form at
    // three waves per SIMD would spill).
    constexpr bool kAlwaysValid = PK || (LDSB && (TWO_B || ONLY_X));
    bool gok[GZ];
#pragma unroll
    for (int j = 0; j < GZ; ++j) gok[j] = kAlwaysValid || gz0 + j < a.G;

    if (kb0 < kb1) {
        const uint8_t* abase = a.A + ((tr * a.KB) << 12) + m0 * 1024 + lane * 16;
        const int8_t* bxp = a.Bx + ((int64_t)(gz0 * LT) << 10) + lane * 16;
        const int8_t* bop = a.Bo + ((int64_t)(gz0 * LT) << 10) + lane * 16;
        const int64_t last_step = kb1 * 4 - 1;
        uint4 av[MT], an[MT];
        v4i bx[GZ][LT], bo[GZ][LT], bxn[GZ][LT], bon[GZ][LT];
        using LD = GenoLds<LT, TWO_B>;
        ST stg[LD::kPerThread];
        auto stage_load = [&](int64_t kb) {
#pragma unroll
            for (int i = 0; i < LD::kPerThread; ++i) {
                const int c = (int)threadIdx.x + 256 * i;
                const int op = c / LD::kChunksPerOp, w = c - op * LD::kChunksPerOp;
                const int t = w / (LT * 64), rem = w - t * (LT * 64);   // step t, limb tiles of group gz0
                const ST* src = reinterpret_cast<const ST*>(op == 0 ? a.Bx : a.Bo);
                stg[i] = src[((kb * 4 + t) * NT + gz0 * LT) * 64 + rem];
            }
        };
        auto stage_store = [&](int buf) {
#pragma unroll
            for (int i = 0; i < LD::kPerThread; ++i) ldsb[buf * LD::kChunks + (int)threadIdx.x + 256 * i] = stg[i];
        };
        // DMA: the B chunks go from global memory straight into LDS (global_load_lds_dwordx4),
        // with no staging registers.  The 64 lanes of a wave load 64 consecutive chunks, which
        // land at the wave's LDS base + 16 lane, the same place stage_store puts them.  The
        // end-of-block wait vmcnt(0) covers them (and the A prefetch, needed next anyway).
        typedef __attribute__((address_space(3))) ST lds_st;
        auto stage_dma = [&](int64_t kb, int buf) {
#pragma unroll
            for (int i = 0; i < LD::kPerThread; ++i) {
                const int c = (int)threadIdx.x + 256 * i;
                const int op = c / LD::kChunksPerOp, w = c - op * LD::kChunksPerOp;
                const int t = w / (LT * 64), rem = w - t * (LT * 64);
                const ST* src = reinterpret_cast<const ST*>(op == 0 ? a.Bx : a.Bo);
                const int dst = __builtin_amdgcn_readfirstlane(buf * LD::kChunks + 256 * i + 64 * wave);
                __builtin_amdgcn_global_load_lds(src + ((kb * 4 + t) * NT + gz0 * LT) * 64 + rem,
                                                 (lds_st*)ldsb + dst, 16, 0, 0);
            }
        };
        constexpr int kWaitVm0 = 0x0F70;   // s_waitcnt vmcnt(0)
#pragma unroll
        for (int m = 0; m < MT; ++m) av[m] = *reinterpret_cast<const uint4*>(abase + (kb0 << 12) + m * 1024);
        if (LDSB && DMA) {
            stage_dma(kb0, 0);
            __builtin_amdgcn_s_waitcnt(kWaitVm0);
            __syncthreads();
        } else if (LDSB) {
            stage_load(kb0);
            stage_store(0);
            __syncthreads();
        } else {
#pragma unroll
            for (int j = 0; j < GZ; ++j)
#pragma unroll
                for (int l = 0; l < LT; ++l) {
                    if (!gok[j]) continue;
                    bx[j][l] = *reinterpret_cast<const v4i*>(bxp + ((kb0 * 4 * NT + j * LT + l) << 10));
                    if (TWO_B) bo[j][l] = *reinterpret_cast<const v4i*>(bop + ((kb0 * 4 * NT + j * LT + l) << 10));
                }
        }
        for (int64_t kb = kb0; kb < kb1; ++kb) {
            const int64_t kn = kb + 1 < kb1 ? kb + 1 : kb;
            const int cur = (int)((kb - kb0) & 1);
            if (LDSB && kb + 1 < kb1) {
                if (DMA) stage_dma(kb + 1, cur ^ 1);
                else stage_load(kb + 1);
            }
#pragma unroll
            for (int m = 0; m < MT; ++m) an[m] = *reinterpret_cast<const uint4*>(abase + (kn << 12) + m * 1024);
            // keep the A prefetch at the top of the block (without staging registers the scheduler
            // would sink it next to its use at the end of the block)
            if (DMA) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                if (LDSB) {
                    const ST* lb = ldsb + cur * LD::kChunks;
#pragma unroll
                    for (int l = 0; l < LT; ++l) {
                        bx[0][l] = as_v4i(lb[(t * LT + l) * 64 + lane]);
                        if (TWO_B) {
                            bo[0][l] = as_v4i(lb[LD::kChunksPerOp + (t * LT + l) * 64 + lane]);
                        }
                    }
                }
                int64_t sn = kb * 4 + t + 1;
                if (sn > last_step) sn = last_step;
#pragma unroll
                for (int j = 0; j < GZ; ++j)
#pragma unroll
                    for (int l = 0; l < LT; ++l) {
                        if (LDSB || !gok[j]) continue;
                        bxn[j][l] = *reinterpret_cast<const v4i*>(bxp + ((sn * NT + j * LT + l) << 10));
                        if (TWO_B) bon[j][l] = *reinterpret_cast<const v4i*>(bop + ((sn * NT + j * LT + l) << 10));
                    }
#pragma unroll
                for (int m = 0; m < MT; ++m) {
                    const uint32_t w = t == 0 ? av[m].x : t == 1 ? av[m].y : t == 2 ? av[m].z : av[m].w;
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
                    for (int j = 0; j < GZ; ++j) {
                        if (!gok[j]) continue;
#pragma unroll
                        for (int l = 0; l < LT; ++l) {
                            accX[m][j][l] = __builtin_amdgcn_mfma_i32_16x16x64_i8(X, bx[j][l], accX[m][j][l], 0, 0, 0);
                            if (!ONLY_X && !FUSE)
                                accO[m][j][l] = __builtin_amdgcn_mfma_i32_16x16x64_i8(O, TWO_B ? bo[j][l] : bx[j][l],
                                                                                      accO[m][j][l], 0, 0, 0);
                        }
                        if (FUSE) {
#pragma unroll
                            for (int l = 0; l < LT; ++l)
                                accX[m][j][l] = __builtin_amdgcn_mfma_i32_16x16x64_i8(O, bo[j][l], accX[m][j][l], 0, 0, 0);
                        }
                    }
                }
                if (!LDSB) {
#pragma unroll
                    for (int j = 0; j < GZ; ++j)
#pragma unroll
                        for (int l = 0; l < LT; ++l) {
                            bx[j][l] = bxn[j][l];
                            if (TWO_B) bo[j][l] = bon[j][l];
                        }
                }
            }
#pragma unroll
            for (int m = 0; m < MT; ++m) av[m] = an[m];
            if (LDSB && DMA) {
                __builtin_amdgcn_s_waitcnt(kWaitVm0);
                __syncthreads();
            } else if (LDSB) {
                if (kb + 1 < kb1) stage_store(cur ^ 1);
                __syncthreads();
            }
        }
    }

    // Epilogue.  16x16 MFMA C layout: lane holds column r of its group, rows 4g + v (v = 0..3)
    // of each m-subtile; the L limbs of one column sit in the same lane (tiles gz*L + l), or with
    // PK in lanes c L + l of the same 16-lane group (gathered by the lane of limb 0).
#pragma unroll
    for (int j = 0; j < GZ; ++j) {
        const int64_t rhs = PK ? (int64_t)(r / L) : (int64_t)(gz0 + j) * 16 + r;
        const bool col_ok = PK ? (r % L == 0 && rhs < a.k) : (gok[j] && rhs < a.k);
        double wmax = 0.0;
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t row = rbase + 16 * m + 4 * g + v;
                long long P = 0, Q = 0;
                if (PK) {
#pragma unroll
                    for (int l = L - 1; l >= 0; --l) {
                        P = P * 256 + __shfl(accX[m][j][0][v], lane + l);
                        if (!ONLY_X && !FUSE) Q = Q * 256 + __shfl(accO[m][j][0][v], lane + l);
                    }
                } else {
#pragma unroll
                    for (int l = L - 1; l >= 0; --l) {
                        P = P * 256 + accX[m][j][l][v];
                        if (!ONLY_X && !FUSE) Q = Q * 256 + accO[m][j][l][v];
                    }
                }
                if (ONLY_X && EPI != EPI_KV_G && col_ok && row < a.rows) Q = a.qext[row * 16 * (int64_t)a.G + rhs];
                if (EPI == EPI_PQ) {
                    if (col_ok && row < a.rows) {
                        const int64_t G16 = 16 * (int64_t)a.G;
                        a.pq[(row * 2) * G16 + rhs] = P;
                        a.pq[(row * 2 + 1) * G16 + rhs] = Q;
                    }
                } else if (EPI == EPI_KV_GT) {
                    if (col_ok && a.out_u) {
                        double u = 0.0;
                        if (row < a.rows && a.inv_stds[row] != 0.0)
                            u = ((double)P - a.means[row] * (double)Q) * a.inv_stds[row] * a.scale[rhs];
                        a.W1[row * a.k + rhs] = u;
                    } else if (col_ok) {
                        double w1 = 0.0, w2 = 0.0;
                        if (row < a.rows && a.mult[row] != 0 && a.inv_stds[row] != 0.0) {
                            const double s = a.inv_stds[row], mu = a.means[row];
                            const double u = ((double)P - mu * (double)Q) * s * a.scale[rhs];
                            const double w = (double)a.mult[row] * s * u * a.inv_M;
                            w1 = w;
                            w2 = mu * w;
                        }
                        a.W1[row * a.k + rhs] = w1;
                        a.W2[row * a.k + rhs] = w2;
                        wmax = fmax(wmax, fmax(fabs(w1), fabs(w2)));
                    }
                } else {
                    if (col_ok && row < a.rows)
                        atomicAdd(reinterpret_cast<unsigned long long*>(a.yint + row * a.k + rhs),
                                  (unsigned long long)(ONLY_X || FUSE ? P : P - Q));
                }
            }
        if (EPI == EPI_KV_GT) {
            wmax = fmax(wmax, __shfl_xor(wmax, 16));
            wmax = fmax(wmax, __shfl_xor(wmax, 32));
            if (g == 0 && col_ok && wmax > 0.0)
                atomicMax(a.mx + rhs, (unsigned long long)__double_as_longlong(wmax));
        }
    }
}

// Entry points.  The two-operand form carries twice the B registers; without an explicit
// occupancy target the compiler sinks its B prefetch loads next to their uses (every MFMA step
// then waits on the outstanding A prefetch), so it is pinned to 2 waves per SIMD.
template <int L, int EPI, int MT = 4, int GZ = 1>
__global__ __launch_bounds__(256) void k_geno1(GenoArgs a) {
    geno_body<L, false, EPI, false, MT, GZ>(a);
}
// Genotype-only product (the sparse observed-indicator form).
template <int L, int EPI>
__global__ __launch_bounds__(256) void k_genox(GenoArgs a) {
    geno_body<L, false, EPI, true>(a);
}
template <int L, int EPI, int MT = 4, int GZ = 1>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_geno2(GenoArgs a) {
    geno_body<L, true, EPI, false, MT, GZ>(a);
}
// One group of 16 right-hand sides with the B fragments shared through LDS (GenoLds).  For Z W
// the two products share one accumulator set (FUSE), which leaves room for 3 waves per SIMD
// (two with four limbs).  Packed limbs (PK, k L <= 16) need a quarter of the registers and run
// four waves per SIMD.
//
// DMA: the B fragments are staged by direct global-to-LDS loads, which frees the staging
// registers (24 for Z W, 12 for Z^T V).  The LDS array is a native vector type: with the HIP
// uint4 struct the compiler could not tell the B reads from the in-flight LDS writes and waited
// for every outstanding vector load before the first read of a block.
template <int L, int EPI, bool PK = false, bool DMA = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PK ? 4 : 3, PK ? 4 : 3))) void
k_geno2s(GenoArgs a) {
    __shared__ v4i ldsb[2 * GenoLds<PK ? 1 : L, true>::kChunks];
    geno_body<L, true, EPI, false, 4, 1, true, PK, DMA>(a, ldsb);
}
template <int L, int EPI, bool PK = false, bool DMA = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PK ? 4 : 3, PK ? 4 : 3))) void
k_geno1s(GenoArgs a) {
    __shared__ typename std::conditional<PK || DMA, v4i, uint4>::type ldsb[2 * GenoLds<PK ? 1 : L, false>::kChunks];
    geno_body<L, false, EPI, false, 4, 1, true, PK, DMA>(a, ldsb);
}
// Genotype-only product with the B fragments shared through LDS: one accumulator set of
// 4 x L tiles, four waves per SIMD.
template <int L, int EPI, bool DMA = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_genox_s(GenoArgs a) {
    __shared__ typename std::conditional<DMA, v4i, uint4>::type ldsb[2 * GenoLds<L, false>::kChunks];
    geno_body<L, false, EPI, true, 4, 1, true, false, DMA>(a, ldsb);
}
// Four limbs (two accumulator sets of 4 x 4 tiles) fit only at two waves per SIMD.
template <int L, int EPI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_geno1s2(GenoArgs a) {
    __shared__ uint4 ldsb[2 * GenoLds<L, false>::kChunks];
    geno_body<L, false, EPI, false, 4, 1, true>(a, ldsb);
}

}  // namespace csamd
