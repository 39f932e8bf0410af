// Kernel (b): the matrix-free kinship operator K = Z_sel Z_sel^T / M_sel on the resident packed
// genotypes (PlinkKinshipOperator, src/kinship_operator.cpp:28-374).
//
// Factorisation (kinship_operator.cpp:101-135, :162-179):
//   u_j = s_j (sum_i x~_ij v_i - mu_j sum_i o_ij v_i)         (Z^T V, missing excluded from both)
//   y_i = sum_j x~_ij w_j - sum_j o_ij (mu_j w_j),  w_j = mult_j s_j u_j / M_sel
// Samples: keep lists (any order, repeats) act through a scatter-add of V into the full sample
// space and a gather of the result (K_sel is the principal submatrix of the full K).  Variants:
// a per-SNP multiplicity (0 = not selected), so LOCO masks cost nothing extra.
// Both sums run on int8 MFMA with the genotype operand exact and V / W as 2 or 3 signed 8-bit
// limbs of a per-column fixed point; Z^T V streams the SNP-major copy and writes w, mu w directly
// from its epilogue; Z W streams the sample-major copy and accumulates exact int64 results.
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstring>

#include <mutex>

#include "cs_common.h"
#include "cs_geno_mfma.h"
#include "cs_ops.h"

struct cs_kinship {
    cs_bed* bed = nullptr;
    int64_t n = 0, Msel = 0;
    bool identity = true;
    std::vector<int32_t> keep;      // empty = identity
    std::vector<int32_t> variants;  // empty = all
    int prec = CS_PREC_INT8_L3;
    int missing_mode = 0;           // 0 auto, 1 dense O-products, 2 sparse missing lists
    hipStream_t stream = nullptr;
    csamd::DevBuf d_keep, d_inv_off, d_inv_idx, d_mult;
    // workspaces (grown on demand)
    csamd::DevBuf dX, dY, Vt, BV, W1, W2, B1, B2, Yint, mxW, mx, scl, Vf, PQf, Wf, Yf;
    csamd::DevBuf Vp, B2p, S1, S2, qext;   // sparse-missing form: plain quantised operands, sums
    csamd::DevBuf diag_tab, diag_full;
    cs_bed* rows = nullptr;         // compact copy of the cohort's genotype rows (owned), or null
    std::vector<double> diag_cache;  // diag(K) over the rows, once computed (fixed by keep / variants)
    hipEvent_t ev[6] = {};
    float last_ms[5] = {0, 0, 0, 0, 0};
};

namespace csamd {

// The matrix the genotype products read (the cohort's compact copy when there is one) and
// whether its rows are the operator's rows in order (no scatter / gather through keep).
static cs_bed* row_bed(const cs_kinship* op) { return op->rows ? op->rows : op->bed; }
static bool rows_direct(const cs_kinship* op) { return op->identity || op->rows != nullptr; }

// ------------------------------------------------------------------------------------------
__global__ void k_scatter(const double* __restrict__ V, int64_t ldv, int64_t k,
                          const int32_t* __restrict__ inv_off, const int32_t* __restrict__ inv_idx,
                          int64_t N, int64_t Npad, double* __restrict__ Vt) {
    const int64_t total = Npad * k;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = e / k, c = e - s * k;
        double acc = 0.0;
        if (s < N) {
            if (inv_off) {
                for (int32_t p = inv_off[s]; p < inv_off[s + 1]; ++p) acc += V[inv_idx[p] + c * ldv];
            } else {
                acc = V[s + c * ldv];
            }
        }
        Vt[e] = acc;
    }
}

// Column max |W| of a row-major [rows][k] matrix; one column per blockIdx.y.
__global__ __launch_bounds__(256) void k_colmaxabs(const double* __restrict__ W, int64_t rows,
                                                   int64_t k, unsigned long long* __restrict__ mx) {
    __shared__ double red[256];
    const int64_t c = blockIdx.y;
    double m = 0.0;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < rows;
         r += (int64_t)gridDim.x * blockDim.x)
        m = fmax(m, fabs(W[r * k + c]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(&mx[c], (unsigned long long)__double_as_longlong(red[0]));
}

// scale[c] = max/qmax (1 if the column is zero); inv[c] = 1/scale[c]
__global__ void k_scales(const unsigned long long* __restrict__ mx, int64_t k, double qmax,
                         double* __restrict__ scale, double* __restrict__ inv) {
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= k) return;
    const double m = __longlong_as_double((long long)mx[c]);
    const double s = m > 0.0 ? m / qmax : 1.0;
    scale[c] = s;
    inv[c] = 1.0 / s;
}

// Fixed-point quantisation into L balanced signed 8-bit limbs, stored pre-permuted for the MFMA
// B operand: column c, limb l -> tile (c/16)*L + l, column c%16 of the tile.
// sign = -1 quantises -W (rint and the clamp are symmetric, so the digits give exactly -q(W)).
// packed: limb l of column c in column c L + l of tile 0 (the packed tiling).
__global__ void k_quant(const double* __restrict__ W, int64_t rows, int64_t k, int L, int NT,
                        double qmax, const double* __restrict__ inv, int8_t* __restrict__ B,
                        int32_t* __restrict__ plain = nullptr, int64_t K16 = 0, double sign = 1.0,
                        int packed = 0) {
    const int64_t total = rows * k;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = e / k;
        const int c = (int)(e - row * k);
        double qd = rint(sign * W[e] * inv[c]);
        qd = fmin(fmax(qd, -qmax), qmax);
        long long q = (long long)qd;
        if (plain) plain[row * K16 + c] = (int32_t)q;
        if (!B) continue;
        for (int l = 0; l < L; ++l) {
            int d;
            if (l == L - 1) {
                d = (int)q;
            } else {
                d = (int)(((q + 128) & 255) - 128);
                q = (q - d) >> 8;
            }
            if (packed) B[bperm_offset(row, 0, c * L + l, NT)] = (int8_t)d;
            else B[bperm_offset(row, (c >> 4) * L + l, c & 15, NT)] = (int8_t)d;
        }
    }
}

// The two B operands of Z W in one pass: W1 and -W2 quantised with the common per-column
// inverse scale inv, every byte of both written (zero digits in the padding columns and limbs),
// and the int64 result rows cleared.  A thread builds whole 16-byte B chunks: the 16
// consecutive k-elements (rows) 16 rg .. 16 rg + 15 of one tile column, whose bytes 4 s + q hold
// row 16 rg + 4 q + s (bperm_offset).  Unpacked: thread (rg, column cc < 16 G) writes the L limb
// tiles of cc; packed: thread (rg, j < 16) writes tile column j = c L + l.  Neighbouring threads
// read neighbouring columns of W and store neighbouring chunks.
// TWO = false: W1 into B1 only (W2, B2 and yint unused), the B operand of any product.
template <int L, bool packed, bool TWO = true>
__global__ void k_quant_zw(const double* __restrict__ W1, const double* __restrict__ W2, int64_t rows, int64_t k,
                           int NT, int CW, double qmax, const double* __restrict__ inv,
                           int8_t* __restrict__ B1, int8_t* __restrict__ B2,
                           long long* __restrict__ yint, int64_t ny) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t first = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (TWO)
        for (int64_t e = first; e < ny; e += stride) yint[e] = 0;
    const int64_t total = (rows / 16) * CW;
    for (int64_t e = first; e < total; e += stride) {
        const int64_t rg = e / CW;
        const int j = (int)(e - rg * CW);
        const int c = packed ? j / L : j;
        const int l0 = packed ? j % L : 0;
        const bool ok = packed ? j < k * L : j < k;
        // chunk words: [limb][word s]; byte q of word s = row 16 rg + 4 q + s
        constexpr int NL = packed ? 1 : L;
        uint32_t w1[NL][4], w2[NL][4];
#pragma unroll
        for (int l = 0; l < NL; ++l)
#pragma unroll
            for (int s = 0; s < 4; ++s) w1[l][s] = w2[l][s] = 0u;
        if (ok) {
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int64_t row = rg * 16 + i;
                const double d1 = rint(W1[row * k + c] * inv[c]), d2 = TWO ? rint(-W2[row * k + c] * inv[c]) : 0.0;
                long long q1 = (long long)fmin(fmax(d1, -qmax), qmax);
                long long q2 = (long long)fmin(fmax(d2, -qmax), qmax);
                const int s = i & 3, sh = 8 * (i >> 2);
#pragma unroll
                for (int l = 0; l < L; ++l) {
                    const int e1 = (l == L - 1) ? (int)q1 : (int)(((q1 + 128) & 255) - 128);
                    const int e2 = (l == L - 1) ? (int)q2 : (int)(((q2 + 128) & 255) - 128);
                    q1 = (q1 - e1) >> 8;
                    q2 = (q2 - e2) >> 8;
                    if (packed) {   // keep limb l0 only
                        if (l == l0) {
                            w1[0][s] |= (uint32_t)(uint8_t)e1 << sh;
                            w2[0][s] |= (uint32_t)(uint8_t)e2 << sh;
                        }
                    } else {
                        w1[packed ? 0 : l][s] |= (uint32_t)(uint8_t)e1 << sh;
                        w2[packed ? 0 : l][s] |= (uint32_t)(uint8_t)e2 << sh;
                    }
                }
            }
        }
#pragma unroll
        for (int l = 0; l < NL; ++l) {
            const int64_t o = packed ? bperm_offset(rg * 16, 0, j, NT) : bperm_offset(rg * 16, (c >> 4) * L + l, c & 15, NT);
            *reinterpret_cast<uint4*>(B1 + o) = make_uint4(w1[l][0], w1[l][1], w1[l][2], w1[l][3]);
            if (TWO) *reinterpret_cast<uint4*>(B2 + o) = make_uint4(w2[l][0], w2[l][1], w2[l][2], w2[l][3]);
        }
    }
}

// Exact column sums of a plain int32 [rows][K16] operand into S (zeroed by the caller): block
// (x, y) sums a stride of the rows for columns 16 y .. 16 y + 15 (16 threads read one 64-byte row
// piece) and adds its partial sums with one int64 atomic per column (integer addition, so the
// result does not depend on the order).
__global__ __launch_bounds__(256) void k_colsum_plain(const int32_t* __restrict__ P, int64_t rows, int64_t K16,
                                                      unsigned long long* __restrict__ S) {
    __shared__ long long red[256];
    const int c = threadIdx.x & 15, rr = threadIdx.x >> 4;
    const int64_t col = (int64_t)blockIdx.y * 16 + c;
    long long acc = 0;
    for (int64_t r = (int64_t)blockIdx.x * 16 + rr; r < rows; r += (int64_t)gridDim.x * 16) acc += P[r * K16 + col];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off >= 16; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x < 16) atomicAdd(S + col, (unsigned long long)red[threadIdx.x]);
}

static void colsum_plain(const int32_t* P, int64_t rows, int64_t K16, long long* S, hipStream_t st) {
    CS_HIP(hipMemsetAsync(S, 0, (size_t)K16 * 8, st));
    const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>(1024, (rows + 15) / 16));
    k_colsum_plain<<<dim3(gx, (unsigned)(K16 / 16)), 256, 0, st>>>(P, rows, K16,
                                                                   reinterpret_cast<unsigned long long*>(S));
    CS_HIP(hipGetLastError());
}

// Sum over the entries of one missing list of the rows of a plain int32 [.][K16] operand
// (exact, int64).  One wave per list row; lane = e CW + cq
// holds column quad `quad` (16 bytes) of list entry e of each batch of E = 64 / CW entries.  A
// step gathers U batches and, before adding them, issues the index loads of the next U batches,
// so a wave keeps U E row gathers and U index loads in flight.  The indices are read once
// (non-temporal); the operand rows are cache hits.  The sums are reduced over e (valid in every
// lane).
// W = 64 (one row per wave) or 32 (one row per half-wave: two rows' gathers in flight per wave;
// the reduction shuffles stay inside the half).
template <int U, int CW, int W = 64>
__device__ __forceinline__ void list_row_sums(const int32_t* __restrict__ idx, int64_t p0, int64_t p1,
                                              const int4* __restrict__ P4, uint32_t q4, int64_t quad, int e,
                                              long long acc[4]) {
    static_assert(W == 64 || W == 32, "a row takes a wave or a half-wave");
    static_assert(CW <= W, "at least one entry per batch");
    constexpr int E = W / CW;
    acc[0] = acc[1] = acc[2] = acc[3] = 0;
    if (p1 == p0) return;   // empty row (p0, p1 are the same in every lane of the row): nothing to reduce
    const int4* col = P4 + quad;
    const int32_t* ip = idx + p0 + e;
    int64_t left = p1 - p0 - e;   // entries of this lane's stride still to read, from ip
    int id[U];
#pragma unroll
    for (int u = 0; u < U; ++u) id[u] = E * u < left ? __builtin_nontemporal_load(ip + E * u) : -1;
    while (left > 0) {
        int4 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = id[u] >= 0 ? col[(uint32_t)id[u] * q4] : make_int4(0, 0, 0, 0);
        ip += E * U;
        left -= E * U;
#pragma unroll
        for (int u = 0; u < U; ++u) id[u] = E * u < left ? __builtin_nontemporal_load(ip + E * u) : -1;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            acc[0] += v[u].x;
            acc[1] += v[u].y;
            acc[2] += v[u].z;
            acc[3] += v[u].w;
        }
    }
    for (int s = CW; s < W; s <<= 1)
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[t] += __shfl_xor(acc[t], s);
}

// Row offsets of a block of R < 64 list rows starting at j0, loaded by one instruction (lane
// l <= R holds off[min(j0 + l, rows)]); row r of the block spans [lane r, lane r + 1).
template <int R>
__device__ __forceinline__ int64_t block_offsets(const int64_t* __restrict__ off, int64_t j0, int64_t rows) {
    static_assert(R > 0 && R < 64, "the end of the block sits in lane R");
    const int lane = threadIdx.x & 63;
    const int64_t j = j0 + lane;
    return lane <= R ? off[j < rows ? j : rows] : 0;
}

// Observed-indicator product of the SNP rows from the missing lists:
// qext[j][c] = sum_i o_ij Vq_i[c] = S[c] - sum_{i missing at j} Vq_i[c]   (exact integers).
// CW: a power of two dividing K16 / 4, at most 64.  Each wave takes blocks of R rows.
template <int CW, int R, int W>
__global__ __launch_bounds__(256) void k_sparse_q(const int64_t* __restrict__ off, const int32_t* __restrict__ idx,
                                                  const int32_t* __restrict__ Vp, int64_t K16,
                                                  const long long* __restrict__ S, int64_t rows,
                                                  long long* __restrict__ qext) {
    constexpr int H = 64 / W;   // rows in flight per wave
    const int lane = threadIdx.x & 63, h = lane / W, e = (lane % W) / CW, cq = lane % CW;
    const int64_t q4 = K16 / 4;
    const int4* P4 = reinterpret_cast<const int4*>(Vp);
    for (int64_t j0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * R; j0 < rows; j0 += (int64_t)gridDim.x * 4 * R) {
        const int64_t o = block_offsets<R>(off, j0, rows);
        const int nr = (int)(rows - j0 < R ? rows - j0 : R);
        for (int rp = 0; rp < nr; rp += H) {
            // the offsets are read with uniform control flow; a row past the block end is empty
            const int r = rp + h;
            const int64_t j = j0 + r, p1 = __shfl(o, r + 1 < nr ? r + 1 : nr);
            const int64_t p0 = r < nr ? __shfl(o, r) : p1;
            for (int64_t c0 = 0; c0 < q4; c0 += CW) {
                const int64_t quad = c0 + cq;
                long long acc[4];
                list_row_sums<8, CW, W>(idx, p0, p1, P4, (uint32_t)q4, quad, e, acc);
                if (e == 0 && r < nr)
#pragma unroll
                    for (int t = 0; t < 4; ++t) qext[j * K16 + 4 * quad + t] = S[4 * quad + t] - acc[t];
            }
        }
    }
}

// Z W with the sparse observed-indicator form: yint[i][c] (= X~ W1 from the MFMA kernel) minus
// sum_j o_ij (mu W)_j[c] = S2[c] - sum_{j missing at i} (mu W)_j[c].
template <int CW, int R, int W>
__global__ __launch_bounds__(256) void k_sparse_yadd(const int64_t* __restrict__ off, const int32_t* __restrict__ idx,
                                                     const int32_t* __restrict__ B2p, int64_t K16,
                                                     const long long* __restrict__ S2, int64_t rows, int64_t k,
                                                     long long* __restrict__ yint) {
    constexpr int H = 64 / W;   // rows in flight per wave
    const int lane = threadIdx.x & 63, h = lane / W, e = (lane % W) / CW, cq = lane % CW;
    const int64_t q4 = K16 / 4;
    const int4* P4 = reinterpret_cast<const int4*>(B2p);
    for (int64_t i0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * R; i0 < rows; i0 += (int64_t)gridDim.x * 4 * R) {
        const int64_t o = block_offsets<R>(off, i0, rows);
        const int nr = (int)(rows - i0 < R ? rows - i0 : R);
        for (int rp = 0; rp < nr; rp += H) {
            const int r = rp + h;
            const int64_t i = i0 + r, p1 = __shfl(o, r + 1 < nr ? r + 1 : nr);
            const int64_t p0 = r < nr ? __shfl(o, r) : p1;
            for (int64_t c0 = 0; c0 < q4; c0 += CW) {
                const int64_t quad = c0 + cq;
                long long acc[4];
                list_row_sums<8, CW, W>(idx, p0, p1, P4, (uint32_t)q4, quad, e, acc);
                if (e == 0 && r < nr)
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        const int64_t c = 4 * quad + t;
                        if (c < k) yint[i * k + c] += acc[t] - S2[c];
                    }
            }
        }
    }
}

// Column quads handled together by one wave of the list gathers: the largest power of two that
// divides K16 / 4, at most 64.
static int list_quads(int64_t K16) {
    const int64_t q4 = K16 / 4;
    int cw = 1;
    while (cw < 64 && q4 % (2 * cw) == 0) cw *= 2;
    return cw;
}

// y = scale_c * yint[keep[r]][c] into the column-major output.
__global__ void k_finalize(const long long* __restrict__ yint, int64_t k,
                           const double* __restrict__ scale, const int32_t* __restrict__ keep,
                           int64_t n, double* __restrict__ Y, int64_t ldy) {
    const int64_t total = n * k;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e / k;
        const int c = (int)(e - r * k);
        const int64_t s = keep ? keep[r] : r;
        Y[r + c * ldy] = (double)yint[s * k + c] * scale[c];
    }
}

// ------------------------------------------------------------------------------------------
// Plain FP32 comparator: the same K.V with genotypes decoded to FP32 values and FP32 FMAs on the
// vector ALUs (no MFMA, no fixed-point split), over the same tiled copies.  One wave per 64-row
// tile; lane (g, r) holds rows 16 m + r (m = 0..3) and the 64 k positions of its 16-byte chunk g
// of every 256-k block, so each right-hand-side value read from LDS feeds four rows.  The
// right-hand sides of a k-block are staged in LDS (double-buffered) for the workgroup; the four
// chunk lanes of a row are reduced by shuffles at the end.
//
// ZᵀV (rows = SNPs): a per-SNP table z[code] = {(2-mu)s, 0, (1-mu)s, -mu s} gives the
// standardised value, one FMA per column.  Z W (rows = samples): y = X~ W1 - O W2 with
// W1 = w, W2 = mu w per SNP, two FMAs per column; split along the SNPs into partial rows.
typedef float f32x4n __attribute__((ext_vector_type(4)));   // native vector: stays in registers

template <int KC>
__global__ __launch_bounds__(256) void k_f32_zt(const uint8_t* __restrict__ A, int64_t KB, int64_t rows,
                                                const double* __restrict__ means,
                                                const double* __restrict__ inv_stds,
                                                const f32x4n* __restrict__ V, float* __restrict__ U) {
    constexpr int Q = (KC + 3) / 4;                 // f32x4n per k position
    constexpr int CH = 256 * Q;                     // f32x4n per k-block
    __shared__ f32x4n lb[2][CH];
    const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
    const int64_t tr = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    float t[4][4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int64_t row = tr * 64 + 16 * m + r;
        float mu = 0.f, sd = 0.f;
        if (row < rows && inv_stds[row] != 0.0) {
            mu = (float)means[row];
            sd = (float)inv_stds[row];
        }
        t[m][0] = (2.f - mu) * sd;
        t[m][1] = 0.f;
        t[m][2] = (1.f - mu) * sd;
        t[m][3] = -mu * sd;
    }
    float acc[4][4 * Q];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int c = 0; c < 4 * Q; ++c) acc[m][c] = 0.f;
    const uint8_t* base = A + ((tr * KB) << 12) + lane * 16;
    f32x4n stg[(CH + 255) / 256];
    auto load = [&](int64_t kb) {
#pragma unroll
        for (int i = 0; i < (CH + 255) / 256; ++i) {
            const int e = (int)threadIdx.x + 256 * i;
            if (e < CH) stg[i] = V[kb * CH + e];
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int i = 0; i < (CH + 255) / 256; ++i) {
            const int e = (int)threadIdx.x + 256 * i;
            if (e < CH) lb[buf][e] = stg[i];
        }
    };
    load(0);
    store(0);
    __syncthreads();
    for (int64_t kb = 0; kb < KB; ++kb) {
        const int cur = (int)(kb & 1);
        if (kb + 1 < KB) load(kb + 1);
        uint4 a[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) a[m] = *reinterpret_cast<const uint4*>(base + (kb << 12) + m * 1024);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            uint32_t w[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) w[m] = q == 0 ? a[m].x : q == 1 ? a[m].y : q == 2 ? a[m].z : a[m].w;
#pragma unroll 4
            for (int u = 0; u < 16; ++u) {
                const f32x4n* bp = &lb[cur][(64 * g + 16 * q + u) * Q];
                f32x4n bv[Q];
#pragma unroll
                for (int c = 0; c < Q; ++c) bv[c] = bp[c];
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const uint32_t code = (w[m] >> (2 * u)) & 3u;
                    const float z = (code & 1u) ? ((code & 2u) ? t[m][3] : t[m][1]) : ((code & 2u) ? t[m][2] : t[m][0]);
#pragma unroll
                    for (int c = 0; c < Q; ++c) {
                        acc[m][4 * c] = __builtin_fmaf(z, bv[c].x, acc[m][4 * c]);
                        acc[m][4 * c + 1] = __builtin_fmaf(z, bv[c].y, acc[m][4 * c + 1]);
                        acc[m][4 * c + 2] = __builtin_fmaf(z, bv[c].z, acc[m][4 * c + 2]);
                        acc[m][4 * c + 3] = __builtin_fmaf(z, bv[c].w, acc[m][4 * c + 3]);
                    }
                }
            }
        }
        if (kb + 1 < KB) store(cur ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
#pragma unroll
        for (int c = 0; c < KC; ++c) {
            acc[m][c] += __shfl_xor(acc[m][c], 16);
            acc[m][c] += __shfl_xor(acc[m][c], 32);
        }
        const int64_t row = tr * 64 + 16 * m + r;
        if (g == 0)
#pragma unroll
            for (int c = 0; c < KC; ++c) U[row * KC + c] = acc[m][c];
    }
}

// Z W over SNP split blockIdx.y (kb_per_split k-blocks each): partial rows P[split][row][KC].
template <int KC>
__global__ __launch_bounds__(256) void k_f32_z(const uint8_t* __restrict__ A, int64_t KB, int64_t kb_per_split,
                                               int64_t rows_pad, const f32x4n* __restrict__ W12,
                                               float* __restrict__ P) {
    constexpr int Q = (KC + 3) / 4;
    constexpr int CH = 256 * 2 * Q;                 // W1 and W2 of 256 SNPs
    __shared__ f32x4n lb[2][CH];
    const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
    const int64_t tr = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int64_t kb0 = (int64_t)blockIdx.y * kb_per_split;
    const int64_t kb1 = kb0 + kb_per_split < KB ? kb0 + kb_per_split : KB;
    float acc[4][4 * Q];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int c = 0; c < 4 * Q; ++c) acc[m][c] = 0.f;
    const uint8_t* base = A + ((tr * KB) << 12) + lane * 16;
    f32x4n stg[CH / 256];
    auto load = [&](int64_t kb) {
#pragma unroll
        for (int i = 0; i < CH / 256; ++i) stg[i] = W12[kb * CH + (int)threadIdx.x + 256 * i];
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int i = 0; i < CH / 256; ++i) lb[buf][(int)threadIdx.x + 256 * i] = stg[i];
    };
    if (kb0 < kb1) {
        load(kb0);
        store(0);
    }
    __syncthreads();
    for (int64_t kb = kb0; kb < kb1; ++kb) {
        const int cur = (int)((kb - kb0) & 1);
        if (kb + 1 < kb1) load(kb + 1);
        uint4 a[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) a[m] = *reinterpret_cast<const uint4*>(base + (kb << 12) + m * 1024);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            uint32_t w[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) w[m] = q == 0 ? a[m].x : q == 1 ? a[m].y : q == 2 ? a[m].z : a[m].w;
#pragma unroll 2
            for (int u = 0; u < 16; ++u) {
                const f32x4n* bp = &lb[cur][(64 * g + 16 * q + u) * 2 * Q];
                f32x4n b1[Q], b2[Q];
#pragma unroll
                for (int c = 0; c < Q; ++c) {
                    b1[c] = bp[c];
                    b2[c] = bp[Q + c];
                }
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const uint32_t code = (w[m] >> (2 * u)) & 3u;
                    const float x = code == 0u ? 2.f : (code == 2u ? 1.f : 0.f);
                    const float o = code == 1u ? 0.f : -1.f;
#pragma unroll
                    for (int c = 0; c < Q; ++c) {
                        acc[m][4 * c] = __builtin_fmaf(o, b2[c].x, __builtin_fmaf(x, b1[c].x, acc[m][4 * c]));
                        acc[m][4 * c + 1] = __builtin_fmaf(o, b2[c].y, __builtin_fmaf(x, b1[c].y, acc[m][4 * c + 1]));
                        acc[m][4 * c + 2] = __builtin_fmaf(o, b2[c].z, __builtin_fmaf(x, b1[c].z, acc[m][4 * c + 2]));
                        acc[m][4 * c + 3] = __builtin_fmaf(o, b2[c].w, __builtin_fmaf(x, b1[c].w, acc[m][4 * c + 3]));
                    }
                }
            }
        }
        if (kb + 1 < kb1) store(cur ^ 1);
        __syncthreads();
    }
    float* out = P + (int64_t)blockIdx.y * rows_pad * KC;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
#pragma unroll
        for (int c = 0; c < KC; ++c) {
            acc[m][c] += __shfl_xor(acc[m][c], 16);
            acc[m][c] += __shfl_xor(acc[m][c], 32);
        }
        const int64_t row = tr * 64 + 16 * m + r;
        if (g == 0)
#pragma unroll
            for (int c = 0; c < KC; ++c) out[row * KC + c] = acc[m][c];
    }
}

__global__ void k_to_f32_cols(const double* __restrict__ src, int64_t rows, int64_t k, int KC,
                              float* __restrict__ dst) {
    const int64_t total = rows * KC;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e / KC;
        const int c = (int)(e - r * KC);
        dst[e] = c < k ? (float)src[r * k + c] : 0.f;
    }
}

// W1 = mult s u / M and W2 = mu W1 per SNP, interleaved per 256-SNP block as [kb][j][W1 | W2].
__global__ void k_epi_w_f32(const float* __restrict__ U, int KC, int64_t M, int64_t Mpad, int64_t k,
                            const double* __restrict__ means, const double* __restrict__ inv_stds,
                            const int32_t* __restrict__ mult, float inv_M, float* __restrict__ W12) {
    const int64_t total = Mpad * KC;
    const int Qc = (KC + 3) / 4 * 4;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = e / KC;
        const int c = (int)(e - j * KC);
        float w1 = 0.f, w2 = 0.f;
        if (j < M && c < k && mult[j] != 0 && inv_stds[j] != 0.0) {
            w1 = (float)mult[j] * (float)inv_stds[j] * U[j * KC + c] * inv_M;
            w2 = (float)means[j] * w1;
        }
        float* o = W12 + j * 2 * Qc;
        o[c] = w1;
        o[Qc + c] = w2;
    }
}

// Y[r, c] = sum over the splits of P[split][keep[r]][c], in split order.
__global__ void k_epi_y_f32(const float* __restrict__ P, int64_t nsplit, int64_t rows_pad, int KC, int64_t k,
                            const int32_t* __restrict__ keep, int64_t n, double* __restrict__ Y, int64_t ldy) {
    const int64_t total = n * k;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e / k;
        const int c = (int)(e - r * k);
        const int64_t s = keep ? keep[r] : r;
        float acc = 0.f;
        for (int64_t p = 0; p < nsplit; ++p) acc += P[(p * rows_pad + s) * KC + c];
        Y[r + c * ldy] = (double)acc;
    }
}

// ------------------------------------------------------------------------------------------
// diag(K) over the sample-major copy: table T[j][code] = mult_j s_j^2 (x_code - mu_j)^2 (0 for
// missing), one wave per 64-sample tile summing T over the SNPs in a fixed order per lane (the
// k-block's table rows staged in LDS for the workgroup), the four chunk lanes of a sample reduced
// by shuffles.  The grid covers the padded samples (whole tiles; padding codes are "missing").
__global__ void k_diag_table(int64_t M, const double* __restrict__ means,
                             const double* __restrict__ inv_stds, const int32_t* __restrict__ mult,
                             double* __restrict__ T) {
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= M) return;
    const double s = inv_stds[j], mu = means[j], m = (double)mult[j];
    const double x[4] = {2.0, 0.0, 1.0, 0.0};
    for (int c = 0; c < 4; ++c) {
        const double z = (x[c] - mu) * s;
        T[j * 4 + c] = (c == 1 || s == 0.0 || mult[j] == 0) ? 0.0 : m * z * z;
    }
}

// Per-SNP trace terms mult_j s_j^2 sum_c n_c (x_c - mu_j)^2 from class counts [M][4].
__global__ void k_trace_terms(const int64_t* __restrict__ counts, int64_t M, const double* __restrict__ means,
                              const double* __restrict__ inv_stds, const int32_t* __restrict__ mult,
                              double* __restrict__ t) {
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= M) return;
    const double s = inv_stds[j], mu = means[j];
    if (s == 0.0 || mult[j] == 0) {
        t[j] = 0.0;
        return;
    }
    const double x[4] = {2.0, 0.0, 1.0, 0.0};
    double acc = 0.0;
    for (int c = 0; c < 4; ++c) {
        if (c == 1) continue;   // missing
        const double z = (x[c] - mu) * s;
        acc += (double)counts[j * 4 + c] * (z * z);
    }
    t[j] = (double)mult[j] * acc;
}

__global__ __launch_bounds__(256) void k_diag(const uint8_t* __restrict__ smp, int64_t KB,
                                              int64_t N, int64_t M, const double* __restrict__ T,
                                              double inv_M, double* __restrict__ diag) {
    __shared__ double tab[256 * 4];   // T of the k-block's 256 SNPs, shared by the four waves
    const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
    const int64_t tr = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint8_t* base = smp + ((tr * KB) << 12) + lane * 16;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t kb = 0; kb < KB; ++kb) {
        uint4 w4[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) w4[m] = *reinterpret_cast<const uint4*>(base + (kb << 12) + m * 1024);
        const int64_t js = kb * 256 + threadIdx.x;
#pragma unroll
        for (int c = 0; c < 4; ++c) tab[threadIdx.x * 4 + c] = js < M ? T[js * 4 + c] : 0.0;
        __syncthreads();
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const uint32_t ws[4] = {w4[m].x, w4[m].y, w4[m].z, w4[m].w};
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    acc[m] += tab[(g * 64 + 16 * q + u) * 4 + ((ws[q] >> (2 * u)) & 3u)];
        }
        __syncthreads();
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        for (int off = 16; off < 64; off <<= 1) acc[m] += __shfl_xor(acc[m], off);
        const int64_t s = tr * 64 + 16 * m + r;
        if (g == 0 && s < N) diag[s] = acc[m] * inv_M;
    }
}

// ------------------------------------------------------------------------------------------
// grid as the callers give it: x = padded rows / 256, y = splits, z = G.  The LDS kernels take
// one group per workgroup (with two or more groups on a 1-D XCD-aware grid); without LDS
// sharing (CSAMD_GENO_LDSB=0) two or more groups run with each wave on 32 rows and two groups.
bool geno_lds_enabled() {
    static const bool on = [] {
        const char* e = std::getenv("CSAMD_GENO_LDSB");
        return !(e && e[0] == '0');
    }();
    return on;
}

// Grid of the LDS kernels (one group per workgroup).  With several groups: a 1-D grid in
// XCD-aware order, so the groups of one tile run together on one XCD and read its A tiles from
// HBM about once.
static dim3 lds_grid(dim3 grid, const GenoArgs& a, GenoArgs& f) {
    f = a;
    if (a.G < 2) return grid;
    f.tiles_x = grid.x;
    f.flat_tiles = (int64_t)grid.x * grid.y;
    const int64_t blocks = (f.flat_tiles + 7) / 8 * 8 * a.G;
    if (blocks > 0x7fffffff) throw RuntimeError("launch_geno: grid too large");
    return dim3((unsigned)blocks);
}

// B operands of the LDS kernels (three limbs or fewer) staged by direct global-to-LDS loads
// unless CSAMD_GENO_DMA=0 (read once per process).
static bool geno_dma_enabled() {
    static const bool on = [] {
        const char* e = std::getenv("CSAMD_GENO_DMA");
        return !(e && e[0] == '0');
    }();
    return on;
}

template <int L, bool TWO_B, int EPI>
static void launch_one(dim3 grid, hipStream_t st, const GenoArgs& a, bool packed) {
    // B shared by the workgroup through LDS unless CSAMD_GENO_LDSB=0
    const bool ldsb = geno_lds_enabled();
    if (packed) {
        if (!ldsb || a.G != 1) throw RuntimeError("launch_geno: packed limbs need the LDS kernels and one group");
        const bool dma = geno_dma_enabled();
        if constexpr (TWO_B) {
            if (dma) k_geno2s<L, EPI, true, true><<<grid, 256, 0, st>>>(a);
            else k_geno2s<L, EPI, true><<<grid, 256, 0, st>>>(a);
        } else {
            if (dma) k_geno1s<L, EPI, true, true><<<grid, 256, 0, st>>>(a);
            else k_geno1s<L, EPI, true><<<grid, 256, 0, st>>>(a);
        }
        return;
    }
    if (ldsb) {
        GenoArgs f;
        const dim3 g = lds_grid(grid, a, f);
        if constexpr (TWO_B) {
            // B staged by direct global-to-LDS loads (three limbs or fewer: the four-limb kernels
            // run at two waves per SIMD with registers to spare)
            if (L != 4 && geno_dma_enabled()) k_geno2s<L, EPI, false, true><<<g, 256, 0, st>>>(f);
            else k_geno2s<L, EPI><<<g, 256, 0, st>>>(f);
        } else if constexpr (L == 4) {
            k_geno1s2<L, EPI><<<g, 256, 0, st>>>(f);
        } else {
            if (geno_dma_enabled()) k_geno1s<L, EPI, false, true><<<g, 256, 0, st>>>(f);
            else k_geno1s<L, EPI><<<g, 256, 0, st>>>(f);
        }
        return;
    }
    if constexpr (L == 4) {
        // four limbs: the two-group form exceeds the register file, so one group per z slice
        if (TWO_B) k_geno2<L, EPI><<<grid, 256, 0, st>>>(a);
        else k_geno1<L, EPI><<<grid, 256, 0, st>>>(a);
    } else if (a.G >= 2) {
        const dim3 g2(grid.x * 2, grid.y, (unsigned)((a.G + 1) / 2));
        if (TWO_B) k_geno2<L, EPI, 2, 2><<<g2, 256, 0, st>>>(a);
        else k_geno1<L, EPI, 2, 2><<<g2, 256, 0, st>>>(a);
    } else {
        if (TWO_B) k_geno2<L, EPI><<<grid, 256, 0, st>>>(a);
        else k_geno1<L, EPI><<<grid, 256, 0, st>>>(a);
    }
}

void launch_geno(int L, bool two_b, int epi, dim3 grid, hipStream_t st, const GenoArgs& a, bool only_x,
                 bool packed) {
    if (only_x) {
        if (two_b || packed) throw RuntimeError("launch_geno: the genotype-only product has one unpacked operand");
#define CS_GENOX_CASE(LL, EE)                                \
        if (L == LL && epi == EE) {                          \
            if (geno_lds_enabled()) {                        \
                GenoArgs f;                                  \
                const dim3 g = lds_grid(grid, a, f);         \
                if (geno_dma_enabled())                      \
                    k_genox_s<LL, EE, true><<<g, 256, 0, st>>>(f); \
                else                                         \
                    k_genox_s<LL, EE><<<g, 256, 0, st>>>(f);  \
            } else {                                         \
                k_genox<LL, EE><<<grid, 256, 0, st>>>(a);    \
            }                                                \
            CS_HIP(hipGetLastError());                       \
            return;                                          \
        }
        CS_GENOX_CASE(3, EPI_PQ)
        CS_GENOX_CASE(2, EPI_PQ)
        CS_GENOX_CASE(3, EPI_KV_GT)
        CS_GENOX_CASE(2, EPI_KV_GT)
        CS_GENOX_CASE(3, EPI_KV_G)
        CS_GENOX_CASE(2, EPI_KV_G)
#undef CS_GENOX_CASE
        throw RuntimeError("launch_geno: unsupported configuration");
    }
#define CS_GENO_CASE(LL, TB, EE)                         \
    if (L == LL && two_b == TB && epi == EE) {           \
        launch_one<LL, TB, EE>(grid, st, a, packed);     \
        CS_HIP(hipGetLastError());                       \
        return;                                          \
    }
    CS_GENO_CASE(3, false, EPI_PQ)
    CS_GENO_CASE(2, false, EPI_PQ)
    CS_GENO_CASE(3, false, EPI_KV_GT)
    CS_GENO_CASE(2, false, EPI_KV_GT)
    CS_GENO_CASE(3, true, EPI_KV_G)
    CS_GENO_CASE(2, true, EPI_KV_G)
    CS_GENO_CASE(4, false, EPI_PQ)
    CS_GENO_CASE(4, false, EPI_KV_GT)
    CS_GENO_CASE(4, true, EPI_KV_G)
#undef CS_GENO_CASE
    throw RuntimeError("launch_geno: unsupported configuration");
}

void quantize_cols(const double* const* W, int nW, int64_t rows, int64_t k, const Tiling& tl,
                   hipStream_t st, DevBuf& mx, DevBuf& scl, int8_t* const* B) {
    mx.ensure((size_t)k * sizeof(unsigned long long));
    scl.ensure((size_t)2 * k * sizeof(double));
    quantize_cols_into(W, nW, rows, k, tl, st, mx.as<unsigned long long>(), scl.as<double>(), B);
}

void quantize_cols_into(const double* const* W, int nW, int64_t rows, int64_t k, const Tiling& tl,
                        hipStream_t st, unsigned long long* mxp, double* sclp, int8_t* const* B) {
    const double qmax = limbs_qmax(tl.L);
    CS_HIP(hipMemsetAsync(mxp, 0, (size_t)k * sizeof(unsigned long long), st));
    for (int i = 0; i < nW; ++i)
        k_colmaxabs<<<dim3(64, (unsigned)k), 256, 0, st>>>(W[i], rows, k, mxp);
    k_scales<<<grid_for(k), 256, 0, st>>>(mxp, k, qmax, sclp, sclp + k);
    // every byte of each operand from whole 16-byte chunks (rows is a multiple of 256)
    if (rows % 16) throw RuntimeError("quantize_cols: rows must be a multiple of 16");
    const int cw = tl.packed ? 16 : 16 * tl.G;
    const unsigned gq = grid_for(rows / 16 * cw);
    const double* inv = sclp + k;
    for (int i = 0; i < nW; ++i) {
#define CS_QC(LL, PP)                                                                                      \
        if (tl.L == LL && tl.packed == PP)                                                                 \
            k_quant_zw<LL, PP, false><<<gq, 256, 0, st>>>(W[i], nullptr, rows, k, tl.NT, cw, qmax, inv, B[i], \
                                                          nullptr, nullptr, 0);
        CS_QC(2, false) CS_QC(3, false) CS_QC(4, false) CS_QC(2, true) CS_QC(3, true) CS_QC(4, true)
#undef CS_QC
    }
    CS_HIP(hipGetLastError());
}

// Sparse observed-indicator form when the missing genotypes are few: the MFMA kernels then run
// only the genotype products, and O^T V / O (mu W) come from column sums and the missing lists.
// Both forms give bitwise identical results.
static bool use_sparse_missing(cs_kinship* op, int64_t k, int L) {
    cs_bed* b = row_bed(op);
    if (op->missing_mode == 1 || op->prec == CS_PREC_INT8_L4) return false;   // int32 plain operand
    if (op->missing_mode == 0) {
        // Measured at n = 50k, m = 500k, k = 16: the genotype-only products take 1.3 + 1.2 ms
        // against 1.8 + 1.9 ms for the dense kernels, and each list gather reads one 64-byte row
        // per missing genotype per 16 columns (0.3 + 0.3 ms at 0.1 % missing, 1.0 + 1.4 ms at
        // 0.5 %): the apply takes 2.9 ms at 0 %, 3.35 ms at 0.1 % against 3.8 ms dense, so
        // break-even is near 0.2 % missing.  With k L <= 16 the dense products run with packed
        // limbs and are bandwidth-bound, so there is nothing to save.
        if (k * L <= 16) return false;
        const double frac = (double)missing_total(b) / ((double)b->N * (double)b->M);
        if (!(frac <= 0.0015)) return false;
    }
    ensure_missing_lists(b);
    return true;
}

// Device workspace of one dense-form int8 apply with k columns (kv_work_layout gives the offsets
// of its regions in one caller-provided block, each aligned to 256 bytes).
struct KvWork {
    double* Vt = nullptr;              // Npad x k scattered V
    int8_t* BV = nullptr;              // V limbs (B operand of Z^T V)
    unsigned long long* mx = nullptr;  // k column maxima
    double* scl = nullptr;             // 2 k scales and inverses
    double *W1 = nullptr, *W2 = nullptr;   // Mpad x k epilogue outputs w, mu w
    unsigned long long* mxW = nullptr; // k column maxima of W1 / W2
    int8_t *B1 = nullptr, *B2 = nullptr;   // W limbs (B operands of Z W)
    long long* Yint = nullptr;         // Npad x k exact int64 result rows
};

// Byte offsets of the KvWork regions for k columns; returns the total size.
static size_t kv_work_layout(int64_t Npad, int64_t Mpad, int64_t k, const Tiling& tl, size_t off[10]) {
    const size_t sz[10] = {(size_t)(Npad * k) * 8, (size_t)(Npad * tl.NT * 16), (size_t)k * 8, (size_t)(2 * k) * 8,
                           (size_t)(Mpad * k) * 8, (size_t)(Mpad * k) * 8, (size_t)k * 8,
                           (size_t)(Mpad * tl.NT * 16), (size_t)(Mpad * tl.NT * 16), (size_t)(Npad * k) * 8};
    size_t t = 0;
    for (int i = 0; i < 10; ++i) {
        off[i] = t;
        t += (sz[i] + 255) / 256 * 256;
    }
    return t;
}

static KvWork kv_work_at(unsigned char* base, const size_t off[10]) {
    KvWork w;
    w.Vt = (double*)(base + off[0]);
    w.BV = (int8_t*)(base + off[1]);
    w.mx = (unsigned long long*)(base + off[2]);
    w.scl = (double*)(base + off[3]);
    w.W1 = (double*)(base + off[4]);
    w.W2 = (double*)(base + off[5]);
    w.mxW = (unsigned long long*)(base + off[6]);
    w.B1 = (int8_t*)(base + off[7]);
    w.B2 = (int8_t*)(base + off[8]);
    w.Yint = (long long*)(base + off[9]);
    return w;
}

// One int8 apply.  ws == NULL: the operator's own workspaces (grown on demand) and the
// missing-genotype form of its mode; otherwise the dense form entirely in the caller's workspace,
// with no allocation (both forms give bitwise identical results).
// Modes of apply_int8: the full apply K X, Z^T V alone (PlinkReader::multiply_Zt_v), Z U alone
// (multiply_Z_v).
enum KvMode { KV_APPLY = 0, KV_ZT = 1, KV_Z = 2 };

// U (M x k column-major, ldu) from the row-major [Mpad][k] epilogue output
__global__ void k_rows_to_cols(const double* __restrict__ Wr, int64_t M, int64_t k, double* __restrict__ U,
                               int64_t ldu) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < M * k; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = e / k, c = e - j * k;
        U[j + c * ldu] = Wr[e];
    }
}

// Forward operands of Z U: W1[j][c] = s_j U[j][c], W2 = mu_j W1 (zero for inv_std == 0 and the
// padding rows), and their per-column maxima.
__global__ void k_w_from_u(const double* __restrict__ U, int64_t ldu, int64_t M, int64_t Mpad, int64_t k,
                           const double* __restrict__ means, const double* __restrict__ inv_stds,
                           double* __restrict__ W1, double* __restrict__ W2, unsigned long long* __restrict__ mx) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < Mpad * k; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = e / k, c = e - j * k;
        double w = 0.0, w2 = 0.0;
        if (j < M && inv_stds[j] != 0.0) {
            w = inv_stds[j] * U[j + c * ldu];
            w2 = means[j] * w;
        }
        W1[e] = w;
        W2[e] = w2;
        const double m = fmax(fabs(w), fabs(w2));
        if (m > 0.0) atomicMax(&mx[c], (unsigned long long)__double_as_longlong(m));
    }
}

static void apply_int8(cs_kinship* op, const double* dX, int64_t ldx, int64_t k, double* dY,
                       int64_t ldy, const KvWork* ws = nullptr, int mode = KV_APPLY) {
    // A last partial group of r columns with r L <= 16 runs as its own apply with packed limbs
    // (one MFMA tile, bandwidth-bound) instead of a dense group of L tiles.  Every column's
    // quantisation scale and integer products are its own, so the results are bitwise those of
    // one apply of all k columns.
    if (!ws && mode == KV_APPLY && k > 16) {
        const int64_t r = k % 16;
        if (r > 0 && r * prec_limbs(op->prec) <= 16) {
            apply_int8(op, dX, ldx, k - r, dY, ldy, nullptr, mode);
            apply_int8(op, dX + (k - r) * ldx, ldx, r, dY + (k - r) * ldy, ldy, nullptr, mode);
            return;
        }
    }
    cs_bed* b = row_bed(op);
    const bool direct = rows_direct(op);
    hipStream_t st = op->stream;
    const int L = prec_limbs(op->prec);
    const double qmax = limbs_qmax(L);
    const bool sparse = (ws || mode != KV_APPLY) ? false : use_sparse_missing(op, k, L);
    const Tiling tl = tiling(k, L, !sparse);   // the genotype-only kernels take unpacked operands
    const int64_t K16 = 16 * (int64_t)tl.G;
    const int32_t* inv_off = direct ? nullptr : op->d_inv_off.as<int32_t>();
    const int32_t* inv_idx = direct ? nullptr : op->d_inv_idx.as<int32_t>();
    KvWork w;
    if (ws) {
        w = *ws;
    } else {
        op->Vt.ensure((size_t)(b->Npad * k) * sizeof(double));
        op->BV.ensure((size_t)(b->Npad * tl.NT * 16));
        op->mx.ensure((size_t)k * sizeof(unsigned long long));
        op->scl.ensure((size_t)2 * k * sizeof(double));
        op->W1.ensure((size_t)(b->Mpad * k) * sizeof(double));
        op->W2.ensure((size_t)(b->Mpad * k) * sizeof(double));
        op->mxW.ensure((size_t)k * 8);
        op->B1.ensure((size_t)(b->Mpad * tl.NT * 16));
        op->Yint.ensure((size_t)(b->Npad * k) * sizeof(long long));
        if (!sparse) op->B2.ensure((size_t)(b->Mpad * tl.NT * 16));
        w.Vt = op->Vt.as<double>();
        w.BV = op->BV.as<int8_t>();
        w.mx = op->mx.as<unsigned long long>();
        w.scl = op->scl.as<double>();
        w.W1 = op->W1.as<double>();
        w.W2 = op->W2.as<double>();
        w.mxW = op->mxW.as<unsigned long long>();
        w.B1 = op->B1.as<int8_t>();
        w.B2 = sparse ? nullptr : op->B2.as<int8_t>();
        w.Yint = op->Yint.as<long long>();
    }

    CS_HIP(hipEventRecord(op->ev[0], st));
    if (mode == KV_Z) {
        // U given (dX: M x k): the forward operands directly, then the second half
        CS_HIP(hipMemsetAsync(w.mxW, 0, (size_t)k * 8, st));
        k_w_from_u<<<grid_for(b->Mpad * k), 256, 0, st>>>(dX, ldx, b->M, b->Mpad, k, b->means.as<double>(),
                                                         b->inv_stds.as<double>(), w.W1, w.W2, w.mxW);
        CS_HIP(hipGetLastError());
        CS_HIP(hipEventRecord(op->ev[1], st));
        CS_HIP(hipEventRecord(op->ev[2], st));
    } else {
    k_scatter<<<grid_for(b->Npad * k), 256, 0, st>>>(dX, ldx, k, inv_off, inv_idx, b->N, b->Npad, w.Vt);
    CS_HIP(hipGetLastError());
    {
        const double* Ws[1] = {w.Vt};
        int8_t* Bs[1] = {w.BV};
        quantize_cols_into(Ws, 1, b->Npad, k, tl, st, w.mx, w.scl, Bs);
    }
    if (sparse) {
        // plain quantised V, its column sums, and O^T V per SNP from the missing lists
        op->Vp.ensure((size_t)(b->Npad * K16) * 4);
        op->S1.ensure((size_t)K16 * 8);
        op->qext.ensure((size_t)(b->Mpad * K16) * 8);
        CS_HIP(hipMemsetAsync(op->Vp.p, 0, (size_t)(b->Npad * K16) * 4, st));
        k_quant<<<grid_for(b->Npad * k), 256, 0, st>>>(w.Vt, b->Npad, k, L, tl.NT, qmax, w.scl + k, nullptr,
                                                      op->Vp.as<int32_t>(), K16);
        colsum_plain(op->Vp.as<int32_t>(), b->Npad, K16, op->S1.as<long long>(), st);
        const unsigned gq = (unsigned)std::min<int64_t>(4096, (b->M + 127) / 128);   // 4 waves x 32 rows
        const int64_t* qo = b->miss_snp_off.as<int64_t>();
        const int32_t* qi = b->miss_snp_idx.as<int32_t>();
        switch (list_quads(K16)) {
            case 4: k_sparse_q<4, 32, 32><<<gq, 256, 0, st>>>(qo, qi, op->Vp.as<int32_t>(), K16, op->S1.as<long long>(), b->M, op->qext.as<long long>()); break;
            case 8: k_sparse_q<8, 32, 32><<<gq, 256, 0, st>>>(qo, qi, op->Vp.as<int32_t>(), K16, op->S1.as<long long>(), b->M, op->qext.as<long long>()); break;
            case 16: k_sparse_q<16, 32, 32><<<gq, 256, 0, st>>>(qo, qi, op->Vp.as<int32_t>(), K16, op->S1.as<long long>(), b->M, op->qext.as<long long>()); break;
            case 32: k_sparse_q<32, 32, 64><<<gq, 256, 0, st>>>(qo, qi, op->Vp.as<int32_t>(), K16, op->S1.as<long long>(), b->M, op->qext.as<long long>()); break;
            default: k_sparse_q<64, 32, 64><<<gq, 256, 0, st>>>(qo, qi, op->Vp.as<int32_t>(), K16, op->S1.as<long long>(), b->M, op->qext.as<long long>()); break;
        }
        CS_HIP(hipGetLastError());
    }
    CS_HIP(hipEventRecord(op->ev[1], st));

    // Z^T V with the fused epilogue: W1 = w, W2 = mu w, and their per-column maxima.
    CS_HIP(hipMemsetAsync(w.mxW, 0, (size_t)k * 8, st));
    GenoArgs a{};
    a.A = b->snp.as<uint8_t>();
    a.KB = b->KBs;
    a.rows = b->M;
    a.kb_total = a.kb_per_split = b->KBs;
    a.G = tl.G;
    a.Bx = a.Bo = w.BV;
    a.k = k;
    a.scale = w.scl;
    a.means = b->means.as<double>();
    a.inv_stds = b->inv_stds.as<double>();
    a.mult = op->d_mult.as<int32_t>();
    a.inv_M = 1.0 / (double)op->Msel;
    a.W1 = w.W1;
    a.W2 = w.W2;
    a.mx = w.mxW;
    a.qext = sparse ? op->qext.as<long long>() : nullptr;
    a.out_u = mode == KV_ZT ? 1 : 0;
    launch_geno(L, false, EPI_KV_GT, dim3((unsigned)(b->Mpad / 256), 1, (unsigned)tl.G), st, a, sparse, tl.packed);
    CS_HIP(hipEventRecord(op->ev[2], st));
    if (mode == KV_ZT) {
        // u over all SNPs into dY (M x k)
        k_rows_to_cols<<<grid_for(b->M * k), 256, 0, st>>>(w.W1, b->M, k, dY, ldy);
        CS_HIP(hipGetLastError());
        for (int e = 3; e < 6; ++e) CS_HIP(hipEventRecord(op->ev[e], st));
        return;
    }
    }

    // common per-column scale of W1 and W2, their limbs
    double* sclW = w.scl;  // the V scales are no longer needed
    k_scales<<<grid_for(k), 256, 0, st>>>(w.mxW, k, qmax, sclW, sclW + k);
    if (sparse) {
        CS_HIP(hipMemsetAsync(w.B1, 0, (size_t)(b->Mpad * tl.NT * 16), st));
        k_quant<<<grid_for(b->Mpad * k), 256, 0, st>>>(w.W1, b->Mpad, k, L, tl.NT, qmax, sclW + k, w.B1, nullptr, 0,
                                                      1.0, tl.packed);
        op->B2p.ensure((size_t)(b->Mpad * K16) * 4);
        op->S2.ensure((size_t)K16 * 8);
        CS_HIP(hipMemsetAsync(op->B2p.p, 0, (size_t)(b->Mpad * K16) * 4, st));
        k_quant<<<grid_for(b->Mpad * k), 256, 0, st>>>(w.W2, b->Mpad, k, L, tl.NT, qmax, sclW + k, nullptr,
                                                      op->B2p.as<int32_t>(), K16);
        colsum_plain(op->B2p.as<int32_t>(), b->Mpad, K16, op->S2.as<long long>(), st);
        CS_HIP(hipMemsetAsync(w.Yint, 0, (size_t)(b->Npad * k) * sizeof(long long), st));
    } else {
        // W1 and -(mu W) (negated: the Z W kernel accumulates X~ W1 and O (-mu W) into the same
        // registers), every byte of both, and the cleared result rows, in one launch
        const int cw = tl.packed ? 16 : 16 * tl.G;
        const unsigned gq = grid_for(b->Mpad / 16 * cw);
        const double *w1 = w.W1, *w2 = w.W2;
        int8_t *b1 = w.B1, *b2 = w.B2;
        long long* yi = w.Yint;
#define CS_QZW(LL, PP)                                                                                 \
        if (L == LL && tl.packed == PP)                                                                \
            k_quant_zw<LL, PP><<<gq, 256, 0, st>>>(w1, w2, b->Mpad, k, tl.NT, cw, qmax, sclW + k, b1, b2, yi, \
                                                   b->Npad * k);
        CS_QZW(2, false) CS_QZW(3, false) CS_QZW(4, false) CS_QZW(2, true) CS_QZW(3, true) CS_QZW(4, true)
#undef CS_QZW
    }
    CS_HIP(hipGetLastError());
    CS_HIP(hipEventRecord(op->ev[3], st));

    // Z W over the sample-major copy, split along the SNPs, exact int64 accumulation.
    const int64_t kbt = b->KBm;
    const int64_t rblocks = b->Npad / 256;
    // SNP splits: at least ~1536 workgroups, and a count that fills the last round of the
    // device's resident workgroups as fully as possible.
    // workgroups per split: one per (row block, group) in launch_one
    const int64_t per_split = rblocks * tl.G;
    static const int64_t cus = [] {
        int dev = 0, n = 256;
        if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev);
        return (int64_t)std::max(n, 1);
    }();
    // resident workgroups per CU: three for the LDS kernel, two for the four-limb one
    const int64_t slots = (L != 4 ? 3 : 2) * cus;
    // int32 accumulators: |x~ d1 + o d2| <= 2*128 + 128 per SNP, so a split holds at most
    // (2^31 - 1) / 384 SNPs (21845 blocks of 256)
    const int64_t kbps_max = ((int64_t)INT32_MAX / 384) / 256;
    const int64_t base = std::min<int64_t>(
        std::max<int64_t>(std::max<int64_t>(1, (1536 + per_split - 1) / per_split), (kbt + kbps_max - 1) / kbps_max),
        kbt);
    int64_t nsplit = base, kbps = (kbt + base - 1) / base;
    double best = -1.0;
    for (int64_t cand = base; cand <= std::min<int64_t>(kbt, 2 * base); ++cand) {
        const int64_t kp = (kbt + cand - 1) / cand, ns = (kbt + kp - 1) / kp;
        const int64_t blocks = per_split * ns;
        const double eff = (double)blocks / (double)(((blocks + slots - 1) / slots) * slots);
        if (eff > best + 1e-9) {
            best = eff;
            nsplit = ns;
            kbps = kp;
        }
    }
    GenoArgs c{};
    c.A = b->smp.as<uint8_t>();
    c.KB = b->KBm;
    c.rows = b->N;
    c.kb_total = kbt;
    c.kb_per_split = kbps;
    c.G = tl.G;
    c.Bx = w.B1;
    c.Bo = sparse ? w.B1 : w.B2;
    c.k = k;
    c.yint = w.Yint;
    if (sparse)
        launch_geno(L, false, EPI_KV_G, dim3((unsigned)rblocks, (unsigned)nsplit, (unsigned)tl.G), st, c, true);
    else
        launch_geno(L, true, EPI_KV_G, dim3((unsigned)rblocks, (unsigned)nsplit, (unsigned)tl.G), st, c, false,
                    tl.packed);
    CS_HIP(hipEventRecord(op->ev[4], st));
    if (sparse) {
        const unsigned gy = (unsigned)std::min<int64_t>(4096, (b->N + 31) / 32);     // 4 waves x 8 rows
        const int64_t* yo = b->miss_smp_off.as<int64_t>();
        const int32_t* yi = b->miss_smp_idx.as<int32_t>();
        switch (list_quads(K16)) {
            case 4: k_sparse_yadd<4, 8, 32><<<gy, 256, 0, st>>>(yo, yi, op->B2p.as<int32_t>(), K16, op->S2.as<long long>(), b->N, k, w.Yint); break;
            case 8: k_sparse_yadd<8, 8, 32><<<gy, 256, 0, st>>>(yo, yi, op->B2p.as<int32_t>(), K16, op->S2.as<long long>(), b->N, k, w.Yint); break;
            case 16: k_sparse_yadd<16, 8, 32><<<gy, 256, 0, st>>>(yo, yi, op->B2p.as<int32_t>(), K16, op->S2.as<long long>(), b->N, k, w.Yint); break;
            case 32: k_sparse_yadd<32, 8, 64><<<gy, 256, 0, st>>>(yo, yi, op->B2p.as<int32_t>(), K16, op->S2.as<long long>(), b->N, k, w.Yint); break;
            default: k_sparse_yadd<64, 8, 64><<<gy, 256, 0, st>>>(yo, yi, op->B2p.as<int32_t>(), K16, op->S2.as<long long>(), b->N, k, w.Yint); break;
        }
        CS_HIP(hipGetLastError());
    }
    k_finalize<<<grid_for(op->n * k), 256, 0, st>>>(w.Yint, k, sclW, direct ? nullptr : op->d_keep.as<int32_t>(),
                                                   op->n, dY, ldy);
    CS_HIP(hipGetLastError());
    CS_HIP(hipEventRecord(op->ev[5], st));
}

template <int KC>
static void apply_f32_cols(cs_kinship* op, int64_t k, double* dY, int64_t ldy) {
    cs_bed* b = row_bed(op);
    hipStream_t st = op->stream;
    constexpr int Qc = (KC + 3) / 4 * 4;
    op->Vf.ensure((size_t)(b->Npad * Qc) * sizeof(float));
    k_to_f32_cols<<<grid_for(b->Npad * Qc), 256, 0, st>>>(op->Vt.as<double>(), b->Npad, k, Qc, op->Vf.as<float>());
    CS_HIP(hipEventRecord(op->ev[1], st));
    op->Wf.ensure((size_t)(b->Mpad * 2 * Qc) * sizeof(float));
    float* W12 = op->Wf.as<float>();   // per SNP [W1 | W2], Qc floats each
    op->PQf.ensure((size_t)(b->Mpad * KC) * sizeof(float));   // u = Z^T V, Mpad x KC
    k_f32_zt<KC><<<(unsigned)(b->Mpad / 256), 256, 0, st>>>(b->snp.as<uint8_t>(), b->KBs, b->M,
                                                           b->means.as<double>(), b->inv_stds.as<double>(),
                                                           reinterpret_cast<const f32x4n*>(op->Vf.as<float>()),
                                                           op->PQf.as<float>());
    CS_HIP(hipGetLastError());
    CS_HIP(hipEventRecord(op->ev[2], st));
    k_epi_w_f32<<<grid_for(b->Mpad * KC), 256, 0, st>>>(op->PQf.as<float>(), KC, b->M, b->Mpad, k,
                                                       b->means.as<double>(), b->inv_stds.as<double>(),
                                                       op->d_mult.as<int32_t>(), (float)(1.0 / (double)op->Msel), W12);
    CS_HIP(hipGetLastError());
    CS_HIP(hipEventRecord(op->ev[3], st));
    // SNP splits: about 2048 workgroups, at least 8 k-blocks each
    const int64_t rblocks = b->Npad / 256;
    const int64_t nsplit = std::max<int64_t>(1, std::min<int64_t>((2048 + rblocks - 1) / rblocks, b->KBm / 8));
    const int64_t kbps = (b->KBm + nsplit - 1) / nsplit;
    const int64_t ns = (b->KBm + kbps - 1) / kbps;
    op->Yf.ensure((size_t)(ns * b->Npad * KC) * sizeof(float));
    k_f32_z<KC><<<dim3((unsigned)rblocks, (unsigned)ns), 256, 0, st>>>(
        b->smp.as<uint8_t>(), b->KBm, kbps, b->Npad, reinterpret_cast<const f32x4n*>(W12), op->Yf.as<float>());
    CS_HIP(hipGetLastError());
    CS_HIP(hipEventRecord(op->ev[4], st));
    k_epi_y_f32<<<grid_for(op->n * k), 256, 0, st>>>(op->Yf.as<float>(), ns, b->Npad, KC, k,
                                                    rows_direct(op) ? nullptr : op->d_keep.as<int32_t>(), op->n,
                                                    dY, ldy);
    CS_HIP(hipGetLastError());
}

static void apply_f32(cs_kinship* op, const double* dX, int64_t ldx, int64_t k, double* dY,
                      int64_t ldy) {
    cs_bed* b = row_bed(op);
    hipStream_t st = op->stream;
    const int32_t* inv_off = rows_direct(op) ? nullptr : op->d_inv_off.as<int32_t>();
    const int32_t* inv_idx = rows_direct(op) ? nullptr : op->d_inv_idx.as<int32_t>();
    for (int64_t c0 = 0; c0 < k; c0 += 16) {
        const int64_t kc = std::min<int64_t>(16, k - c0);
        CS_HIP(hipEventRecord(op->ev[0], st));
        op->Vt.ensure((size_t)(b->Npad * kc) * sizeof(double));
        k_scatter<<<grid_for(b->Npad * kc), 256, 0, st>>>(dX + c0 * ldx, ldx, kc, inv_off, inv_idx,
                                                          b->N, b->Npad, op->Vt.as<double>());
        if (kc <= 1) apply_f32_cols<1>(op, kc, dY + c0 * ldy, ldy);
        else if (kc <= 4) apply_f32_cols<4>(op, kc, dY + c0 * ldy, ldy);
        else apply_f32_cols<16>(op, kc, dY + c0 * ldy, ldy);
        CS_HIP(hipEventRecord(op->ev[5], st));
    }
}

// Fixed workspace of cs_kinship_apply_device_noalloc for callers that cannot allocate (code under
// graph capture, serving loops, harnesses that forbid hipMalloc).  Each device holds its own copy
// from module load until the process ends.  In host code the array's name is not a device
// address: kernels and copies take the per-device pointer from hipGetSymbolAddress.
constexpr size_t kKvScratchBytes = (size_t)256 << 20;
static __device__ __align__(256) unsigned char scratch[kKvScratchBytes];

// The workspace's address on the current device, looked up once per device (first call there).
static unsigned char* scratch_ptr() {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static unsigned char* ptrs[kMaxDevices] = {};
    int dev = 0;
    CS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw RuntimeError("cs_kinship_apply_device_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!ptrs[dev]) {
        void* p = nullptr;
        CS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw RuntimeError("cs_kinship_apply_device_noalloc: no device address for the workspace");
        ptrs[dev] = (unsigned char*)p;
    }
    return ptrs[dev];
}

static void record_timings(cs_kinship* op) {
    for (int p = 0; p < 5; ++p) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, op->ev[p], op->ev[p + 1]) != hipSuccess) ms = 0.f;
        op->last_ms[p] = ms;
    }
}

static void apply_device_impl(cs_kinship* op, const double* dX, int64_t ldx, int64_t k, double* dY,
                              int64_t ldy) {
    if (k < 0) throw RuntimeError("PlinkKinshipOperator::apply: dimension mismatch");
    if (k == 0) return;
    if (ldx < op->n || ldy < op->n) throw RuntimeError("PlinkKinshipOperator::apply: dimension mismatch");
    if (op->prec == CS_PREC_FP32) apply_f32(op, dX, ldx, k, dY, ldy);
    else apply_int8(op, dX, ldx, k, dY, ldy);
    CS_HIP(hipStreamSynchronize(op->stream));
    record_timings(op);
}

}  // namespace csamd

using namespace csamd;

extern "C" {

int cs_kinship_create(cs_bed* bed, const int32_t* keep, int64_t n_keep, const int32_t* variants,
                      int64_t n_variants, cs_kinship** out) {
    return guarded([&] {
        if (!out) throw InvalidArgument("cs_kinship_create: null output handle");
        if (!bed) throw RuntimeError("PlinkKinshipOperator: null reader");
        const int64_t N = bed->N, M = bed->M;
        auto* op = new cs_kinship();
        try {
            op->bed = bed;
            if (keep && n_keep > 0) {
                op->keep.assign(keep, keep + n_keep);
                for (int32_t s : op->keep)
                    if (s < 0 || s >= N) throw RuntimeError("PlinkKinshipOperator: sample index out of range");
                bool id = n_keep == N;
                for (int64_t i = 0; id && i < N; ++i) id = op->keep[i] == i;
                op->identity = id;
            }
            if (variants && n_variants > 0) {
                op->variants.assign(variants, variants + n_variants);
                for (int32_t j : op->variants)
                    if (j < 0 || j >= M) throw RuntimeError("PlinkKinshipOperator: variant index out of range");
            }
            op->n = op->keep.empty() ? N : (int64_t)op->keep.size();
            op->Msel = op->variants.empty() ? M : (int64_t)op->variants.size();
            if (op->Msel == 0) throw RuntimeError("PlinkKinshipOperator: no variants selected");
            ensure_stats(bed);          // reader_->computeStats(1) (kinship_operator.cpp:67)
            ensure_sample_major(bed);
            CS_HIP(hipStreamCreateWithFlags(&op->stream, hipStreamNonBlocking));
            for (auto& e : op->ev) CS_HIP(hipEventCreate(&e));
            std::vector<int32_t> mult((size_t)M, op->variants.empty() ? 1 : 0);
            for (int32_t j : op->variants) mult[(size_t)j] += 1;
            op->d_mult.alloc((size_t)M * sizeof(int32_t));
            CS_HIP(hipMemcpy(op->d_mult.p, mult.data(), (size_t)M * sizeof(int32_t), hipMemcpyHostToDevice));
            if (!op->identity) {
                op->d_keep.alloc((size_t)op->n * sizeof(int32_t));
                CS_HIP(hipMemcpy(op->d_keep.p, op->keep.data(), (size_t)op->n * sizeof(int32_t),
                                 hipMemcpyHostToDevice));
                std::vector<int32_t> off((size_t)N + 1, 0), idx((size_t)op->n);
                for (int32_t s : op->keep) off[(size_t)s + 1]++;
                for (int64_t s = 0; s < N; ++s) off[(size_t)s + 1] += off[(size_t)s];
                std::vector<int32_t> fill(off.begin(), off.end() - 1);
                for (int64_t r = 0; r < op->n; ++r) idx[(size_t)fill[(size_t)op->keep[(size_t)r]]++] = (int32_t)r;
                op->d_inv_off.alloc(off.size() * sizeof(int32_t));
                op->d_inv_idx.alloc(idx.size() * sizeof(int32_t));
                CS_HIP(hipMemcpy(op->d_inv_off.p, off.data(), off.size() * sizeof(int32_t), hipMemcpyHostToDevice));
                CS_HIP(hipMemcpy(op->d_inv_idx.p, idx.data(), idx.size() * sizeof(int32_t), hipMemcpyHostToDevice));
            }
        } catch (...) {
            if (op->stream) (void)hipStreamDestroy(op->stream);
            for (auto& e : op->ev) if (e) (void)hipEventDestroy(e);
            delete op;
            throw;
        }
        *out = op;
    });
}

int cs_kinship_destroy(cs_kinship* op) {
    return guarded([&] {
        if (!op) return;
        if (op->stream) (void)hipStreamSynchronize(op->stream);
        for (auto& e : op->ev) if (e) (void)hipEventDestroy(e);
        hipStream_t s = op->stream;
        cs_bed* rows = op->rows;
        delete op;
        if (rows) bed_free(rows);
        if (s) (void)hipStreamDestroy(s);
    });
}

int cs_kinship_compact_rows(cs_kinship* op) {
    return guarded([&] {
        if (!op) throw InvalidArgument("cs_kinship_compact_rows: null handle");
        kinship_compact_rows(op);
    });
}

int cs_kinship_device_bytes(const cs_kinship* op, int64_t* bytes) {
    return guarded([&] {
        if (!op || !bytes) throw InvalidArgument("cs_kinship_device_bytes: null argument");
        int64_t t = 0;
        for (const DevBuf* b : {&op->d_keep, &op->d_inv_off, &op->d_inv_idx, &op->d_mult, &op->dX, &op->dY, &op->Vt,
                                &op->BV, &op->W1, &op->W2, &op->B1, &op->B2, &op->Yint, &op->mxW, &op->mx, &op->scl,
                                &op->Vf, &op->PQf, &op->Wf, &op->Yf, &op->Vp, &op->B2p, &op->S1, &op->S2, &op->qext,
                                &op->diag_tab, &op->diag_full})
            t += (int64_t)b->bytes;
        if (op->rows) t += bed_rows_copy_bytes(op->bed, op->n);
        *bytes = t;
    });
}

int cs_kinship_rows(const cs_kinship* op, int64_t* rows) {
    return guarded([&] {
        if (!op || !rows) throw InvalidArgument("cs_kinship_rows: null argument");
        *rows = op->n;
    });
}

int cs_kinship_num_variants(const cs_kinship* op, int64_t* m) {
    return guarded([&] {
        if (!op || !m) throw InvalidArgument("cs_kinship_num_variants: null argument");
        *m = op->Msel;
    });
}

int cs_kinship_set_precision(cs_kinship* op, int mode) {
    return guarded([&] {
        if (!op) throw InvalidArgument("cs_kinship_set_precision: null handle");
        if (mode != CS_PREC_INT8_L3 && mode != CS_PREC_INT8_L2 && mode != CS_PREC_INT8_L4 && mode != CS_PREC_FP32)
            throw InvalidArgument("cs_kinship_set_precision: unknown mode");
        op->prec = mode;
    });
}

}  // extern "C"

int csamd::kinship_get_precision(const cs_kinship* op) { return op->prec; }

// The compact copy holds the cohort's rows in the operator's order, so V needs no scatter and
// K V no gather.  Every per-column quantity of the products is unchanged: V's column maxima (the
// scattered copy only adds zero rows), the exact integer sums over the samples, W's per-SNP
// values and maxima, and the per-row results.  With repeated samples the scatter would sum V's
// rows before quantising, so such cohorts stay on the full-sample products.
void csamd::kinship_compact_rows(cs_kinship* op) {
    if (op->identity || op->rows) return;
    std::vector<uint8_t> seen((size_t)op->bed->N, 0);
    for (int32_t s : op->keep)
        if (seen[(size_t)s]++) throw InvalidArgument("PlinkKinshipOperator: a compact cohort lists a sample twice");
    CS_HIP(hipStreamSynchronize(op->stream));
    op->rows = bed_rows_copy(op->bed, op->keep.data(), op->n);
    kinship_release_workspace(op);   // sized for the full-sample rows
}

bool csamd::kinship_is_compact(const cs_kinship* op) { return op->rows != nullptr; }

cs_bed* csamd::kinship_rows_bed(cs_kinship* op) { return op->rows; }

int64_t csamd::kinship_work_bytes(int64_t Npad, int64_t Mpad, int prec, int64_t k) {
    if (k <= 0) return 0;
    const int L = prec_limbs(prec);
    const Tiling tl = tiling(k, L, true), tu = tiling(k, L, false);
    size_t off[10];
    // dense form, plus what the sparse missing form adds (plain operands, sums, per-SNP terms)
    const int64_t K16 = 16 * (int64_t)tu.G;
    int64_t t = (int64_t)std::max(kv_work_layout(Npad, Mpad, k, tl, off), kv_work_layout(Npad, Mpad, k, tu, off));
    t += Npad * K16 * 4 + Mpad * K16 * 12 + 2 * K16 * 8;
    // FP32 comparator: V and W in FP32, u, and the per-split partial rows (about 2048 workgroups)
    if (prec == CS_PREC_FP32) t += Npad * 16 * 4 + Mpad * 16 * 12 + std::max<int64_t>(2 * 2048 * 256, 2 * Npad) * 16 * 4;
    return t;
}

// The K.V apply enqueued on the caller's stream, with no wait on the host: the serving loop puts
// every active request's iteration on one stream and waits once per step.  (The FP32 comparator
// precision runs synchronously on the operator's stream.)
void csamd::kinship_apply_async(cs_kinship* op, const double* dX, int64_t ldx, int64_t k, double* dY, int64_t ldy,
                                hipStream_t st) {
    if (op->prec == CS_PREC_FP32) {
        CS_HIP(hipStreamSynchronize(st));
        apply_device_impl(op, dX, ldx, k, dY, ldy);
        return;
    }
    if (k < 0 || (k > 0 && (ldx < op->n || ldy < op->n)))
        throw RuntimeError("PlinkKinshipOperator::apply: dimension mismatch");
    if (k == 0) return;
    hipStream_t own = op->stream;
    op->stream = st;
    try {
        apply_int8(op, dX, ldx, k, dY, ldy);
    } catch (...) {
        op->stream = own;
        throw;
    }
    op->stream = own;
}

void csamd::kinship_release_workspace(cs_kinship* op) {
    CS_HIP(hipStreamSynchronize(op->stream));
    for (DevBuf* b : {&op->dX, &op->dY, &op->Vt, &op->BV, &op->W1, &op->W2, &op->B1, &op->B2, &op->Yint, &op->mxW,
                      &op->mx, &op->scl, &op->Vf, &op->PQf, &op->Wf, &op->Yf, &op->Vp, &op->B2p, &op->S1, &op->S2,
                      &op->qext})
        b->release();
}

extern "C" {

int cs_kinship_set_missing_mode(cs_kinship* op, int mode) {
    return guarded([&] {
        if (!op) throw InvalidArgument("cs_kinship_set_missing_mode: null handle");
        if (mode < 0 || mode > 2) throw InvalidArgument("cs_kinship_set_missing_mode: mode must be 0, 1 or 2");
        op->missing_mode = mode;
    });
}

int cs_kinship_apply_device(cs_kinship* op, const double* dX, int64_t ldx, int64_t k, double* dY,
                            int64_t ldy) {
    return guarded([&] {
        if (!op) throw InvalidArgument("cs_kinship_apply_device: null handle");
        if (k > 0 && (!dX || !dY)) throw InvalidArgument("cs_kinship_apply_device: null buffer");
        apply_device_impl(op, dX, ldx, k, dY, ldy);
    });
}

int cs_kinship_apply_device_noalloc(cs_kinship* op, const double* dX, int64_t ldx, int64_t k, double* dY,
                                    int64_t ldy) {
    return guarded([&] {
        if (!op) throw InvalidArgument("cs_kinship_apply_device_noalloc: null handle");
        if (k < 0 || (k > 0 && (ldx < op->n || ldy < op->n)))
            throw RuntimeError("PlinkKinshipOperator::apply: dimension mismatch");
        if (k == 0) return;
        if (!dX || !dY) throw InvalidArgument("cs_kinship_apply_device_noalloc: null buffer");
        if (op->prec == CS_PREC_FP32)
            throw InvalidArgument("cs_kinship_apply_device_noalloc: the FP32 comparator precision is not supported");
        const Tiling tl = tiling(k, prec_limbs(op->prec), true);
        size_t off[10];
        const size_t need = kv_work_layout(row_bed(op)->Npad, row_bed(op)->Mpad, k, tl, off);
        if (need > kKvScratchBytes)
            throw InvalidArgument("cs_kinship_apply_device_noalloc: " + std::to_string(need) +
                                  " bytes of workspace needed for this operator and k, the fixed workspace holds " +
                                  std::to_string(kKvScratchBytes));
        const KvWork w = kv_work_at(scratch_ptr(), off);
        apply_int8(op, dX, ldx, k, dY, ldy, &w);
        CS_HIP(hipStreamSynchronize(op->stream));
        record_timings(op);
    });
}

// PlinkReader::multiply_Zt_v / multiply_Z_v (plink_reader.cpp:241-438) over the operator's rows
// and all SNPs: the first or second half of the int8 apply.
static void multiply_impl(cs_kinship* op, int mode, const double* dIn, int64_t ldi, int64_t k, double* dOut,
                          int64_t ldo) {
    if (op->prec == CS_PREC_FP32) throw InvalidArgument("multiply_Z: the FP32 comparator precision is not supported");
    const int64_t rin = mode == KV_ZT ? op->n : op->bed->M, rout = mode == KV_ZT ? op->bed->M : op->n;
    if (k < 0 || (k > 0 && (ldi < rin || ldo < rout)))
        throw RuntimeError(mode == KV_ZT ? "Dimension mismatch in multiply_Zt_v" : "Dimension mismatch in multiply_Z_v");
    if (k == 0) return;
    ensure_stats(op->bed);
    apply_int8(op, dIn, ldi, k, dOut, ldo, nullptr, mode);
    CS_HIP(hipStreamSynchronize(op->stream));
}

static void multiply_host(cs_kinship* op, int mode, const double* In, int64_t ldi, int64_t k, double* Out,
                          int64_t ldo) {
    const int64_t rin = mode == KV_ZT ? op->n : op->bed->M, rout = mode == KV_ZT ? op->bed->M : op->n;
    if (k < 0 || (k > 0 && (ldi < rin || ldo < rout)))
        throw RuntimeError(mode == KV_ZT ? "Dimension mismatch in multiply_Zt_v" : "Dimension mismatch in multiply_Z_v");
    if (k == 0) return;
    if (!In || !Out) throw InvalidArgument("multiply_Z: null buffer");
    op->dX.ensure((size_t)(rin * k) * sizeof(double));
    op->dY.ensure((size_t)(rout * k) * sizeof(double));
    CS_HIP(hipMemcpy2DAsync(op->dX.p, (size_t)rin * 8, In, (size_t)ldi * 8, (size_t)rin * 8, (size_t)k,
                            hipMemcpyHostToDevice, op->stream));
    multiply_impl(op, mode, op->dX.as<double>(), rin, k, op->dY.as<double>(), rout);
    CS_HIP(hipMemcpy2DAsync(Out, (size_t)ldo * 8, op->dY.p, (size_t)rout * 8, (size_t)rout * 8, (size_t)k,
                            hipMemcpyDeviceToHost, op->stream));
    CS_HIP(hipStreamSynchronize(op->stream));
}

int cs_kinship_multiply_zt(cs_kinship* op, const double* V, int64_t ldv, int64_t k, double* U, int64_t ldu) {
    return guarded([&] {
        if (!op) throw InvalidArgument("cs_kinship_multiply_zt: null handle");
        multiply_host(op, KV_ZT, V, ldv, k, U, ldu);
    });
}

int cs_kinship_multiply_z(cs_kinship* op, const double* U, int64_t ldu, int64_t k, double* Y, int64_t ldy) {
    return guarded([&] {
        if (!op) throw InvalidArgument("cs_kinship_multiply_z: null handle");
        multiply_host(op, KV_Z, U, ldu, k, Y, ldy);
    });
}

int cs_kinship_multiply_zt_device(cs_kinship* op, const double* dV, int64_t ldv, int64_t k, double* dU,
                                  int64_t ldu) {
    return guarded([&] {
        if (!op) throw InvalidArgument("cs_kinship_multiply_zt_device: null handle");
        if (k > 0 && (!dV || !dU)) throw InvalidArgument("cs_kinship_multiply_zt_device: null buffer");
        multiply_impl(op, KV_ZT, dV, ldv, k, dU, ldu);
    });
}

int cs_kinship_multiply_z_device(cs_kinship* op, const double* dU, int64_t ldu, int64_t k, double* dY,
                                 int64_t ldy) {
    return guarded([&] {
        if (!op) throw InvalidArgument("cs_kinship_multiply_z_device: null handle");
        if (k > 0 && (!dU || !dY)) throw InvalidArgument("cs_kinship_multiply_z_device: null buffer");
        multiply_impl(op, KV_Z, dU, ldu, k, dY, ldy);
    });
}

int cs_kinship_apply(cs_kinship* op, const double* X, int64_t ldx, int64_t k, double* Y,
                     int64_t ldy) {
    return guarded([&] {
        if (!op) throw InvalidArgument("cs_kinship_apply: null handle");
        if (k < 0 || (k > 0 && (ldx < op->n || ldy < op->n)))
            throw RuntimeError("PlinkKinshipOperator::apply: dimension mismatch");
        if (k == 0) return;
        if (!X || !Y) throw InvalidArgument("cs_kinship_apply: null buffer");
        const int64_t n = op->n;
        op->dX.ensure((size_t)(n * k) * sizeof(double));
        op->dY.ensure((size_t)(n * k) * sizeof(double));
        CS_HIP(hipMemcpy2DAsync(op->dX.p, (size_t)n * sizeof(double), X, (size_t)ldx * sizeof(double),
                                (size_t)n * sizeof(double), (size_t)k, hipMemcpyHostToDevice, op->stream));
        apply_device_impl(op, op->dX.as<double>(), n, k, op->dY.as<double>(), n);
        CS_HIP(hipMemcpy2DAsync(Y, (size_t)ldy * sizeof(double), op->dY.p, (size_t)n * sizeof(double),
                                (size_t)n * sizeof(double), (size_t)k, hipMemcpyDeviceToHost, op->stream));
        CS_HIP(hipStreamSynchronize(op->stream));
    });
}

int cs_kinship_diag(cs_kinship* op, double* diag) {
    return guarded([&] {
        if (!op || !diag) throw InvalidArgument("cs_kinship_diag: null argument");
        if (!op->diag_cache.empty()) {
            std::copy(op->diag_cache.begin(), op->diag_cache.end(), diag);
            return;
        }
        cs_bed* b = op->bed;
        hipStream_t st = op->stream;
        op->diag_tab.ensure((size_t)b->M * 4 * sizeof(double));
        op->diag_full.ensure((size_t)b->Npad * sizeof(double));
        k_diag_table<<<grid_for(b->M), 256, 0, st>>>(b->M, b->means.as<double>(), b->inv_stds.as<double>(),
                                                    op->d_mult.as<int32_t>(), op->diag_tab.as<double>());
        k_diag<<<(unsigned)(b->Npad / 256), 256, 0, st>>>(b->smp.as<uint8_t>(), b->KBm, b->N, b->M,
                                                         op->diag_tab.as<double>(),
                                                         1.0 / (double)op->Msel, op->diag_full.as<double>());
        CS_HIP(hipGetLastError());
        std::vector<double> full((size_t)b->N);
        CS_HIP(hipMemcpyAsync(full.data(), op->diag_full.p, (size_t)b->N * sizeof(double),
                              hipMemcpyDeviceToHost, st));
        CS_HIP(hipStreamSynchronize(st));
        for (int64_t r = 0; r < op->n; ++r) diag[r] = full[(size_t)(op->keep.empty() ? r : op->keep[(size_t)r])];
        op->diag_cache.assign(diag, diag + op->n);
    });
}

// trace(K) as the reference sums it (kinship_operator.cpp:306-343): per selected SNP the sum of
// z^2 over the kept samples, then over the SNPs.  The per-SNP sum comes from the class counts of
// the kept samples (keep multiplicities as weights): sum_c n_c(j) (x_c - mu_j)^2 s_j^2.
int cs_kinship_trace(cs_kinship* op, double* trace) {
    return guarded([&] {
        if (!op || !trace) throw InvalidArgument("cs_kinship_trace: null argument");
        cs_bed* b = op->bed;
        hipStream_t st = op->stream;
        std::vector<int32_t> w((size_t)b->N, 0);
        if (op->keep.empty()) std::fill(w.begin(), w.end(), 1);
        else
            for (int32_t s : op->keep) w[(size_t)s] += 1;
        DevBuf counts((size_t)b->M * 4 * sizeof(int64_t)), terms((size_t)b->M * sizeof(double));
        weighted_counts(b, w.data(), (int64_t)w.size(), counts.as<int64_t>(), st);
        k_trace_terms<<<grid_for(b->M), 256, 0, st>>>(counts.as<int64_t>(), b->M, b->means.as<double>(),
                                                     b->inv_stds.as<double>(), op->d_mult.as<int32_t>(),
                                                     terms.as<double>());
        CS_HIP(hipGetLastError());
        std::vector<double> t((size_t)b->M);
        CS_HIP(hipMemcpyAsync(t.data(), terms.p, (size_t)b->M * sizeof(double), hipMemcpyDeviceToHost, st));
        CS_HIP(hipStreamSynchronize(st));
        double total = 0.0;
        for (double v : t) total += v;
        *trace = total / (double)op->Msel;
    });
}

int cs_kinship_last_timing(const cs_kinship* op, double* ms5) {
    return guarded([&] {
        if (!op || !ms5) throw InvalidArgument("cs_kinship_last_timing: null argument");
        for (int p = 0; p < 5; ++p) ms5[p] = op->last_ms[p];
    });
}

}  // extern "C"
