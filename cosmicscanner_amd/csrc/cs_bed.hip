// Resident packed genotypes: upload / synthetic generation, padded SNP-major layout, the
// sample-major transpose, and kernel (a) -- decode + per-SNP statistics.
//
// HBM layout.  Two resident copies in MFMA tiles (cs_geno_mfma.h): SNP-major (row j = SNP,
// sample i at byte i>>2, bits 2(i&3) of the row -- the .bed order, plink_reader.cpp:84) and
// sample-major (row i = sample, SNP j at byte j>>2, bits 2(j&3)).  Samples are padded to
// Npad = roundup(N, 256) and SNPs to Mpad = roundup(M, 256); every padding slot holds code 01
// (missing), so it contributes to no sum.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "cs_common.h"
#include "cs_geno_mfma.h"
#include "cs_ops.h"

namespace csamd {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

// Plain streaming copy (16 bytes per lane), the reference kernel for HBM byte counters.
__global__ __launch_bounds__(256) void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                int64_t n16) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n16;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// ------------------------------------------------------------------------------ relayout
// Copy `rows` raw rows (Braw bytes each, starting at SNP r0) into the tiled SNP-major matrix;
// slots of samples >= N in the last byte become code 01.  Everything else was pre-filled 0x55.
__global__ void k_relayout(const uint8_t* __restrict__ raw, int64_t r0, int64_t rows, int64_t Braw,
                           int64_t N, uint8_t* __restrict__ dst, int64_t KB) {
    const int64_t total = rows * Braw;
    const int tail = (int)(N & 3);
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e / Braw, b = e - r * Braw;
        uint8_t v = raw[e];
        if (tail && b == Braw - 1) {
            const uint8_t keep_mask = (uint8_t)((1u << (2 * tail)) - 1u);
            v = (uint8_t)((v & keep_mask) | (kMissingByte & ~keep_mask));
        }
        dst[tiled_offset(r0 + r, b, KB)] = v;
    }
}

// Inverse for download: tiled SNP-major -> raw rows, padding slots of the last byte zeroed.
__global__ void k_unlayout(const uint8_t* __restrict__ src, int64_t KB, int64_t rows, int64_t Braw,
                           int64_t N, uint8_t* __restrict__ raw) {
    const int64_t total = rows * Braw;
    const int tail = (int)(N & 3);
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e / Braw, b = e - r * Braw;
        uint8_t v = src[tiled_offset(r, b, KB)];
        if (tail && b == Braw - 1) v &= (uint8_t)((1u << (2 * tail)) - 1u);
        raw[e] = v;
    }
}

// The same for a list of SNPs: row i of raw is SNP snps[i].
__global__ void k_unlayout_snps(const uint8_t* __restrict__ src, int64_t KB, const int32_t* __restrict__ snps,
                                int64_t ns, int64_t Braw, int64_t N, uint8_t* __restrict__ raw) {
    const int64_t total = ns * Braw;
    const int tail = (int)(N & 3);
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / Braw, b = e - i * Braw;
        uint8_t v = src[tiled_offset(snps[i], b, KB)];
        if (tail && b == Braw - 1) v &= (uint8_t)((1u << (2 * tail)) - 1u);
        raw[e] = v;
    }
}

// ------------------------------------------------------------------------------ synthetic
// Counter-based HWE generator: byte-identical to the oracle's CPU generator (same integer
// hash, same IEEE operations; the library is compiled with -ffp-contract=off).
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t hash3(uint64_t seed, uint64_t j, uint64_t i) {
    return mix64(mix64(mix64(seed + 0x9e3779b97f4a7c15ULL) ^ (j * 0xd1b54a32d192ed03ULL)) ^
                 (i * 0x8cb92ba72f3d8dd7ULL + 0x632be59bd9b4e019ULL));
}
__device__ __forceinline__ double u01(uint64_t h) { return (double)(h >> 11) * 0x1.0p-53; }

__global__ void k_synth(uint64_t seed, int64_t N, int64_t M, double pmin, double pmax,
                        double miss, uint8_t* __restrict__ dst, int64_t KB) {
    const int64_t Braw = (N + 3) / 4;
    const int64_t total = M * Braw;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = e / Braw, b = e - j * Braw;
        const double p = pmin + (pmax - pmin) * u01(hash3(seed, (uint64_t)j, ~0ULL));
        const double q = 1.0 - p;
        const double p0 = q * q;
        const double p1 = p0 + 2.0 * p * q;
        uint8_t byte = 0;
        for (int s = 0; s < 4; ++s) {
            const int64_t i = b * 4 + s;
            int code = 1;  // padding slots: missing
            if (i < N) {
                const uint64_t h = hash3(seed, (uint64_t)j, (uint64_t)i);
                if (u01(h) < miss) {
                    code = 1;
                } else {
                    const double v = u01(mix64(h ^ 0xa0761d6478bd642fULL));
                    const int x = v < p0 ? 0 : (v < p1 ? 1 : 2);
                    code = x == 2 ? 0 : (x == 1 ? 2 : 3);
                }
            }
            byte |= (uint8_t)(code << (2 * s));
        }
        dst[tiled_offset(j, b, KB)] = byte;
    }
}

// ------------------------------------------------------------------------------ transpose
// Tiled SNP-major (Mpad x Npad) -> tiled sample-major (Npad x Mpad), 256 x 256 genotype blocks
// through LDS.  blockIdx.x: 256-sample block, blockIdx.y: 256-SNP block.
__global__ __launch_bounds__(256) void k_transpose(const uint8_t* __restrict__ snp, int64_t KBs,
                                                   uint8_t* __restrict__ smp, int64_t KBm) {
    __shared__ uint32_t tile[256][17];  // 256 SNP rows x 64 bytes (+1 dword pad vs conflicts)
    const int t = threadIdx.x;
    const int64_t j0 = (int64_t)blockIdx.y * 256;   // SNPs
    const int64_t s0 = (int64_t)blockIdx.x * 256;   // samples
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint4 v = *reinterpret_cast<const uint4*>(snp + tiled_offset(j0 + t, s0 / 4 + 16 * q, KBs));
        tile[t][4 * q] = v.x;
        tile[t][4 * q + 1] = v.y;
        tile[t][4 * q + 2] = v.z;
        tile[t][4 * q + 3] = v.w;
    }
    __syncthreads();
    // thread t -> sample s0 + t; output dword q covers SNPs j0 + 16q .. 16q+15
    const int dw = t >> 4, sh = (t & 15) * 2;   // sample t lives in dword t/16, bits 2(t%16)
    uint32_t out[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        uint32_t word = 0;
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const uint32_t code = (tile[q * 16 + u][dw] >> sh) & 3u;
            word |= code << (2 * u);
        }
        out[q] = word;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
        *reinterpret_cast<uint4*>(smp + tiled_offset(s0 + t, j0 / 4 + 16 * q, KBm)) =
            make_uint4(out[4 * q], out[4 * q + 1], out[4 * q + 2], out[4 * q + 3]);
}

// ------------------------------------------------------------------------------ kernel (a)
// All-sample statistics.  One wave per 64-SNP row tile streams the tile sequence (the access
// pattern of the MFMA kernels); each lane counts the three observed codes of rows 16m + r over
// its 16-byte chunk g, the four chunk lanes of a row are reduced by shuffles, and the
// reference's FP64 formulae (plink_reader.cpp:205-234) run with correctly rounded IEEE division
// and square root and no contraction, so values are bit-identical to the CPU restatement.
__global__ __launch_bounds__(256) void k_stats(const uint8_t* __restrict__ snp, int64_t KB,
                                               int64_t M, int64_t N, int64_t* __restrict__ counts,
                                               double* __restrict__ means,
                                               double* __restrict__ inv_stds) {
    const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
    const int64_t tr = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (tr * 64 >= M) return;
    const uint8_t* base = snp + ((tr * KB) << 12) + lane * 16;
    uint32_t c00[4] = {0, 0, 0, 0}, c10[4] = {0, 0, 0, 0}, c11[4] = {0, 0, 0, 0};
    for (int64_t kb = 0; kb < KB; ++kb) {
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const uint4 w4 = *reinterpret_cast<const uint4*>(base + (kb << 12) + m * 1024);
            const uint32_t ws[4] = {w4.x, w4.y, w4.z, w4.w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t L = ws[q] & 0x55555555u, H = (ws[q] >> 1) & 0x55555555u;
                c00[m] += __popc(~(L | H) & 0x55555555u);
                c10[m] += __popc(H & ~L & 0x55555555u);
                c11[m] += __popc(H & L);
            }
        }
    }
#pragma unroll
    for (int m = 0; m < 4; ++m)
        for (int off = 16; off < 64; off <<= 1) {
            c00[m] += __shfl_xor(c00[m], off);
            c10[m] += __shfl_xor(c10[m], off);
            c11[m] += __shfl_xor(c11[m], off);
        }
    if (g != 0) return;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int64_t j = tr * 64 + 16 * m + r;
        if (j >= M) continue;
        const int64_t obs = (int64_t)c00[m] + c10[m] + c11[m];
        counts[j * 4 + 0] = c00[m];
        counts[j * 4 + 1] = N - obs;
        counts[j * 4 + 2] = c10[m];
        counts[j * 4 + 3] = c11[m];
        if (obs > 0) {
            const double sum = 2.0 * (double)c00[m] + (double)c10[m];
            const double mean = sum / (double)obs;
            const double p = mean / 2.0;
            const double var = 2.0 * p * (1.0 - p);
            means[j] = mean;
            inv_stds[j] = (var < 1e-9) ? 0.0 : 1.0 / sqrt(var);
        } else {
            means[j] = 0.0;
            inv_stds[j] = 0.0;
        }
    }
}

static void init_layout(cs_bed* b, int64_t N, int64_t M) {
    if (N <= 0 || M <= 0) throw InvalidArgument("cs_bed: n_samples and n_snps must be positive");
    CS_HIP(hipGetDevice(&b->device));
    b->N = N;
    b->M = M;
    b->Braw = (N + 3) / 4;
    b->Npad = round_up(N, kPad);
    b->Mpad = round_up(M, kPad);
    b->KBs = b->Npad / 256;
    b->KBm = b->Mpad / 256;
    CS_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    b->snp.alloc((size_t)(b->Mpad * b->Npad / 4));
    CS_HIP(hipMemsetAsync(b->snp.p, kMissingByte, b->snp.bytes, b->stream));
}

static void upload(cs_bed* b, const uint8_t* body) {
    // Staged in chunks of whole rows through a pinned buffer and a device staging area.
    const int64_t chunk_rows = std::max<int64_t>(1, (int64_t)(256ll << 20) / b->Braw);
    const size_t chunk_bytes = (size_t)(chunk_rows * b->Braw);
    DevBuf stage(chunk_bytes);
    void* pinned = nullptr;
    CS_HIP(hipHostMalloc(&pinned, chunk_bytes, hipHostMallocDefault));
    try {
        for (int64_t r0 = 0; r0 < b->M; r0 += chunk_rows) {
            const int64_t rows = std::min(chunk_rows, b->M - r0);
            const size_t nb = (size_t)(rows * b->Braw);
            CS_HIP(hipStreamSynchronize(b->stream));
            std::memcpy(pinned, body + r0 * b->Braw, nb);
            CS_HIP(hipMemcpyAsync(stage.p, pinned, nb, hipMemcpyHostToDevice, b->stream));
            k_relayout<<<grid_for(rows * b->Braw), 256, 0, b->stream>>>(
                stage.as<uint8_t>(), r0, rows, b->Braw, b->N, b->snp.as<uint8_t>(), b->KBs);
            CS_HIP(hipGetLastError());
        }
        CS_HIP(hipStreamSynchronize(b->stream));
    } catch (...) {
        (void)hipHostFree(pinned);
        throw;
    }
    CS_HIP(hipHostFree(pinned));
}

// Row gather of a tiled matrix: destination row r (< rows_pad) takes source row keep[r] when
// r < n, code 01 (missing) beyond.  One lane per 16-byte chunk, in destination address order:
// chunk e = ((rt KB + kb) << 8) + (m 64 + g 16 + rr) holds k-block kb, chunk g of row 64 rt + 16 m + rr.
__global__ __launch_bounds__(256) void k_gather_rows(const uint4* __restrict__ src, int64_t KB,
                                                     const int32_t* __restrict__ keep, int64_t n,
                                                     int64_t rows_pad, uint4* __restrict__ dst) {
    const int64_t chunks = rows_pad / 64 * KB * 256;
    const uint32_t miss = 0x55555555u;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < chunks;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t tile = e >> 8;
        const int w = (int)(e & 255);
        const int64_t rt = tile / KB, kb = tile - rt * KB;
        const int64_t row = rt * 64 + ((w >> 6) << 4) + (w & 15);
        uint4 v = make_uint4(miss, miss, miss, miss);
        if (row < n) {
            const int64_t s = keep[row];
            v = src[(((s >> 6) * KB + kb) << 8) + (((s >> 4) & 3) << 6) + (w & 48) + (s & 15)];
        }
        dst[e] = v;
    }
}

void ensure_sample_major(cs_bed* b) {
    if (b->smp.p) return;
    b->smp.alloc((size_t)(b->Npad * b->Mpad / 4));
    dim3 grid((unsigned)(b->Npad / 256), (unsigned)(b->Mpad / 256));
    k_transpose<<<grid, 256, 0, b->stream>>>(b->snp.as<uint8_t>(), b->KBs, b->smp.as<uint8_t>(), b->KBm);
    CS_HIP(hipGetLastError());
    CS_HIP(hipStreamSynchronize(b->stream));
}

void ensure_stats(cs_bed* b) {
    if (b->have_stats) return;
    b->counts.alloc((size_t)b->M * 4 * sizeof(int64_t));
    b->means.alloc((size_t)b->M * sizeof(double));
    b->inv_stds.alloc((size_t)b->M * sizeof(double));
    k_stats<<<(unsigned)(b->Mpad / 256), 256, 0, b->stream>>>(
        b->snp.as<uint8_t>(), b->KBs, b->M, b->N, b->counts.as<int64_t>(), b->means.as<double>(),
        b->inv_stds.as<double>());
    CS_HIP(hipGetLastError());
    CS_HIP(hipStreamSynchronize(b->stream));
    b->have_stats = true;
}

// Weighted class counts from popcounts.  The weights (non-negative multiplicities) are split into
// binary planes, each packed like a genotype row (bit 2i set where sample i's plane bit is 1);
// for a 16-genotype word with low bits lo = w & 0x55.. and high bits hi = (w >> 1) & 0x55..:
// n11 = popc(hi & lo & pm), n10 = popc(hi & pm) - n11, n01 = popc(lo & pm) - n11,
// n00 = popc(pm) - n01 - n10 - n11, each weighted by 2^plane.
// tiles (optional): the 64-SNP row tiles to count, ntiles of them (wave w of block b takes entry
// 4 b + w); without a list every row tile is counted.
__global__ __launch_bounds__(256) void k_wcounts_pc(const uint8_t* __restrict__ snp, int64_t KB, int64_t M,
                                                    const uint32_t* __restrict__ planes, int nplanes,
                                                    int64_t* __restrict__ counts, const int32_t* __restrict__ tiles,
                                                    int64_t ntiles) {
    const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
    int64_t tr = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (tiles) {
        if (tr >= ntiles) return;
        tr = tiles[tr];
    }
    if (tr * 64 >= M) return;
    const uint8_t* base = snp + ((tr * KB) << 12) + lane * 16;
    long long c[4][4];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int t = 0; t < 4; ++t) c[m][t] = 0;
    for (int64_t kb = 0; kb < KB; ++kb) {
        uint4 w4[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) w4[m] = *reinterpret_cast<const uint4*>(base + (kb << 12) + m * 1024);
        for (int p = 0; p < nplanes; ++p) {
            const uint4 pm4 = *reinterpret_cast<const uint4*>(planes + ((p * KB + kb) * 4 + g) * 4);
            const uint32_t pm[4] = {pm4.x, pm4.y, pm4.z, pm4.w};
            int tot = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) tot += __builtin_popcount(pm[q]);
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                int h = 0, l = 0, hl = 0;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const uint32_t wd = q == 0 ? w4[m].x : q == 1 ? w4[m].y : q == 2 ? w4[m].z : w4[m].w;
                    const uint32_t lo = wd & 0x55555555u, hi = (wd >> 1) & 0x55555555u;
                    h += __builtin_popcount(hi & pm[q]);
                    l += __builtin_popcount(lo & pm[q]);
                    hl += __builtin_popcount(hi & lo & pm[q]);
                }
                const long long sc = 1LL << p;
                c[m][3] += (long long)hl * sc;
                c[m][2] += (long long)(h - hl) * sc;
                c[m][1] += (long long)(l - hl) * sc;
                c[m][0] += (long long)(tot - h - l + hl) * sc;
            }
        }
    }
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int t = 0; t < 4; ++t)
            for (int off = 16; off < 64; off <<= 1) c[m][t] += __shfl_xor(c[m][t], off);
    if (g != 0) return;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int64_t j = tr * 64 + 16 * m + r;
        if (j >= M) continue;
#pragma unroll
        for (int t = 0; t < 4; ++t) counts[j * 4 + t] = c[m][t];
    }
}

// Weight planes of the class-count kernel: planes[p][kb][g][q], word q of chunk g of k-block kb
// covers samples 256 kb + 64 g + 16 q + u; returns the plane count (0: every weight is 0).
static int weight_planes(cs_bed* b, const int32_t* w, int64_t n_w, std::vector<uint32_t>& planes) {
    if (n_w > b->Npad) throw InvalidArgument("class counts: more weights than samples");
    int32_t wmax = 0;
    for (int64_t i = 0; i < n_w; ++i) {
        if (w[i] < 0) throw InvalidArgument("class counts: negative weight");
        wmax = std::max(wmax, w[i]);
    }
    int np = 0;
    while (np < 31 && (wmax >> np) > 0) ++np;
    const int64_t KB = b->KBs;
    planes.assign((size_t)(np * KB * 16), 0u);
    for (int64_t i = 0; i < n_w; ++i) {
        const int32_t wi = w[i];
        if (!wi) continue;
        const int64_t kb = i >> 8, rr = i & 255, g = rr >> 6, q = (rr >> 4) & 3, u = rr & 15;
        for (int p = 0; p < np; ++p)
            if ((wi >> p) & 1) planes[(size_t)(((p * KB + kb) * 4 + g) * 4 + q)] |= 1u << (2 * u);
    }
    return np;
}

void weighted_counts(cs_bed* b, const int32_t* w, int64_t n_w, int64_t* d_counts, hipStream_t st) {
    std::vector<uint32_t> planes;
    const int np = weight_planes(b, w, n_w, planes);
    if (np == 0) {
        CS_HIP(hipMemsetAsync(d_counts, 0, (size_t)b->M * 32, st));
        return;
    }
    DevBuf dp(planes.size() * 4);
    CS_HIP(hipMemcpyAsync(dp.p, planes.data(), planes.size() * 4, hipMemcpyHostToDevice, st));
    k_wcounts_pc<<<(unsigned)(b->Mpad / 256), 256, 0, st>>>(b->snp.as<uint8_t>(), b->KBs, b->M, dp.as<uint32_t>(), np,
                                                           d_counts, nullptr, 0);
    CS_HIP(hipGetLastError());
    CS_HIP(hipStreamSynchronize(st));
}

__global__ void k_gather_counts(const int64_t* __restrict__ counts, const int32_t* __restrict__ snps, int64_t ns,
                                int64_t* __restrict__ out) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e < ns * 4) out[e] = counts[(int64_t)snps[e >> 2] * 4 + (e & 3)];
}

void weighted_counts_of(cs_bed* b, const int32_t* w, int64_t n_w, const int32_t* snps, int64_t ns, int64_t* h_counts,
                        hipStream_t st) {
    if (ns <= 0) return;
    std::vector<uint32_t> planes;
    const int np = weight_planes(b, w, n_w, planes);
    if (np == 0) {
        std::fill(h_counts, h_counts + ns * 4, (int64_t)0);
        return;
    }
    std::vector<int32_t> tiles;
    for (int64_t i = 0; i < ns; ++i) tiles.push_back(snps[i] / 64);
    std::sort(tiles.begin(), tiles.end());
    tiles.erase(std::unique(tiles.begin(), tiles.end()), tiles.end());
    const int64_t nt = (int64_t)tiles.size();
    DevBuf dp(planes.size() * 4), dt((size_t)nt * 4), dsn((size_t)ns * 4), counts((size_t)b->M * 32),
        sel((size_t)ns * 32);
    CS_HIP(hipMemcpyAsync(dp.p, planes.data(), planes.size() * 4, hipMemcpyHostToDevice, st));
    CS_HIP(hipMemcpyAsync(dt.p, tiles.data(), (size_t)nt * 4, hipMemcpyHostToDevice, st));
    CS_HIP(hipMemcpyAsync(dsn.p, snps, (size_t)ns * 4, hipMemcpyHostToDevice, st));
    k_wcounts_pc<<<(unsigned)((nt + 3) / 4), 256, 0, st>>>(b->snp.as<uint8_t>(), b->KBs, b->M, dp.as<uint32_t>(), np,
                                                          counts.as<int64_t>(), dt.as<int32_t>(), nt);
    k_gather_counts<<<grid_for(ns * 4), 256, 0, st>>>(counts.as<int64_t>(), dsn.as<int32_t>(), ns, sel.as<int64_t>());
    CS_HIP(hipGetLastError());
    CS_HIP(hipMemcpyAsync(h_counts, sel.p, (size_t)ns * 32, hipMemcpyDeviceToHost, st));
    CS_HIP(hipStreamSynchronize(st));
}

// Missing-genotype lists.  One wave per row of a tiled matrix (SNP-major: rows are SNPs and the
// elements samples; sample-major: the reverse).  Lane l of an iteration covers k-block 16 it + l/4,
// chunk l % 4: 64 consecutive elements, so lane order is element order.
__global__ __launch_bounds__(256) void k_miss_count(const uint8_t* __restrict__ A, int64_t KB, int64_t rows,
                                                   int64_t limit, int32_t* __restrict__ cnt) {
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= rows) return;
    const int64_t rt = row >> 6;
    const int m = (int)((row & 63) >> 4), r = (int)(row & 15);
    int c = 0;
    for (int64_t kb0 = 0; kb0 < KB; kb0 += 16) {
        const int64_t kb = kb0 + (lane >> 2);
        const int g = lane & 3;
        if (kb < KB) {
            const uint4 w4 = *reinterpret_cast<const uint4*>(A + ((rt * KB + kb) << 12) + ((m * 64 + g * 16 + r) << 4));
            const uint32_t ws[4] = {w4.x, w4.y, w4.z, w4.w};
            const int64_t e0 = kb * 256 + g * 64;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                // code 01 <=> low bit set and high bit clear
                const uint32_t w = ws[t];
                uint32_t miss = (w & 0x55555555u) & ~((w >> 1) & 0x55555555u);
                const int64_t base = e0 + 16 * t;
                if (base + 16 > limit) {
                    const int64_t keep = limit - base;
                    miss = keep <= 0 ? 0u : (keep >= 16 ? miss : (miss & ((1u << (2 * keep)) - 1u)));
                }
                c += __builtin_popcount(miss);
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if (lane == 0) cnt[row] = c;
}

__global__ __launch_bounds__(256) void k_miss_fill(const uint8_t* __restrict__ A, int64_t KB, int64_t rows,
                                                  int64_t limit, const int64_t* __restrict__ off,
                                                  int32_t* __restrict__ idx) {
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= rows) return;
    const int64_t rt = row >> 6;
    const int m = (int)((row & 63) >> 4), r = (int)(row & 15);
    int64_t pos = off[row];
    for (int64_t kb0 = 0; kb0 < KB; kb0 += 16) {
        const int64_t kb = kb0 + (lane >> 2);
        const int g = lane & 3;
        uint32_t ws[4] = {0u, 0u, 0u, 0u};
        const int64_t e0 = kb * 256 + g * 64;
        if (kb < KB) {
            const uint4 w4 = *reinterpret_cast<const uint4*>(A + ((rt * KB + kb) << 12) + ((m * 64 + g * 16 + r) << 4));
            ws[0] = w4.x; ws[1] = w4.y; ws[2] = w4.z; ws[3] = w4.w;
        }
        uint32_t miss[4];
        int c = 0;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            uint32_t mm = (ws[t] & 0x55555555u) & ~((ws[t] >> 1) & 0x55555555u);
            const int64_t base = e0 + 16 * t;
            if (kb >= KB) mm = 0u;
            else if (base + 16 > limit) {
                const int64_t keep = limit - base;
                mm = keep <= 0 ? 0u : (keep >= 16 ? mm : (mm & ((1u << (2 * keep)) - 1u)));
            }
            miss[t] = mm;
            c += __builtin_popcount(mm);
        }
        // exclusive prefix over the wave
        int x = c;
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(x, d);
            if (lane >= d) x += y;
        }
        const int total = __shfl(x, 63);
        int64_t p = pos + (x - c);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            uint32_t mm = miss[t];
            while (mm) {
                const int bit = __builtin_ctz(mm);
                idx[p++] = (int32_t)(e0 + 16 * t + (bit >> 1));
                mm &= mm - 1u;
            }
        }
        pos += total;
    }
}

static void build_lists(const uint8_t* A, int64_t KB, int64_t rows, int64_t limit, DevBuf& off, DevBuf& idx,
                        int64_t& nnz, hipStream_t st) {
    DevBuf cnt((size_t)std::max<int64_t>(1, rows) * 4);
    const unsigned nb = (unsigned)((rows + 3) / 4);
    k_miss_count<<<nb, 256, 0, st>>>(A, KB, rows, limit, cnt.as<int32_t>());
    CS_HIP(hipGetLastError());
    std::vector<int32_t> hc((size_t)rows);
    CS_HIP(hipMemcpyAsync(hc.data(), cnt.p, (size_t)rows * 4, hipMemcpyDeviceToHost, st));
    CS_HIP(hipStreamSynchronize(st));
    std::vector<int64_t> ho((size_t)rows + 1, 0);
    for (int64_t i = 0; i < rows; ++i) ho[(size_t)i + 1] = ho[(size_t)i] + hc[(size_t)i];
    nnz = ho[(size_t)rows];
    off.alloc(ho.size() * 8);
    idx.alloc((size_t)std::max<int64_t>(1, nnz) * 4);
    CS_HIP(hipMemcpyAsync(off.p, ho.data(), ho.size() * 8, hipMemcpyHostToDevice, st));
    k_miss_fill<<<nb, 256, 0, st>>>(A, KB, rows, limit, off.as<int64_t>(), idx.as<int32_t>());
    CS_HIP(hipGetLastError());
    CS_HIP(hipStreamSynchronize(st));
}

// Sum of the per-SNP missing counts (class 01 of the M x 4 counts), one int64 atomic per block.
__global__ __launch_bounds__(256) void k_sum_missing(const int64_t* __restrict__ counts, int64_t M,
                                                     unsigned long long* __restrict__ total) {
    __shared__ long long red[256];
    long long acc = 0;
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < M; j += (int64_t)gridDim.x * blockDim.x)
        acc += counts[j * 4 + 1];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(total, (unsigned long long)red[0]);
}

int64_t missing_total(cs_bed* b) {
    if (b->miss_total >= 0) return b->miss_total;
    ensure_stats(b);
    DevBuf t(8);
    CS_HIP(hipMemsetAsync(t.p, 0, 8, b->stream));
    const unsigned nb = (unsigned)std::max<int64_t>(1, std::min<int64_t>(1024, (b->M + 255) / 256));
    k_sum_missing<<<nb, 256, 0, b->stream>>>(b->counts.as<int64_t>(), b->M, t.as<unsigned long long>());
    CS_HIP(hipGetLastError());
    int64_t h = 0;
    CS_HIP(hipMemcpyAsync(&h, t.p, 8, hipMemcpyDeviceToHost, b->stream));
    CS_HIP(hipStreamSynchronize(b->stream));
    b->miss_total = h;
    return h;
}

int64_t bed_rows_copy_bytes(const cs_bed* parent, int64_t n) {
    const int64_t npad = round_up(n, kPad);
    return 2 * (npad * parent->Mpad / 4) + parent->M * 48;
}

cs_bed* bed_rows_copy(cs_bed* parent, const int32_t* keep, int64_t n) {
    if (n <= 0) throw InvalidArgument("cohort copy: empty cohort");
    for (int64_t r = 0; r < n; ++r)
        if (keep[r] < 0 || keep[r] >= parent->N) throw RuntimeError("cohort copy: sample index out of range");
    ensure_stats(parent);
    ensure_sample_major(parent);
    auto* b = new cs_bed();
    try {
        CS_HIP(hipGetDevice(&b->device));
        b->N = n;
        b->M = parent->M;
        b->Braw = (n + 3) / 4;
        b->Npad = round_up(n, kPad);
        b->Mpad = parent->Mpad;
        b->KBs = b->Npad / 256;
        b->KBm = parent->KBm;
        CS_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        DevBuf dkeep((size_t)n * 4);
        CS_HIP(hipMemcpyAsync(dkeep.p, keep, (size_t)n * 4, hipMemcpyHostToDevice, b->stream));
        b->smp.alloc((size_t)(b->Npad * b->Mpad / 4));
        k_gather_rows<<<grid_for(b->Npad / 64 * b->KBm * 256), 256, 0, b->stream>>>(
            parent->smp.as<uint4>(), b->KBm, dkeep.as<int32_t>(), n, b->Npad, b->smp.as<uint4>());
        CS_HIP(hipGetLastError());
        // SNP-major copy by the tiled transpose (input rows: the cohort's samples)
        b->snp.alloc((size_t)(b->Mpad * b->Npad / 4));
        k_transpose<<<dim3((unsigned)(b->Mpad / 256), (unsigned)(b->Npad / 256)), 256, 0, b->stream>>>(
            b->smp.as<uint8_t>(), b->KBm, b->snp.as<uint8_t>(), b->KBs);
        CS_HIP(hipGetLastError());
        // the cohort's own class counts (missing fraction), the parent's all-sample standardisation
        ensure_stats(b);
        CS_HIP(hipMemcpyAsync(b->means.p, parent->means.p, (size_t)b->M * 8, hipMemcpyDeviceToDevice, b->stream));
        CS_HIP(hipMemcpyAsync(b->inv_stds.p, parent->inv_stds.p, (size_t)b->M * 8, hipMemcpyDeviceToDevice,
                              b->stream));
        CS_HIP(hipStreamSynchronize(b->stream));
    } catch (...) {
        if (b->stream) {
            (void)hipStreamSynchronize(b->stream);
            (void)hipStreamDestroy(b->stream);
        }
        delete b;
        throw;
    }
    return b;
}

void bed_free(cs_bed* b) {
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    hipStream_t s = b->stream;
    delete b;
    if (s) (void)hipStreamDestroy(s);
}

void ensure_missing_lists(cs_bed* b) {
    if (b->miss_nnz >= 0) return;
    ensure_sample_major(b);
    int64_t n1 = 0, n2 = 0;
    build_lists(b->snp.as<uint8_t>(), b->KBs, b->M, b->N, b->miss_snp_off, b->miss_snp_idx, n1, b->stream);
    build_lists(b->smp.as<uint8_t>(), b->KBm, b->N, b->M, b->miss_smp_off, b->miss_smp_idx, n2, b->stream);
    if (n1 != n2) throw RuntimeError("missing-genotype lists disagree between the two layouts");
    b->miss_nnz = n1;
}

}  // namespace csamd

using namespace csamd;

extern "C" {

const char* cs_last_error(void) { return g_last_error.c_str(); }
int cs_version(void) { return 100; }

int cs_dev_alloc(int64_t bytes, void** dptr) {
    return guarded([&] {
        if (!dptr || bytes < 0) throw InvalidArgument("cs_dev_alloc: bad argument");
        CS_HIP(hipMalloc(dptr, (size_t)std::max<int64_t>(bytes, 1)));
    });
}

int cs_dev_free(void* dptr) {
    return guarded([&] {
        if (dptr) CS_HIP(hipFree(dptr));
    });
}

int cs_dev_upload(void* dst, const void* src, int64_t bytes) {
    return guarded([&] {
        if (bytes > 0) CS_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
    });
}

int cs_dev_download(void* dst, const void* src, int64_t bytes) {
    return guarded([&] {
        if (bytes > 0) CS_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
    });
}

int cs_dev_copy(void* dst, const void* src, int64_t bytes) {
    return guarded([&] {
        if (bytes < 0 || bytes % 16) throw InvalidArgument("cs_dev_copy: bytes must be a multiple of 16");
        if (bytes == 0) return;
        k_copy16<<<8192, 256>>>(static_cast<const uint4*>(src), static_cast<uint4*>(dst), bytes / 16);
        CS_HIP(hipGetLastError());
        CS_HIP(hipDeviceSynchronize());
    });
}

int cs_dev_synchronize(void) {
    return guarded([&] { CS_HIP(hipDeviceSynchronize()); });
}

int cs_bed_create(const uint8_t* body, int64_t n_samples, int64_t n_snps, cs_bed** out) {
    return guarded([&] {
        if (!out) throw InvalidArgument("cs_bed_create: null output handle");
        if (!body) throw RuntimeError("cs_bed_create: null genotype buffer");
        auto* b = new cs_bed();
        try {
            init_layout(b, n_samples, n_snps);
            upload(b, body);
        } catch (...) {
            if (b->stream) (void)hipStreamDestroy(b->stream);
            delete b;
            throw;
        }
        *out = b;
    });
}

int cs_bed_create_from_file_image(const uint8_t* image, int64_t image_bytes, int64_t n_samples,
                                  int64_t n_snps, cs_bed** out) {
    return guarded([&] {
        if (!image || image_bytes < 3) throw RuntimeError("Invalid .bed file (too small)");
        if (image[0] != 0x6C || image[1] != 0x1B) throw RuntimeError("Invalid .bed file magic number");
        if (image[2] != 0x01)
            throw RuntimeError("Only SNP-major mode (0x01) .bed files are supported");
        const int64_t need = 3 + n_snps * ((n_samples + 3) / 4);
        if (image_bytes < need) throw RuntimeError("Invalid .bed file (truncated)");
        int rc = cs_bed_create(image + 3, n_samples, n_snps, out);
        if (rc != CS_OK) throw RuntimeError(cs_last_error());
    });
}

int cs_bed_create_synthetic(uint64_t seed, int64_t n_samples, int64_t n_snps, double pmin,
                            double pmax, double miss, cs_bed** out) {
    return guarded([&] {
        if (!out) throw InvalidArgument("cs_bed_create_synthetic: null output handle");
        if (!(pmin >= 0.0 && pmax <= 1.0 && pmin <= pmax) || !(miss >= 0.0 && miss < 1.0))
            throw InvalidArgument("cs_bed_create_synthetic: bad frequency or missing rate");
        auto* b = new cs_bed();
        try {
            init_layout(b, n_samples, n_snps);
            const int64_t total = n_snps * ((n_samples + 3) / 4);
            k_synth<<<grid_for(total), 256, 0, b->stream>>>(seed, n_samples, n_snps, pmin, pmax, miss,
                                                           b->snp.as<uint8_t>(), b->KBs);
            CS_HIP(hipGetLastError());
            CS_HIP(hipStreamSynchronize(b->stream));
        } catch (...) {
            if (b->stream) (void)hipStreamDestroy(b->stream);
            delete b;
            throw;
        }
        *out = b;
    });
}

int cs_bed_destroy(cs_bed* bed) {
    return guarded([&] {
        if (!bed) return;
        bed_free(bed);
        dev_cache_release();   // the blocks of this matrix's work are sized for it
    });
}

int cs_dev_release_cache(void) {
    return guarded([&] { dev_cache_release(); });
}

int cs_bed_dims(const cs_bed* bed, int64_t* n_samples, int64_t* n_snps) {
    return guarded([&] {
        if (!bed) throw InvalidArgument("cs_bed_dims: null handle");
        if (n_samples) *n_samples = bed->N;
        if (n_snps) *n_snps = bed->M;
    });
}

int cs_bed_download(cs_bed* bed, uint8_t* body) {
    return guarded([&] {
        if (!bed || !body) throw InvalidArgument("cs_bed_download: null argument");
        DevBuf raw((size_t)(bed->M * bed->Braw));
        k_unlayout<<<grid_for(bed->M * bed->Braw), 256, 0, bed->stream>>>(
            bed->snp.as<uint8_t>(), bed->KBs, bed->M, bed->Braw, bed->N, raw.as<uint8_t>());
        CS_HIP(hipGetLastError());
        CS_HIP(hipMemcpyAsync(body, raw.p, (size_t)(bed->M * bed->Braw), hipMemcpyDeviceToHost,
                              bed->stream));
        CS_HIP(hipStreamSynchronize(bed->stream));
    });
}

int cs_bed_download_snps(cs_bed* bed, const int32_t* snps, int64_t n_snps, uint8_t* body) {
    return guarded([&] {
        if (!bed || (n_snps > 0 && (!snps || !body))) throw InvalidArgument("cs_bed_download_snps: null argument");
        for (int64_t i = 0; i < n_snps; ++i)
            if (snps[i] < 0 || snps[i] >= bed->M) throw RuntimeError("cs_bed_download_snps: variant index out of range");
        if (n_snps == 0) return;
        DevBuf dsn((size_t)n_snps * 4), raw((size_t)(n_snps * bed->Braw));
        CS_HIP(hipMemcpyAsync(dsn.p, snps, (size_t)n_snps * 4, hipMemcpyHostToDevice, bed->stream));
        k_unlayout_snps<<<grid_for(n_snps * bed->Braw), 256, 0, bed->stream>>>(
            bed->snp.as<uint8_t>(), bed->KBs, dsn.as<int32_t>(), n_snps, bed->Braw, bed->N, raw.as<uint8_t>());
        CS_HIP(hipGetLastError());
        CS_HIP(hipMemcpyAsync(body, raw.p, (size_t)(n_snps * bed->Braw), hipMemcpyDeviceToHost, bed->stream));
        CS_HIP(hipStreamSynchronize(bed->stream));
    });
}

int cs_bed_build_sample_major(cs_bed* bed) {
    return guarded([&] {
        if (!bed) throw InvalidArgument("cs_bed_build_sample_major: null handle");
        ensure_sample_major(bed);
    });
}

int cs_bed_device_bytes(const cs_bed* bed, int64_t* bytes) {
    return guarded([&] {
        if (!bed || !bytes) throw InvalidArgument("cs_bed_device_bytes: null argument");
        *bytes = (int64_t)(bed->snp.bytes + bed->smp.bytes + bed->counts.bytes + bed->means.bytes +
                           bed->inv_stds.bytes);
    });
}

int cs_bed_compute_stats(cs_bed* bed) {
    return guarded([&] {
        if (!bed) throw InvalidArgument("cs_bed_compute_stats: null handle");
        ensure_stats(bed);
    });
}

int cs_bed_get_stats(cs_bed* bed, double* means, double* inv_stds) {
    return guarded([&] {
        if (!bed) throw InvalidArgument("cs_bed_get_stats: null handle");
        ensure_stats(bed);
        if (means)
            CS_HIP(hipMemcpy(means, bed->means.p, (size_t)bed->M * sizeof(double), hipMemcpyDeviceToHost));
        if (inv_stds)
            CS_HIP(hipMemcpy(inv_stds, bed->inv_stds.p, (size_t)bed->M * sizeof(double),
                             hipMemcpyDeviceToHost));
    });
}

int cs_bed_class_counts(cs_bed* bed, const int32_t* weights, int64_t* counts) {
    return guarded([&] {
        if (!bed || !counts) throw InvalidArgument("cs_bed_class_counts: null argument");
        if (!weights) {
            ensure_stats(bed);
            CS_HIP(hipMemcpy(counts, bed->counts.p, (size_t)bed->M * 4 * sizeof(int64_t),
                             hipMemcpyDeviceToHost));
            return;
        }
        DevBuf c((size_t)bed->M * 4 * sizeof(int64_t));
        weighted_counts(bed, weights, bed->N, c.as<int64_t>(), bed->stream);
        CS_HIP(hipMemcpyAsync(counts, c.p, c.bytes, hipMemcpyDeviceToHost, bed->stream));
        CS_HIP(hipStreamSynchronize(bed->stream));
    });
}

}  // extern "C"
