/*
 * cs_oracle.c -- FP64 / exact-integer CPU restatement of the reference's genotype hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by nothing else.  The product
 * library (cosmicscanner_amd/) never links or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates.  BED pointers
 * here point PAST the three magic bytes, SNP j at bed + j*ceil(N/4) (plink_reader.cpp:84).
 *
 * Parity pinning (what checks each function beyond retyping): see the header of
 * oracle/__init__.py, which lists every function with the tests that pin it.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef __int128 i128;

/* Code map (plink_reader.cpp:11-22, kinship_operator.cpp:18-20):
 * 00 -> 2 copies of .bim column-5 allele (A1), 01 -> missing, 10 -> 1, 11 -> 0. */
static const double kDosage[4] = {2.0, 0.0, 1.0, 0.0};
static const int kDosI[4] = {2, 0, 1, 0};
static const int kObs[4] = {1, 0, 1, 1};

static inline int code_at(const uint8_t* col, int64_t s) {
    return (col[s >> 2] >> ((s & 3) * 2)) & 3;
}

static inline int64_t bytes_per_snp(int64_t N) { return (N + 3) / 4; }

/* ---------------------------------------------------------------------------------------
 * A3: PlinkReader::computeStats (plink_reader.cpp:187-239).  All N samples.
 * mean = sum/count; p = mean/2; var = 2p(1-p); var < 1e-9 -> inv_std 0; count 0 -> 0,0. */
void orc_stats(const uint8_t* bed, int64_t N, int64_t M, double* means, double* inv_stds,
               int nthreads) {
    const int64_t B = bytes_per_snp(N);
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t j = 0; j < M; ++j) {
        const uint8_t* col = bed + j * B;
        double sum = 0.0;
        int count = 0;
        for (int64_t i = 0; i < N; ++i) {
            int c = code_at(col, i);
            if (c == 1) continue;
            sum += kDosage[c];
            count++;
        }
        if (count > 0) {
            double mean = sum / count;
            double p = mean / 2.0;
            double var = 2.0 * p * (1.0 - p);
            means[j] = mean;
            inv_stds[j] = (var < 1e-9) ? 0.0 : 1.0 / sqrt(var);
        } else {
            means[j] = 0.0;
            inv_stds[j] = 0.0;
        }
    }
}

/* Weighted per-SNP class counts: cnt[j*4 + code] = sum_s w_s [code(s,j) == code].
 * w == NULL means weight 1 for every sample.  Building block of A3 / A16 / K8. */
void orc_class_counts(const uint8_t* bed, int64_t N, int64_t M, const int32_t* w,
                      int64_t* cnt, int nthreads) {
    const int64_t B = bytes_per_snp(N);
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t j = 0; j < M; ++j) {
        const uint8_t* col = bed + j * B;
        int64_t c4[4] = {0, 0, 0, 0};
        for (int64_t i = 0; i < N; ++i) c4[code_at(col, i)] += w ? w[i] : 1;
        for (int t = 0; t < 4; ++t) cnt[j * 4 + t] = c4[t];
    }
}

/* ---------------------------------------------------------------------------------------
 * A5-A7: PlinkKinshipOperator::apply_block (kinship_operator.cpp:225-299).
 * z(r, t) = values[code(keep[r], variants[t])] with values = {(2-mu)s, 0, (1-mu)s, -mu s};
 * Y = (1/M) Z (Z^T X).  keep == NULL -> identity; variants == NULL -> all SNPs.
 * X, Y column-major n x k (Eigen default).  Y may alias X. */
void orc_kinship_apply(const uint8_t* bed, int64_t N, const double* means,
                       const double* inv_stds, const int32_t* keep, int64_t n,
                       const int32_t* variants, int64_t Msel, const double* X, int64_t k,
                       double* Y, int nthreads) {
    const int64_t B = bytes_per_snp(N);
    const double scaling = 1.0 / (double)Msel;
    double* U = (double*)calloc((size_t)(Msel * k > 0 ? Msel * k : 1), sizeof(double));
    /* U = Z^T X, one SNP per iteration, samples in order. */
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t t = 0; t < Msel; ++t) {
        const int64_t j = variants ? variants[t] : t;
        const double s = inv_stds[j];
        if (s == 0.0) continue;
        const double mu = means[j];
        const double vals[4] = {(2.0 - mu) * s, 0.0, (1.0 - mu) * s, (0.0 - mu) * s};
        const uint8_t* col = bed + j * B;
        for (int64_t c = 0; c < k; ++c) {
            const double* x = X + c * n;
            double acc = 0.0;
            for (int64_t r = 0; r < n; ++r) acc += vals[code_at(col, keep ? keep[r] : r)] * x[r];
            U[t * k + c] = acc;
        }
    }
    /* Y = Z U, each output row owned by one thread, SNPs summed in order. */
    double* R = (double*)calloc((size_t)(n * k > 0 ? n * k : 1), sizeof(double));
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t r = 0; r < n; ++r) {
        const int64_t smp = keep ? keep[r] : r;
        double acc[64];
        for (int64_t c0 = 0; c0 < k; c0 += 64) {
            const int64_t kc = (k - c0) < 64 ? (k - c0) : 64;
            for (int64_t c = 0; c < kc; ++c) acc[c] = 0.0;
            for (int64_t t = 0; t < Msel; ++t) {
                const int64_t j = variants ? variants[t] : t;
                const double s = inv_stds[j];
                if (s == 0.0) continue;
                const double mu = means[j];
                const int code = code_at(bed + j * B, smp);
                if (code == 1) continue;
                const double z = (kDosage[code] - mu) * s;
                const double* u = U + t * k + c0;
                for (int64_t c = 0; c < kc; ++c) acc[c] += z * u[c];
            }
            for (int64_t c = 0; c < kc; ++c) R[(c0 + c) * n + r] = acc[c] * scaling;
        }
    }
    memcpy(Y, R, (size_t)(n * k) * sizeof(double));
    free(R);
    free(U);
}

/* A7 as the reference computes it (kinship_operator.cpp:225-299), the CPU baseline's form: blocks
 * of 512 selected SNPs are decoded to FP64 z (n x 512, one column per SNP, inv_std == 0 -> zero
 * column), projected = z^T X (512 x k), result += z projected, and finally result x 1/M.  The two
 * products run over 256-row tiles with OpenMP (per-thread partial sums of z^T X, added in thread
 * order), so the same inputs give the same bits for a given thread count.  Same values as
 * orc_kinship_apply up to FP64 summation order. */
#define ORC_KB 512
#define ORC_RT 256
typedef double orc_v4d __attribute__((vector_size(32)));

__attribute__((target_clones("avx2", "default")))
static void orc_blk_zt_x(const double* Z, int64_t n, int cnt, const double* X, int64_t k, double* part,
                         int64_t r0, int64_t r1) {
    /* part[t * k + c] += sum_{r in [r0, r1)} Z[t n + r] X[c n + r]: the X tile is transposed to
     * [r][16 columns], and two SNP rows accumulate 16 columns each in registers, sequentially in r */
    orc_v4d xt[ORC_RT * 4];
    for (int64_t c0 = 0; c0 < k; c0 += 16) {
        const int kc = (int)((k - c0) < 16 ? (k - c0) : 16);
        for (int64_t a = r0; a < r1; a += ORC_RT) {
            const int64_t b = a + ORC_RT < r1 ? a + ORC_RT : r1;
            for (int64_t r = a; r < b; ++r)
                for (int c = 0; c < 16; ++c) ((double*)xt)[(r - a) * 16 + c] = c < kc ? X[(c0 + c) * n + r] : 0.0;
            for (int t = 0; t < cnt; t += 2) {
                const double* z0 = Z + (int64_t)t * n;
                const double* z1 = Z + (int64_t)(t + 1 < cnt ? t + 1 : t) * n;
                orc_v4d a0[4] = {{0}}, a1[4] = {{0}};
                for (int64_t r = a; r < b; ++r) {
                    const orc_v4d* xr = xt + (r - a) * 4;
                    const double s0 = z0[r], s1 = z1[r];
                    for (int q = 0; q < 4; ++q) {
                        a0[q] += s0 * xr[q];
                        a1[q] += s1 * xr[q];
                    }
                }
                for (int c = 0; c < kc; ++c) part[t * k + c0 + c] += a0[c / 4][c % 4];
                if (t + 1 < cnt)
                    for (int c = 0; c < kc; ++c) part[(t + 1) * k + c0 + c] += a1[c / 4][c % 4];
            }
        }
    }
}

__attribute__((target_clones("avx2", "default")))
static void orc_blk_z_p(const double* Z, int64_t n, int cnt, const double* P, int64_t k, double* R, int64_t r0,
                        int64_t r1) {
    /* R[c n + r] += sum_t Z[t n + r] P[t k + c] for r in [r0, r1): 8 rows x 8 columns held in
     * registers over all t (t in order) */
    const int64_t r8 = r0 + (r1 - r0) / 8 * 8;
    for (int tb = 0; tb < cnt; tb += 64)   /* 64 SNP columns at a time: few pages per row chunk */
    for (int64_t r = r0; r < r8; r += 8) {
        const int te = tb + 64 < cnt ? tb + 64 : cnt;
        for (int64_t c0 = 0; c0 < k; c0 += 8) {
            const int kc = (int)((k - c0) < 8 ? (k - c0) : 8);
            orc_v4d lo[8], hi[8];
            for (int c = 0; c < 8; ++c) {
                if (c < kc) {
                    memcpy(&lo[c], R + (c0 + c) * n + r, 32);
                    memcpy(&hi[c], R + (c0 + c) * n + r + 4, 32);
                } else {
                    lo[c] = hi[c] = (orc_v4d){0, 0, 0, 0};
                }
            }
            for (int t = tb; t < te; ++t) {
                const double* z = Z + (int64_t)t * n + r;
                orc_v4d z0, z1;
                memcpy(&z0, z, 32);
                memcpy(&z1, z + 4, 32);
                const double* pt = P + t * k + c0;
                for (int c = 0; c < 8; ++c) {
                    const double p = c < kc ? pt[c] : 0.0;
                    lo[c] += z0 * p;
                    hi[c] += z1 * p;
                }
            }
            for (int c = 0; c < kc; ++c) {
                memcpy(R + (c0 + c) * n + r, &lo[c], 32);
                memcpy(R + (c0 + c) * n + r + 4, &hi[c], 32);
            }
        }
    }
    for (int64_t r = r8; r < r1; ++r)
        for (int64_t c = 0; c < k; ++c) {
            double acc = R[c * n + r];
            for (int t = 0; t < cnt; ++t) acc += Z[(int64_t)t * n + r] * P[t * k + c];
            R[c * n + r] = acc;
        }
}

void orc_kinship_apply_blocked(const uint8_t* bed, int64_t N, const double* means, const double* inv_stds,
                               const int32_t* keep, int64_t n, const int32_t* variants, int64_t Msel,
                               const double* X, int64_t k, double* Y, int nthreads) {
    const int64_t B = bytes_per_snp(N);
    if (nthreads < 1) nthreads = 1;
    double* Z = (double*)malloc((size_t)(n * ORC_KB) * sizeof(double));
    double* P = (double*)malloc((size_t)(ORC_KB * (k > 0 ? k : 1)) * sizeof(double));
    double* parts = (double*)malloc((size_t)(nthreads * ORC_KB * (k > 0 ? k : 1)) * sizeof(double));
    double* R = (double*)calloc((size_t)(n * k > 0 ? n * k : 1), sizeof(double));
    for (int64_t first = 0; first < Msel; first += ORC_KB) {
        const int cnt = (int)((Msel - first) < ORC_KB ? (Msel - first) : ORC_KB);
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int j = 0; j < cnt; ++j) {
            const int64_t snp = variants ? variants[first + j] : first + j;
            const double mean = means[snp], s = inv_stds[snp];
            double* dest = Z + (int64_t)j * n;
            if (s == 0.0) {
                memset(dest, 0, (size_t)n * sizeof(double));
                continue;
            }
            const double vals[4] = {(2.0 - mean) * s, 0.0, (1.0 - mean) * s, -mean * s};
            const uint8_t* col = bed + snp * B;
            for (int64_t r = 0; r < n; ++r) dest[r] = vals[code_at(col, keep ? keep[r] : r)];
        }
        memset(parts, 0, (size_t)(nthreads * ORC_KB * k) * sizeof(double));
#pragma omp parallel num_threads(nthreads)
        {
            int th = 0, nth = 1;
#ifdef _OPENMP
            th = omp_get_thread_num();
            nth = omp_get_num_threads();
#endif
            const int64_t per = (n + nth - 1) / nth;
            const int64_t r0 = th * per, r1 = (r0 + per) < n ? (r0 + per) : n;
            if (r0 < r1) orc_blk_zt_x(Z, n, cnt, X, k, parts + (int64_t)th * ORC_KB * k, r0, r1);
#pragma omp barrier
#pragma omp single
            {
                for (int64_t e = 0; e < (int64_t)cnt * k; ++e) {
                    double v = 0.0;
                    for (int t2 = 0; t2 < nth; ++t2) v += parts[(int64_t)t2 * ORC_KB * k + e];
                    P[e] = v;
                }
            }
            if (r0 < r1) orc_blk_z_p(Z, n, cnt, P, k, R, r0, r1);
        }
    }
    const double scaling = 1.0 / (double)Msel;
    for (int64_t e = 0; e < n * k; ++e) Y[e] = R[e] * scaling;
    free(R);
    free(parts);
    free(P);
    free(Z);
}

/* A8: trace (kinship_operator.cpp:306-344) -- (1/M) sum_t sum_r z^2, missing skipped. */
double orc_kinship_trace(const uint8_t* bed, int64_t N, const double* means,
                         const double* inv_stds, const int32_t* keep, int64_t n,
                         const int32_t* variants, int64_t Msel) {
    const int64_t B = bytes_per_snp(N);
    double total = 0.0;
    for (int64_t t = 0; t < Msel; ++t) {
        const int64_t j = variants ? variants[t] : t;
        const double s = inv_stds[j];
        if (s == 0.0) continue;
        const double mu = means[j];
        const uint8_t* col = bed + j * B;
        double sum_sq = 0.0;
        for (int64_t r = 0; r < n; ++r) {
            const int code = code_at(col, keep ? keep[r] : r);
            if (code == 1) continue;
            const double z = (kDosage[code] - mu) * s;
            sum_sq += z * z;
        }
        total += sum_sq;
    }
    return total / (double)Msel;
}

/* A8: getDiagonal (kinship_operator.cpp:346-368) -- diag[r] += (1/M) z^2 in SNP order. */
void orc_kinship_diag(const uint8_t* bed, int64_t N, const double* means,
                      const double* inv_stds, const int32_t* keep, int64_t n,
                      const int32_t* variants, int64_t Msel, double* diag) {
    const int64_t B = bytes_per_snp(N);
    const double scaling = 1.0 / (double)Msel;
    for (int64_t r = 0; r < n; ++r) diag[r] = 0.0;
    for (int64_t t = 0; t < Msel; ++t) {
        const int64_t j = variants ? variants[t] : t;
        const double s = inv_stds[j];
        if (s == 0.0) continue;
        const double mu = means[j];
        const uint8_t* col = bed + j * B;
        for (int64_t r = 0; r < n; ++r) {
            const int code = code_at(col, keep ? keep[r] : r);
            if (code == 1) continue;
            const double z = (kDosage[code] - mu) * s;
            diag[r] += scaling * z * z;
        }
    }
}

/* ---------------------------------------------------------------------------------------
 * A10: GrmBuilder::compute single-population path (grm_builder.cpp:63-399), alg 0/3 (Yang,
 * z scaled by float 1/sqrt(2pq), G /= M) and alg 1 (VanRaden, centred, G /= sum float 2pq).
 * Per SNP: integer counts over ALL samples; freq = sum_x/(2 n_obs) (double); MAF filter
 * against float min_maf and 1e-8 (:218); mean = (float)(2 freq) (:240); 2pq = (float)(2 f(1-f))
 * (:234); scale = (float)(1/sqrt(pq2)) for alg 0 (:242-244).  sum_2pq accumulates the floats in
 * double per 2048-SNP block, then across blocks (:266-272, :347).
 * Per-SNP outputs (valid flag, float mean, float scale, float 2pq) are also returned because
 * the GPU path must reproduce exactly the same filtered SNP set.
 * G is N x N row-major FP64 (full square). */
/* calculateHWE (libCosmicCore/src/qc_utils.cpp:9-77): the exact Hardy-Weinberg test of
 * Wigginton, Cutler and Abecasis (2005) -- heterozygote distribution given the allele counts by
 * the recurrences from its mode, p = mass of the outcomes no more likely than the observed one. */
double orc_hwe_p(long long hom1, long long hets, long long hom2) {
    const long long homr = hom1 < hom2 ? hom1 : hom2;
    const long long rare = 2 * homr + hets, n = hom1 + hets + hom2;
    if (n == 0) return 1.0;
    long long mid = rare * (2 * n - rare) / (2 * n);
    if ((mid % 2) != (rare % 2)) mid++;
    double* pr = (double*)calloc((size_t)rare + 1, sizeof(double));
    pr[mid] = 1.0;
    double total = 1.0;
    long long r = (rare - mid) / 2, c = n - mid - r;
    for (long long h = mid; h >= 2; h -= 2) {
        pr[h - 2] = pr[h] * h * (h - 1.0) / (4.0 * (r + 1.0) * (c + 1.0));
        total += pr[h - 2];
        r++;
        c++;
    }
    r = (rare - mid) / 2;
    c = n - mid - r;
    for (long long h = mid; h <= rare - 2; h += 2) {
        pr[h + 2] = pr[h] * 4.0 * r * c / ((h + 2.0) * (h + 1.0));
        total += pr[h + 2];
        r--;
        c--;
    }
    const double target = pr[hets];
    double p = 0.0;
    for (long long h = 0; h <= rare; ++h) {
        if ((h % 2) != (rare % 2)) continue;
        if (pr[h] <= target) p += pr[h];
    }
    free(pr);
    p /= total;
    return p < 1.0 ? p : 1.0;
}

void orc_grm_snp_params(const uint8_t* bed, int64_t N, int64_t M, int alg, float min_maf, float min_hwe,
                        int8_t* valid, float* mean_f, float* scale_f, float* pq2_f,
                        double* sum_2pq_out, int64_t* n_valid_out, int nthreads) {
    const int64_t Bs = bytes_per_snp(N);
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t j = 0; j < M; ++j) {
        const uint8_t* col = bed + j * Bs;
        long long sum_x = 0, n_obs = 0, n_hom1 = 0, n_het = 0, n_hom2 = 0;
        for (int64_t i = 0; i < N; ++i) {
            int c = code_at(col, i);
            if (c == 1) continue;
            n_obs++;
            sum_x += kDosI[c];
            if (c == 0) n_hom1++;
            else if (c == 2) n_het++;
            else n_hom2++;
        }
        valid[j] = 0;
        mean_f[j] = 0.0f;
        scale_f[j] = 1.0f;
        pq2_f[j] = 0.0f;
        if (n_obs == 0) continue;
        double freq = (double)sum_x / (2.0 * n_obs);
        double maf = freq < 1.0 - freq ? freq : 1.0 - freq;
        if (maf < (double)min_maf || fabs(maf) < 1e-8) continue;
        if (min_hwe > 0.0f && orc_hwe_p(n_hom1, n_het, n_hom2) < (double)min_hwe) continue;
        pq2_f[j] = (float)(1.0f * 2.0 * freq * (1.0 - freq));
        mean_f[j] = (float)(freq * 2.0);
        if (alg == 0 || alg == 3) {
            double pq2 = 2.0 * freq * (1.0 - freq);
            scale_f[j] = (float)((pq2 > 1e-9) ? (1.0 / sqrt(pq2)) : 0.0);
        }
        valid[j] = 1;
    }
    double total = 0.0;
    int64_t nv = 0;
    for (int64_t b = 0; b < M; b += 2048) {
        double block = 0.0;
        int64_t e = b + 2048 < M ? b + 2048 : M;
        for (int64_t j = b; j < e; ++j)
            if (valid[j]) { block += pq2_f[j]; nv++; }
        total += block;
    }
    *sum_2pq_out = total;
    *n_valid_out = nv;
}

void orc_grm(const uint8_t* bed, int64_t N, int64_t M, int alg, float min_maf, float min_hwe, double* G,
             double* sum_2pq_out, int64_t* n_valid_out, int nthreads) {
    const int64_t Bs = bytes_per_snp(N);
    int8_t* valid = (int8_t*)malloc((size_t)M);
    float* mean_f = (float*)malloc((size_t)M * sizeof(float));
    float* scale_f = (float*)malloc((size_t)M * sizeof(float));
    float* pq2_f = (float*)malloc((size_t)M * sizeof(float));
    double sum_2pq;
    int64_t nv;
    orc_grm_snp_params(bed, N, M, alg, min_maf, min_hwe, valid, mean_f, scale_f, pq2_f, &sum_2pq, &nv,
                       nthreads);
    memset(G, 0, (size_t)(N * N) * sizeof(double));
    /* Z fill in FP64 from the float mean / scale (grm_builder.cpp:300-342), then G += Z Z^T. */
    const int64_t bs = 256;
    double* Z = (double*)malloc((size_t)(N * bs) * sizeof(double));
    for (int64_t b = 0; b < M; b += bs) {
        int64_t cnt = 0;
        for (int64_t j = b; j < M && j < b + bs; ++j) {
            if (!valid[j]) continue;
            const double mean = mean_f[j], scale = scale_f[j];
            const double vm[4] = {(2.0 - mean) * scale, 0.0, (1.0 - mean) * scale,
                                  (0.0 - mean) * scale};
            const uint8_t* col = bed + j * Bs;
            for (int64_t i = 0; i < N; ++i) Z[i * bs + cnt] = vm[code_at(col, i)];
            cnt++;
        }
        if (!cnt) continue;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 4)
        for (int64_t i = 0; i < N; ++i) {
            const double* zi = Z + i * bs;
            for (int64_t k = 0; k <= i; ++k) {
                const double* zk = Z + k * bs;
                double acc = 0.0;
                for (int64_t t = 0; t < cnt; ++t) acc += zi[t] * zk[t];
                G[i * N + k] += acc;
            }
        }
    }
    for (int64_t i = 0; i < N; ++i)
        for (int64_t k = 0; k <= i; ++k) {
            double g = G[i * N + k];
            if (alg == 0 || alg == 3) { if (nv > 0) g /= (double)nv; }
            else { if (sum_2pq > 0) g /= sum_2pq; }
            G[i * N + k] = g;
            G[k * N + i] = g;
        }
    *sum_2pq_out = sum_2pq;
    *n_valid_out = nv;
    free(Z); free(valid); free(mean_f); free(scale_f); free(pq2_f);
}

/* ---------------------------------------------------------------------------------------
 * Exact centred GRM (alg 1 semantics).  With the float mean mu_j = A_j * 2^-SB (A_j an integer
 * for a shift SB >= 29 chosen per matrix, orc_grm_exact_prepare), the reference's z_ij =
 * o_ij (x_ij - mu_j) is exactly U_ij * 2^-SB with the integer U_ij = o_ij (2^SB x_ij - A_j),
 * so G_raw = sum_j z z' = I * 2^-2SB with the exact 128-bit integer I = sum_j U_ij U_kj.
 * The functions below return, per entry, the correctly rounded FP64 of G_raw and the correctly
 * rounded FP32 of G_raw / sum_2pq (the GCTA .grm.bin value, grm_builder.cpp:1977-2026, without
 * the reference's GEMM rounding). */
static double i128_to_f64_rn(i128 v) {
    if (v == 0) return 0.0;
    int neg = v < 0;
    unsigned __int128 a = neg ? -(unsigned __int128)v : (unsigned __int128)v;
    int top = 127;
    while (!((a >> top) & 1)) top--;
    double r;
    if (top <= 52) {
        r = (double)(uint64_t)a;
    } else {
        int sh = top - 52;
        unsigned __int128 q = a >> sh;
        unsigned __int128 rem = a & ((((unsigned __int128)1) << sh) - 1);
        unsigned __int128 half = ((unsigned __int128)1) << (sh - 1);
        if (rem > half || (rem == half && (q & 1))) q += 1;
        r = ldexp((double)(uint64_t)q, sh);
    }
    return neg ? -r : r;
}

/* Correctly rounded float of (v * 2^-sc) / S for a positive normal double S. */
static float i128_div_f32_rn_sc(i128 v, double S, int sc);
/* Correctly rounded float of (v * 2^-58) / S for a positive normal double S. */
static float i128_div_f32_rn(i128 v, double S) { return i128_div_f32_rn_sc(v, S, 58); }
static float i128_div_f32_rn_sc(i128 v, double S, int sc) {
    if (v == 0) return 0.0f;
    int neg = v < 0;
    unsigned __int128 a = neg ? -(unsigned __int128)v : (unsigned __int128)v;
    int se;
    double fm = frexp(S, &se);                       /* S = fm * 2^se, fm in [0.5,1) */
    uint64_t D = (uint64_t)ldexp(fm, 53);            /* S = D * 2^(se-53) exactly */
    int top = 127;
    while (!((a >> top) & 1)) top--;
    /* quotient bits ~ top + 1 - 53 + sh; want >= 27 */
    int sh = 27 + 53 - top;
    if (sh < 0) sh = 0;
    unsigned __int128 num = a << sh;
    unsigned __int128 Q = num / D;
    int sticky = (num % D) != 0;
    /* value = Q * 2^(-sh - 58 - (se - 53)) (+ sticky) */
    int qtop = 127;
    while (!((Q >> qtop) & 1)) qtop--;
    int drop = qtop - 23;                             /* keep 24 significant bits */
    unsigned __int128 keepq = Q >> drop;
    unsigned __int128 rem = Q & ((((unsigned __int128)1) << drop) - 1);
    unsigned __int128 half = ((unsigned __int128)1) << (drop - 1);
    if (rem > half || (rem == half && (sticky || (keepq & 1)))) keepq += 1;
    else if (rem == half && !sticky && !(keepq & 1)) { /* tie to even: keep */ }
    int e2 = drop - sh - sc - (se - 53);
    float r = ldexpf((float)(uint64_t)keepq, e2);
    return neg ? -r : r;
}

/* Exact centred GRM entries (alg 1 semantics, as orc_grm_exact_centred) at any float mean: with
 * SB the smallest shift >= 29 that makes every valid mu_j 2^SB an integer, z_ij 2^SB =
 * U_ij = o_ij (x_ij 2^SB - mu_j 2^SB) exactly, and I = sum_j U_ij U_kj in int128.  Returns SB
 * (or -1 if SB would exceed 40 bits or M 2^22 SNPs, where int128 could overflow). */
static int orc_grm_exact_prepare(const uint8_t* bed, int64_t N, int64_t M, float min_maf, float min_hwe, int8_t* valid,
                                 int64_t* A, double* sum_2pq, int64_t* nv, int nthreads) {
    float* mean_f = (float*)malloc((size_t)M * sizeof(float));
    float* scale_f = (float*)malloc((size_t)M * sizeof(float));
    float* pq2_f = (float*)malloc((size_t)M * sizeof(float));
    orc_grm_snp_params(bed, N, M, 1, min_maf, min_hwe, valid, mean_f, scale_f, pq2_f, sum_2pq, nv, nthreads);
    int sb = 29;
    for (int64_t j = 0; j < M; ++j) {
        if (!valid[j]) continue;
        while (sb <= 40) {
            const double a = ldexp((double)mean_f[j], sb);
            if (a == floor(a)) break;
            sb++;
        }
    }
    if (sb > 40 || M > ((int64_t)1 << 22)) sb = -1;
    for (int64_t j = 0; j < M; ++j) A[j] = (sb > 0 && valid[j]) ? (int64_t)ldexp((double)mean_f[j], sb) : 0;
    free(mean_f); free(scale_f); free(pq2_f);
    return sb;
}

/* Tile of the exact centred GRM: rows x cols sample lists over all SNPs -> the correctly rounded
 * FP64 raw value and the correctly rounded FP32 of raw / sum_2pq per entry (row-major nr x nc).
 * Return: SB, or -1 (range, see orc_grm_exact_prepare). */
int orc_grm_exact_tile(const uint8_t* bed, int64_t N, int64_t M, float min_maf, float min_hwe, const int32_t* rows,
                       int64_t nr,
                       const int32_t* cols, int64_t nc, double* raw_f64, float* norm_f32, double* sum_2pq_out,
                       int64_t* n_valid_out, int nthreads) {
    const int64_t Bs = bytes_per_snp(N);
    int8_t* valid = (int8_t*)malloc((size_t)M);
    int64_t* A = (int64_t*)malloc((size_t)M * sizeof(int64_t));
    double sum_2pq;
    int64_t nv;
    const int sb = orc_grm_exact_prepare(bed, N, M, min_maf, min_hwe, valid, A, &sum_2pq, &nv, nthreads);
    if (sb < 0) { free(valid); free(A); return -1; }
    if (nthreads < 1) nthreads = 1;
    i128* acc = (i128*)calloc((size_t)(nthreads * nr * nc), sizeof(i128));
#pragma omp parallel num_threads(nthreads)
    {
        int th = 0, nth = 1;
#ifdef _OPENMP
        th = omp_get_thread_num();
        nth = omp_get_num_threads();
#endif
        i128* my = acc + (int64_t)th * nr * nc;
        int64_t* ur = (int64_t*)malloc((size_t)nr * sizeof(int64_t));
        int64_t* uc = (int64_t*)malloc((size_t)nc * sizeof(int64_t));
        for (int64_t j = th; j < M; j += nth) {
            if (!valid[j]) continue;
            const uint8_t* col = bed + j * Bs;
            for (int64_t a = 0; a < nr; ++a) {
                const int c = code_at(col, rows[a]);
                ur[a] = kObs[c] ? ((int64_t)kDosI[c] << sb) - A[j] : 0;
            }
            for (int64_t b = 0; b < nc; ++b) {
                const int c = code_at(col, cols[b]);
                uc[b] = kObs[c] ? ((int64_t)kDosI[c] << sb) - A[j] : 0;
            }
            for (int64_t a = 0; a < nr; ++a) {
                const i128 x = ur[a];
                if (x == 0) continue;
                i128* row = my + a * nc;
                for (int64_t b = 0; b < nc; ++b) row[b] += x * (i128)uc[b];
            }
        }
        free(ur);
        free(uc);
    }
    for (int64_t e = 0; e < nr * nc; ++e) {
        i128 v = 0;
        for (int t = 0; t < nthreads; ++t) v += acc[(int64_t)t * nr * nc + e];
        raw_f64[e] = ldexp(i128_to_f64_rn(v), -2 * sb);
        norm_f32[e] = sum_2pq > 0 ? i128_div_f32_rn_sc(v, sum_2pq, 2 * sb) : (float)raw_f64[e];
    }
    *sum_2pq_out = sum_2pq;
    *n_valid_out = nv;
    free(acc); free(A); free(valid);
    return sb;
}

/* Tile of the standardised GRM (alg 0 / 3, grm_builder.cpp:241-243, 300-304, 386-387) in FP64:
 * z = (x - fl32 mean) fl32 scale over the valid SNPs, G = Z Z^T / n_valid, rows x cols. */
void orc_grm_std_tile(const uint8_t* bed, int64_t N, int64_t M, float min_maf, float min_hwe, const int32_t* rows,
                      int64_t nr,
                      const int32_t* cols, int64_t nc, double* G, int64_t* n_valid_out, int nthreads) {
    const int64_t Bs = bytes_per_snp(N);
    int8_t* valid = (int8_t*)malloc((size_t)M);
    float* mean_f = (float*)malloc((size_t)M * sizeof(float));
    float* scale_f = (float*)malloc((size_t)M * sizeof(float));
    float* pq2_f = (float*)malloc((size_t)M * sizeof(float));
    double sum_2pq;
    int64_t nv;
    orc_grm_snp_params(bed, N, M, 0, min_maf, min_hwe, valid, mean_f, scale_f, pq2_f, &sum_2pq, &nv, nthreads);
    if (nthreads < 1) nthreads = 1;
    double* acc = (double*)calloc((size_t)(nthreads * nr * nc), sizeof(double));
#pragma omp parallel num_threads(nthreads)
    {
        int th = 0, nth = 1;
#ifdef _OPENMP
        th = omp_get_thread_num();
        nth = omp_get_num_threads();
#endif
        double* my = acc + (int64_t)th * nr * nc;
        double* zr = (double*)malloc((size_t)nr * sizeof(double));
        double* zc = (double*)malloc((size_t)nc * sizeof(double));
        for (int64_t j = th; j < M; j += nth) {
            if (!valid[j]) continue;
            const double mean = mean_f[j], scale = scale_f[j];
            const double vm[4] = {(2.0 - mean) * scale, 0.0, (1.0 - mean) * scale, (0.0 - mean) * scale};
            const uint8_t* col = bed + j * Bs;
            for (int64_t a = 0; a < nr; ++a) zr[a] = vm[code_at(col, rows[a])];
            for (int64_t b = 0; b < nc; ++b) zc[b] = vm[code_at(col, cols[b])];
            for (int64_t a = 0; a < nr; ++a)
                for (int64_t b = 0; b < nc; ++b) my[a * nc + b] += zr[a] * zc[b];
        }
        free(zr);
        free(zc);
    }
    for (int64_t e = 0; e < nr * nc; ++e) {
        double v = 0.0;
        for (int t = 0; t < nthreads; ++t) v += acc[(int64_t)t * nr * nc + e];
        G[e] = nv > 0 ? v / (double)nv : v;
    }
    *n_valid_out = nv;
    free(acc); free(valid); free(mean_f); free(scale_f); free(pq2_f);
}

/* Diagonal of the exact centred GRM for every sample (as orc_grm_exact_tile with rows = cols). */
int orc_grm_exact_diag(const uint8_t* bed, int64_t N, int64_t M, float min_maf, float min_hwe, double* raw_f64,
                       float* norm_f32,
                       int nthreads) {
    const int64_t Bs = bytes_per_snp(N);
    int8_t* valid = (int8_t*)malloc((size_t)M);
    int64_t* A = (int64_t*)malloc((size_t)M * sizeof(int64_t));
    double sum_2pq;
    int64_t nv;
    const int sb = orc_grm_exact_prepare(bed, N, M, min_maf, min_hwe, valid, A, &sum_2pq, &nv, nthreads);
    if (sb < 0) { free(valid); free(A); return -1; }
    if (nthreads < 1) nthreads = 1;
    i128* acc = (i128*)calloc((size_t)(nthreads * N), sizeof(i128));
#pragma omp parallel num_threads(nthreads)
    {
        int th = 0, nth = 1;
#ifdef _OPENMP
        th = omp_get_thread_num();
        nth = omp_get_num_threads();
#endif
        i128* my = acc + (int64_t)th * N;
        for (int64_t j = th; j < M; j += nth) {
            if (!valid[j]) continue;
            const uint8_t* col = bed + j * Bs;
            for (int64_t i = 0; i < N; ++i) {
                const int c = code_at(col, i);
                const int64_t u = kObs[c] ? ((int64_t)kDosI[c] << sb) - A[j] : 0;
                my[i] += (i128)u * u;
            }
        }
    }
    for (int64_t i = 0; i < N; ++i) {
        i128 v = 0;
        for (int t = 0; t < nthreads; ++t) v += acc[(int64_t)t * N + i];
        raw_f64[i] = ldexp(i128_to_f64_rn(v), -2 * sb);
        norm_f32[i] = sum_2pq > 0 ? i128_div_f32_rn_sc(v, sum_2pq, 2 * sb) : (float)raw_f64[i];
    }
    free(acc); free(A); free(valid);
    return sb;
}

/* The whole exact centred GRM (n x n, both triangles) at any float mean (orc_grm_exact_prepare):
 * per entry the correctly rounded FP64 of G_raw and the correctly rounded FP32 of G_raw / sum_2pq.
 * Return value: 0 ok, -1 out of range (orc_grm_exact_prepare). */
int orc_grm_exact_centred(const uint8_t* bed, int64_t N, int64_t M, float min_maf, float min_hwe,
                          double* raw_f64, float* norm_f32, double* sum_2pq_out,
                          int64_t* n_valid_out, int nthreads) {
    const int64_t Bs = bytes_per_snp(N);
    int8_t* valid = (int8_t*)malloc((size_t)M);
    int64_t* A = (int64_t*)malloc((size_t)M * sizeof(int64_t));
    double sum_2pq;
    int64_t nv;
    const int sb = orc_grm_exact_prepare(bed, N, M, min_maf, min_hwe, valid, A, &sum_2pq, &nv, nthreads);
    if (sb < 0) { free(valid); free(A); return -1; }
    /* U as int64 per (sample, valid SNP), SNP-contiguous per sample. */
    int64_t* idx = (int64_t*)malloc((size_t)(nv > 0 ? nv : 1) * sizeof(int64_t));
    int64_t t = 0;
    for (int64_t j = 0; j < M; ++j) if (valid[j]) idx[t++] = j;
    int64_t* U = (int64_t*)malloc((size_t)(N * (nv > 0 ? nv : 1)) * sizeof(int64_t));
    for (int64_t q = 0; q < nv; ++q) {
        const int64_t j = idx[q];
        const uint8_t* col = bed + j * Bs;
        for (int64_t i = 0; i < N; ++i) {
            int c = code_at(col, i);
            U[i * nv + q] = kObs[c] ? ((int64_t)kDosI[c] << sb) - A[j] : 0;
        }
    }
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 2)
    for (int64_t i = 0; i < N; ++i) {
        for (int64_t k = 0; k <= i; ++k) {
            i128 acc = 0;
            const int64_t* ui = U + i * nv;
            const int64_t* uk = U + k * nv;
            for (int64_t q = 0; q < nv; ++q) acc += (i128)ui[q] * (i128)uk[q];
            const double raw = ldexp(i128_to_f64_rn(acc), -2 * sb);
            const float nf = sum_2pq > 0 ? i128_div_f32_rn_sc(acc, sum_2pq, 2 * sb) : (float)raw;
            raw_f64[i * N + k] = raw; raw_f64[k * N + i] = raw;
            norm_f32[i * N + k] = nf; norm_f32[k * N + i] = nf;
        }
    }
    *sum_2pq_out = sum_2pq;
    *n_valid_out = nv;
    free(U); free(idx); free(A); free(valid);
    return 0;
}

/* ---------------------------------------------------------------------------------------
 * A16: genetics::prepareAssociationMarker, additive model (genotype_encoding.h:33-81), over a
 * keep list (lmm_engine.cpp:1500-1514): observed count, A1 freq = sum/(2 n_obs), MAF pass if
 * n_obs >= 3 and min(f, 1-f) >= min_maf; mean-impute missing then centre.
 * Per SNP outputs: freq, n_obs, pass (0/1); g (n) only when g_out != NULL. */
void orc_marker(const uint8_t* bed, int64_t N, int64_t j, const int32_t* keep, int64_t n,
                double min_maf, double* freq, int64_t* n_obs, int* pass, double* g_out) {
    const uint8_t* col = bed + j * bytes_per_snp(N);
    double dosage_sum = 0.0;
    int64_t obs = 0;
    for (int64_t r = 0; r < n; ++r) {
        int c = code_at(col, keep ? keep[r] : r);
        if (c == 1) continue;
        dosage_sum += kDosage[c];
        obs++;
    }
    *n_obs = obs;
    *pass = 0;
    *freq = 0.0;
    if (obs == 0) return;
    double f = dosage_sum / (2.0 * obs);
    *freq = f;
    double maf = f < 1.0 - f ? f : 1.0 - f;
    *pass = (obs >= 3 && maf >= min_maf) ? 1 : 0;
    if (g_out) {
        const double mean = dosage_sum / obs;
        for (int64_t r = 0; r < n; ++r) {
            int c = code_at(col, keep ? keep[r] : r);
            double v = (c == 1) ? mean : kDosage[c];
            g_out[r] = v - mean;
        }
    }
}

/* Projected per-SNP sums used by every scan form: for SNP j and column c of R (n x nr,
 * column-major, over the keep list), dot[j*nr + c] = sum_r g_rj R_rc with g the prepared
 * marker; gtg[j] = g^T g.  Pass / freq / n_obs as orc_marker.  Loops over SNPs in parallel. */
void orc_scan_sums(const uint8_t* bed, int64_t N, int64_t M, const int32_t* keep, int64_t n,
                   double min_maf, const double* R, int64_t nr, double* freq, int64_t* n_obs,
                   int32_t* pass, double* dot, double* gtg, int nthreads) {
#pragma omp parallel num_threads(nthreads)
    {
        double* g = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
#pragma omp for schedule(static)
        for (int64_t j = 0; j < M; ++j) {
            int ps;
            orc_marker(bed, N, j, keep, n, min_maf, &freq[j], &n_obs[j], &ps, g);
            pass[j] = ps;
            if (n_obs[j] == 0) {
                for (int64_t c = 0; c < nr; ++c) dot[j * nr + c] = 0.0;
                gtg[j] = 0.0;
                continue;
            }
            double s2 = 0.0;
            for (int64_t r = 0; r < n; ++r) s2 += g[r] * g[r];
            gtg[j] = s2;
            for (int64_t c = 0; c < nr; ++c) {
                const double* x = R + c * n;
                double acc = 0.0;
                for (int64_t r = 0; r < n; ++r) acc += g[r] * x[r];
                dot[j * nr + c] = acc;
            }
        }
        free(g);
    }
}

/* ---------------------------------------------------------------------------------------
 * A17: gamma_q / chisq_prob (lmm_engine.cpp:22-64): series for x < a+1, Lentz continued
 * fraction otherwise, relative tolerance 1e-7, at most 100 iterations; stat <= 0 -> 1. */
double orc_gamma_q(double a, double x) {
    if (x < 0.0 || a <= 0.0) return 1.0;
    if (x < a + 1.0) {
        double ap = a, sum = 1.0 / a, del = sum;
        for (int i = 0; i < 100; ++i) {
            ap += 1.0;
            del *= x / ap;
            sum += del;
            if (fabs(del) < fabs(sum) * 1e-7)
                return 1.0 - sum * exp(-x + a * log(x) - lgamma(a));
        }
    } else {
        double b = x + 1.0 - a, c = 1.0 / 1e-30, d = 1.0 / b, h = d;
        for (int i = 1; i <= 100; ++i) {
            double an = -i * (i - a);
            b += 2.0;
            d = an * d + b;
            if (fabs(d) < 1e-30) d = 1e-30;
            c = b + an / c;
            if (fabs(c) < 1e-30) c = 1e-30;
            d = 1.0 / d;
            double del = d * c;
            h *= del;
            if (fabs(del - 1.0) < 1e-7) return exp(-x + a * log(x) - lgamma(a)) * h;
        }
    }
    return 0.0;
}

double orc_chisq_prob(double stat, int df) {
    if (stat <= 0) return 1.0;
    return orc_gamma_q(df / 2.0, stat / 2.0);
}

void orc_chisq_prob_vec(const double* stat, int64_t n, int df, double* p) {
    for (int64_t i = 0; i < n; ++i) p[i] = orc_chisq_prob(stat[i], df);
}

/* ---------------------------------------------------------------------------------------
 * Synthetic HWE genotype generator (not a reference function).  Counter-
 * based so the GPU generator writes identical bytes: SNP j draws p_j ~ U(pmin, pmax) from
 * hash(seed, j, ~0); sample i of SNP j draws u from hash(seed, j, i); missing if u < miss,
 * else Binomial(2, p_j) by inverse CDF on a second draw.  Output: raw BED body (no magic),
 * SNP-major, ceil(N/4) bytes per SNP, pad codes 00 (PLINK convention: unused bits zero). */
static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
static inline uint64_t hash3(uint64_t seed, uint64_t j, uint64_t i) {
    return mix64(mix64(mix64(seed + 0x9e3779b97f4a7c15ULL) ^ (j * 0xd1b54a32d192ed03ULL)) ^
                 (i * 0x8cb92ba72f3d8dd7ULL + 0x632be59bd9b4e019ULL));
}
static inline double u01(uint64_t h) { return (double)(h >> 11) * 0x1.0p-53; }

int orc_gen_code(uint64_t seed, int64_t j, int64_t i, double pmin, double pmax, double miss) {
    double p = pmin + (pmax - pmin) * u01(hash3(seed, (uint64_t)j, ~0ULL));
    uint64_t h = hash3(seed, (uint64_t)j, (uint64_t)i);
    double u = u01(h);
    if (u < miss) return 1;
    double v = u01(mix64(h ^ 0xa0761d6478bd642fULL));
    double q = 1.0 - p;
    double p0 = q * q;            /* 0 copies of A1 */
    double p1 = p0 + 2.0 * p * q; /* <= 1 copy */
    int x = v < p0 ? 0 : (v < p1 ? 1 : 2);
    return x == 2 ? 0 : (x == 1 ? 2 : 3);
}

void orc_gen_hwe_bed(uint64_t seed, int64_t N, int64_t M, double pmin, double pmax,
                     double miss, uint8_t* bed, int nthreads) {
    const int64_t B = bytes_per_snp(N);
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t j = 0; j < M; ++j) {
        uint8_t* col = bed + j * B;
        for (int64_t b = 0; b < B; ++b) {
            uint8_t byte = 0;
            for (int s = 0; s < 4; ++s) {
                int64_t i = b * 4 + s;
                if (i >= N) break;
                byte |= (uint8_t)(orc_gen_code(seed, j, i, pmin, pmax, miss) << (2 * s));
            }
            col[b] = byte;
        }
    }
}
