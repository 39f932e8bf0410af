/*
 * cosmicscanner_amd.h -- C-ABI of the MI355X genotype hot path (PLINK BED decode, matrix-free
 * kinship operator K.V, GRM build, per-SNP LMM scan).
 *
 * Plain C: pointers, sizes and status codes; no C++ or torch types.  Every function returns
 * CS_OK (0) or a negative status; cs_last_error() gives the thread-local message.  Status
 * CS_ERR_INVALID_ARGUMENT corresponds to the reference's std::invalid_argument, CS_ERR_RUNTIME to
 * its std::runtime_error (the C++ adapter in cosmicscanner_amd_kinship.hpp rethrows them).
 *
 * Host arrays are borrowed for the duration of a call.  Matrices are column-major with a
 * leading dimension (Eigen's default layout).  All device state lives in the handles.
 */
#ifndef COSMICSCANNER_AMD_H
#define COSMICSCANNER_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CS_OK 0
#define CS_ERR_INVALID_ARGUMENT (-1)
#define CS_ERR_RUNTIME (-2)

/* Precision modes of the K.V operator and the projected scan.
 * INT8_L3: int8 MFMA, exact genotype operand, right-hand sides as 3 signed 8-bit limbs
 *          (23-bit fixed point per column), int32 accumulation -- the default.
 * INT8_L2: as above with 2 limbs (15-bit fixed point per column).
 * INT8_L4: as above with 4 limbs (31-bit fixed point per column); the PCG's continuation
 *          below rel_tol 1e-5 runs on it.
 * FP32:    plain FP32 VALU kernels (decode to FP32, FMA, no MFMA, no split) -- the comparator. */
#define CS_PREC_INT8_L3 3
#define CS_PREC_INT8_L2 2
#define CS_PREC_INT8_L4 4
#define CS_PREC_FP32 32

typedef struct cs_bed cs_bed;
typedef struct cs_kinship cs_kinship;
typedef struct cs_lmm cs_lmm;
typedef struct cs_mvlmm cs_mvlmm;
typedef struct cs_lowrank cs_lowrank;
typedef struct cs_server cs_server;

/* Thread-local message of the last failed call on this thread ("" if none). */
const char* cs_last_error(void);
/* Library version as major*10000 + minor*100 + patch. */
int cs_version(void);

/* ---- Device buffers (for device-resident right-hand sides and solver vectors; no reference
 * counterpart: the reference keeps everything in host memory). */
int cs_dev_alloc(int64_t bytes, void** dptr);
int cs_dev_free(void* dptr);
int cs_dev_upload(void* dst, const void* src, int64_t bytes);
int cs_dev_download(void* dst, const void* src, int64_t bytes);
int cs_dev_synchronize(void);
/* Plain device-to-device copy kernel (also the calibration reference for HBM byte counters). */
int cs_dev_copy(void* dst, const void* src, int64_t bytes);
/* Frees the library's cache of released device blocks (blocks of 8 MB and more are kept for
 * reuse, up to half of the device's memory per process or CSAMD_DEVICE_CACHE_BYTES; destroying a
 * genotype matrix or a server also frees it).  No reference counterpart. */
int cs_dev_release_cache(void);

/* ---- Genotypes (replaces PlinkReader::load/openBed, libCosmicCore/src/plink_reader.cpp:28-92,
 * with the packed matrix resident in HBM).  `body` is the .bed content after the 3 magic bytes:
 * n_snps rows of ceil(n_samples/4) bytes, SNP-major.  Padding slots of each row's last byte are
 * ignored. */
int cs_bed_create(const uint8_t* body, int64_t n_samples, int64_t n_snps, cs_bed** out);
/* Same, from a whole .bed file image including the magic (checks 6C 1B 01 as
 * plink_reader.cpp:74-80 does). */
int cs_bed_create_from_file_image(const uint8_t* image, int64_t image_bytes, int64_t n_samples,
                                  int64_t n_snps, cs_bed** out);
/* Synthetic HWE genotypes generated on the device (no host transfer): p_j ~ U(pmin, pmax),
 * missing with probability `miss`, else Binomial(2, p_j). */
int cs_bed_create_synthetic(uint64_t seed, int64_t n_samples, int64_t n_snps, double pmin,
                            double pmax, double miss, cs_bed** out);
int cs_bed_destroy(cs_bed* bed);
int cs_bed_dims(const cs_bed* bed, int64_t* n_samples, int64_t* n_snps);
/* Copy the packed body back (n_snps * ceil(n_samples/4) bytes; padding slots zeroed). */
int cs_bed_download(cs_bed* bed, uint8_t* body);
/* The same for the listed SNPs only: row i of body (ceil(n_samples/4) bytes) is SNP snps[i]. */
int cs_bed_download_snps(cs_bed* bed, const int32_t* snps, int64_t n_snps, uint8_t* body);
/* Build the resident sample-major copy (used by the K.V forward product and GRM); idempotent. */
int cs_bed_build_sample_major(cs_bed* bed);
/* Bytes of HBM held by the handle. */
int cs_bed_device_bytes(const cs_bed* bed, int64_t* bytes);

/* ---- Kernel (a): decode + statistics.
 * PlinkReader::computeStats (plink_reader.cpp:187-239): all-sample mean and
 * 1/sqrt(2p(1-p)) per SNP (0 if var < 1e-9 or no observed genotype).  Idempotent. */
int cs_bed_compute_stats(cs_bed* bed);
int cs_bed_get_stats(cs_bed* bed, double* means, double* inv_stds);
/* Weighted class counts per SNP: counts[j*4 + code] = sum_s w_s [code(s,j) == code], codes
 * 0 (2 copies of A1), 1 (missing), 2 (1 copy), 3 (0 copies).  weights may be NULL (all 1). */
int cs_bed_class_counts(cs_bed* bed, const int32_t* weights, int64_t* counts);

/* ---- Kernel (b): matrix-free kinship operator K = Z_sel Z_sel^T / M_sel
 * (PlinkKinshipOperator, src/kinship_operator.h:19-93; constructor kinship_operator.cpp:28-68).
 * keep: n sample indices (any order, repeats allowed) or NULL/0 for all samples; variants: SNP
 * indices or NULL/0 for all SNPs.  Copies both lists; the bed handle must outlive the
 * operator.  Errors as the reference: index out of range / no variants -> CS_ERR_RUNTIME. */
int cs_kinship_create(cs_bed* bed, const int32_t* keep, int64_t n_keep, const int32_t* variants,
                      int64_t n_variants, cs_kinship** out);
int cs_kinship_destroy(cs_kinship* op);
/* rows() == cols() (kinship_operator.h:27-28) and numVariants() (:31). */
int cs_kinship_rows(const cs_kinship* op, int64_t* rows);
int cs_kinship_num_variants(const cs_kinship* op, int64_t* m);
int cs_kinship_set_precision(cs_kinship* op, int mode);
/* How the int8 operator treats O = observed-genotype indicators (same results either way, both
 * exact integer sums): 1 dense (an O product on the MFMA units next to each genotype product),
 * 2 sparse (genotype products only; O^T V and O (mu W) from column sums minus the missing
 * entries, from lists built once per genotype handle), 0 auto: sparse when at most 0.15 % of
 * the genotypes are missing and k L > 16 (L = limbs), else dense.  The list gathers read one
 * 64-byte row per missing genotype and 16 columns; at n = 50k, m = 500k, k = 16 they break even
 * with the O products near 0.2 % missing. */
int cs_kinship_set_missing_mode(cs_kinship* op, int mode);
/* No counterpart in the reference (an HBM-for-work trade of this library): run the operator's
 * genotype products on a compact copy of its cohort's rows, gathered once from the resident
 * matrix (2 * roundup(n, 256) * roundup(M, 256) / 4 bytes more), so that each apply reads and
 * multiplies n instead of N sample rows.  The copy keeps the all-sample standardisation, and
 * results are bitwise those of the full-sample products.  No-op for the all-sample cohort;
 * CS_ERR_INVALID_ARGUMENT when keep repeats a sample. */
int cs_kinship_compact_rows(cs_kinship* op);
/* Device bytes the operator holds: its lists, grown workspaces and compact copy. */
int cs_kinship_device_bytes(const cs_kinship* op, int64_t* bytes);
/* apply / apply_block / applyTranspose (kinship_operator.cpp:211-304): Y = K X for X n x k
 * column-major with leading dimension ldx; Y may alias X; k == 0 is allowed.  ldx < n or
 * n_threads-like misuse -> CS_ERR_RUNTIME "dimension mismatch". */
int cs_kinship_apply(cs_kinship* op, const double* X, int64_t ldx, int64_t k, double* Y,
                     int64_t ldy);
/* Same with X, Y in device memory (no host transfer); for device-resident solvers. */
int cs_kinship_apply_device(cs_kinship* op, const double* dX, int64_t ldx, int64_t k, double* dY,
                            int64_t ldy);
/* PlinkReader::multiply_Zt_v and multiply_Z_v (libCosmicCore/src/plink_reader.cpp:241-438, the
 * keep-list overloads when the operator has a keep list): with z_ij = (x_ij - mean_j) inv_std_j
 * over ALL SNPs of the genotypes (the operator's variant selection does not apply; missing and
 * inv_std == 0 give 0), U = Z^T V (M x k from n x k) and Y = Z U (n x k from M x k), column-major
 * with leading dimensions.  Same int8 products as the apply (exact genotypes, 23-bit operands at
 * INT8_L3).  Host-buffer and device-buffer forms; wrong leading dimensions -> CS_ERR_RUNTIME
 * "Dimension mismatch in multiply_Zt_v" / "... multiply_Z_v". */
int cs_kinship_multiply_zt(cs_kinship* op, const double* V, int64_t ldv, int64_t k, double* U, int64_t ldu);
int cs_kinship_multiply_z(cs_kinship* op, const double* U, int64_t ldu, int64_t k, double* Y, int64_t ldy);
int cs_kinship_multiply_zt_device(cs_kinship* op, const double* dV, int64_t ldv, int64_t k, double* dU,
                                  int64_t ldu);
int cs_kinship_multiply_z_device(cs_kinship* op, const double* dU, int64_t ldu, int64_t k, double* dY,
                                 int64_t ldy);
/* No counterpart in the reference.  cs_kinship_apply_device for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): it allocates nothing and keeps every intermediate (scattered and quantised
 * operands, the Z^T V epilogue output, the int64 result rows) in one fixed 256 MiB device
 * workspace per GPU, held from library load until the process ends.  Runs the dense
 * missing-genotype form, whose results are bitwise identical to the sparse one (so to
 * cs_kinship_apply_device with any mode).  The workspace needs 16 Npad k + 16 Npad G L +
 * 16 Mpad k + 32 Mpad G L bytes (Npad, Mpad: samples and SNPs rounded up to 256; G = ceil(k / 16)
 * column groups, L limbs; packed tiles when k L <= 16), e.g. 190 MB at 50,000 x 500,000 with
 * k = 16; larger requests give CS_ERR_INVALID_ARGUMENT before any device work, as does the FP32
 * comparator precision.  Every call on a GPU shares that GPU's workspace: calls that overlap on
 * one GPU (from several threads or operators) race and must be serialised by the caller. */
int cs_kinship_apply_device_noalloc(cs_kinship* op, const double* dX, int64_t ldx, int64_t k,
                                    double* dY, int64_t ldy);
/* trace() and getDiagonal() (kinship_operator.cpp:306-368). */
int cs_kinship_trace(cs_kinship* op, double* trace);
int cs_kinship_diag(cs_kinship* op, double* diag);
/* Time of the last apply's kernels in ms per phase (0: quantize+scatter, 1: Z^T product,
 * 2: middle epilogue, 3: Z product, 4: final epilogue), from HIP events on the operator's stream. */
int cs_kinship_last_timing(const cs_kinship* op, double* ms5);

/* ---- Kernel (c): per-SNP association scan.
 * Marker preparation as genetics::prepareAssociationMarker (src/genetics/genotype_encoding.h:33-81)
 * on the keep list (lmm_engine.cpp:1500-1514): n_obs, A1 frequency, MAF pass (n_obs >= 3 and
 * min(f,1-f) >= min_maf), mean-impute and centre -> g_j.  R is n x nr column-major (over the keep
 * list).  snps: the SNPs to scan in output order (NULL/0 = all).
 *
 * cs_scan_sums: dot[i*nr + c] = g_{snps[i]}^T R_c and gtg[i] = g^T g (any output may be NULL). */
int cs_scan_sums(cs_bed* bed, const int32_t* keep, int64_t n_keep, const double* R, int64_t ldr,
                 int64_t nr, double min_maf, const int32_t* snps, int64_t n_snps, int prec,
                 double* freq, int64_t* n_obs, int32_t* pass, double* dot, double* gtg);

/* Variance model of the fused scan: with d_c = g^T R_c,
 *   score_t = d_t                                  (t < n_score: one column per trait)
 *   var     = alpha g^T g + sum_{c in [quad_begin, quad_end)} quad_w[c - quad_begin] d_c^2
 *             - q^T cov_C q,   q = d[cov_begin:cov_end]  (at most 32 covariate columns)
 * This covers the rotated dense form (LmmEngine::scan, lmm_engine.cpp:1430-1449: R = [P y,
 * W_proj diag(V^-1/2)], alpha 0, unit weights), the low-rank Woodbury form (scanLowRank,
 * lmm_engine.cpp:1297-1307: alpha = base, weights d_l - base on U, covariate block V^-1 X with
 * cov_C = (X^T V^-1 X)^-1) and the PCG-projected form (alpha = calibration ratio). */
typedef struct cs_scan_model {
    int64_t n_score;
    double alpha;
    int64_t quad_begin, quad_end;
    const double* quad_w;
    int64_t cov_begin, cov_end;
    const double* cov_C;
} cs_scan_model;

/* Fused scan: per SNP and trait beta = score/var, se = var^-1/2, p = chi^2_1 tail of
 * score^2/var (gamma_q of lmm_engine.cpp:22-64), log10p = -log10 p (300 when p == 0).  SNPs
 * failing the MAF filter or with var <= 1e-12 get NaN (reference: not written). Outputs
 * n_snps x n_score row-major; freq / n_obs / pass per SNP. */
int cs_lmm_scan(cs_bed* bed, const int32_t* keep, int64_t n_keep, const double* R, int64_t ldr,
                int64_t nr, double min_maf, const int32_t* snps, int64_t n_snps, int prec,
                const cs_scan_model* model, double* freq, int64_t* n_obs, int32_t* pass,
                double* beta, double* se, double* p, double* log10p);
/* Same with R and every output in device memory (d_R n x nr with leading dimension ldr; outputs
 * as cs_lmm_scan, each may be NULL): the statistics stay in HBM, written by the epilogue kernel
 * directly, with no host transfer -- for device-resident pipelines and serving. */
int cs_lmm_scan_device(cs_bed* bed, const int32_t* keep, int64_t n_keep, const double* d_R, int64_t ldr,
                       int64_t nr, double min_maf, const int32_t* snps, int64_t n_snps, int prec,
                       const cs_scan_model* model, double* d_freq, int64_t* d_n_obs, int32_t* d_pass,
                       double* d_beta, double* d_se, double* d_p, double* d_log10p);
/* Device times of the last scan pass in this process (HIP events on its stream, ms): the weighted
 * class counts, the genotype products g^T R (summed over SNP chunks) and the fused epilogues;
 * packed_bytes: the SNP-major genotype bytes the products read.  Any may be NULL. */
int cs_scan_last_timing(double* ms_counts, double* ms_products, double* ms_epilogue,
                        int64_t* packed_bytes);

/* ---- Kernel (b): dense GRM build.
 * GrmBuilder::compute(const PlinkReader&) (libCosmicCore/src/grm_builder.cpp:63-399) for
 * algorithm 1 (VanRaden, the default, --make-grm-alg 1): per-SNP filter on all samples
 * (min(f,1-f) >= min_maf and > 1e-8, :167-257; with min_hwe > 0 also the exact HWE p-value of
 * the genotype counts >= min_hwe, :220-223, calculateHWE of qc_utils.cpp:9-77), float32 mean 2p
 * and float32 2pq, centred
 * z = x - mean (missing -> 0), G = Z Z^T / sum_2pq with sum_2pq the FP64 sum of the float32 2pq in
 * 2048-SNP block order (:266-272, :392).  Computed exactly in integers on int8 MFMA:
 *   grm_tri_f32: n(n+1)/2 floats, lower triangle row-major (i, j <= i) -- the GCTA .grm.bin
 *                order written by GrmBuilder::save (:1977-2026); each value is the correctly
 *                rounded float of the exact G_ij.  May be NULL.
 *   grm_raw_f64: n x n doubles, the correctly rounded FP64 of Z Z^T (before normalisation,
 *                symmetric).  May be NULL.
 *   sum_2pq / n_valid: the normaliser and the number of SNPs used (.grm.N.bin holds n_valid).
 * Any float32 mean is exact (the integer scale grows to at most 2^40 for rare SNPs); a mean needing
 * more fractional bits, or min_hwe outside [0, 1], -> CS_ERR_INVALID_ARGUMENT.
 * alg 0 or 3 (Yang / GCTA default, grm_builder.cpp:241-243, 386-387): z = (x - fl32(2p)) *
 * fl32(1/sqrt(2p(1-p))), G = Z Z^T / n_valid, computed as FP16 hi/lo split products on MFMA with
 * FP32 accumulation flushed to FP64 every 16384 SNPs (entries within ~1e-6 of the FP64 value
 * relative to the diagonal); grm_raw_f64 is then that FP64 Z Z^T and grm_tri_f32 fl32(raw / n_valid).
 * Other algorithms -> CS_ERR_INVALID_ARGUMENT. */
int cs_grm_build(cs_bed* bed, int alg, float min_maf, float min_hwe, float* grm_tri_f32,
                 double* grm_raw_f64, double* sum_2pq, int64_t* n_valid);
/* Same with the lower triangle written to device memory (n(n+1)/2 floats). */
int cs_grm_build_device(cs_bed* bed, int alg, float min_maf, float min_hwe, float* d_grm_tri_f32,
                        double* sum_2pq, int64_t* n_valid);
/* Raw Z Z^T (the correctly rounded FP64 of the exact integer product, n x n, symmetric) over a
 * SNP subset -- a chromosome's per_chrom_raw in GrmBuilder::computeSplit
 * (grm_builder.cpp:401-785, accumulated at :703-713) -- with the same per-SNP filter and float32
 * means as cs_grm_build.  snps NULL/0 = all SNPs.  raw_f64 is host memory, or device memory when
 * raw_on_device != 0.  sum_2pq_block: the float 2pq of the selected valid SNPs summed in 2048-SNP
 * block order (the global normaliser, :266-272); sum_2pq_seq: the same summed SNP by SNP (the
 * per-chromosome per_chrom_sum_2pq, :718-724). */
int cs_grm_build_raw(cs_bed* bed, int alg, float min_maf, float min_hwe, const int32_t* snps,
                     int64_t n_snps, double* raw_f64, int raw_on_device, double* sum_2pq_block,
                     double* sum_2pq_seq, int64_t* n_valid);
/* Kernel time of the last GRM build on this thread (all int8 products; the int128 combine), the
 * number of int8 product launches and the int8 MFMA ops they issued, from HIP events on the
 * build's stream. */
int cs_grm_last_timing(double* ms_products, double* ms_combine, int* launches, double* mfma_ops);

/* ---- Matrix-free solves.
 * Multi-RHS PCG for (sigma_g K + sigma_e I) X = B with the kinship operator K
 * (MatrixFreePCG::solve, libCosmicCore/src/matrix_free_pcg.h:20-66): the columns share one K.V
 * apply per iteration and converge independently on ||r|| / ||b|| <= rel_tol (max_iter
 * iterations at most).  precond 0: scalar 1 / (sigma_g + sigma_e) (lmm_engine.cpp:306-313);
 * 1: Jacobi with diag(K).  dB, dX: device n x k column-major; warm_start != 0 starts from the
 * given dX.  iterations: iterations run; rel_residual (k, may be NULL): the true residual
 * ||B - A X|| / ||B|| per column, from one more apply.  With an INT8_L3 operator and
 * rel_tol < 1e-5 the solve continues, once the INT8_L3 iteration has met rel_tol, as a
 * warm-started PCG on INT8_L4 operands (the 23-bit operands leave the true residual near
 * 1e-6-1e-5); iterations counts both phases, the true residual is the INT8_L4 one and the
 * operator is back at INT8_L3 on return. */
int cs_pcg_solve(cs_kinship* K, double sigma_g, double sigma_e, const double* dB, int64_t ldb,
                 int64_t k, double* dX, int64_t ldx, double rel_tol, int max_iter, int precond,
                 int warm_start, int* iterations, double* rel_residual);
/* Same with per-column variance components (sigma_g[c], sigma_e[c] for column c). */
int cs_pcg_solve_cols(cs_kinship* K, const double* sigma_g, const double* sigma_e, const double* dB,
                      int64_t ldb, int64_t k, double* dX, int64_t ldx, double rel_tol, int max_iter,
                      int precond, int warm_start, int* iterations, double* rel_residual);

/* Prepared markers (genotype_encoding.h:61-79 on the keep list: mean-imputed, centred; zero
 * columns for SNPs failing n_obs >= 3 / MAF) of `snps` as dense device columns d_G (n x n_snps,
 * leading dimension ldg); pass / freq per SNP (host, may be NULL). */
int cs_markers_device(cs_bed* bed, const int32_t* keep, int64_t n_keep, const int32_t* snps,
                      int64_t n_snps, double min_maf, double* d_G, int64_t ldg, int32_t* pass,
                      double* freq);

/* PCG-projected scan with one (e.g. LOCO) kinship operator K over the keep list:
 * V = sigma_g K + sigma_e I, one multi-RHS PCG solve V [Z_Y, Z_X, Z_C] = [Y, X, G_cal],
 * P y = Z_Y - Z_X (X^T Z_X)^-1 X^T Z_Y (the GLS projection of LmmEngine::scan,
 * src/lmm_engine.cpp:1363-1525), calibration gamma = mean_cal g^T P g / g~^T g~ (g~ = g projected
 * off X), and per SNP and trait: score = g^T P y_t, var = gamma g~^T g~, beta, se, p, log10p
 * (n_snps x T row-major; NaN when not tested).  Exact stage: every tested SNP whose smallest
 * P over the T traits, Bonferroni-adjusted (x T), reaches 10^-exact_log10p -- i.e. whose largest
 * -log10 P reaches exact_log10p + log10 T -- is recomputed with its exact GLS variance g^T P g (a
 * second multi-RHS PCG solve over those markers, tolerance min(100 rel_tol, 1e-4), and the
 * variational form g^T z + z^T r of g^T V^-1 g, accurate to the square of the residual) and score
 * g^T P y_t in FP64.  exact_log10p: 0 -> 5 (the default: suggestive significance), < 0 -> every
 * tested SNP, +inf -> off.  Below the threshold the calibrated statistics stand;
 * dlog10p_bound_below states their |d(-log10 P)| bound from the spread of the calibration
 * markers' ratios (4 standard deviations of the variance ratio at the threshold statistic; 0 when
 * every SNP is exact, NaN when the stage is off or fewer than two markers calibrate).  Y: n x T FP64 (ldy); X: n x c float32
 * (ldx, the intercept included, 1 <= c <= 32).  cal_snps: calibration SNPs (those failing the
 * filter are skipped). */
typedef struct cs_pscan_info {
    double gamma;
    int64_t n_calibration;
    int pcg_iterations;
    double max_rel_residual;
    int64_t n_exact;              /* SNPs recomputed by the exact stage */
    int exact_pcg_iterations;     /* largest iteration count of its solves */
    double calibration_rel_sd;    /* relative standard deviation of the calibration ratios */
    double exact_log10p;          /* threshold used on the largest -log10 P over the traits
                                     (level + log10 T; -1: every SNP; inf: off) */
    double dlog10p_bound_below;   /* stated bound below the threshold */
} cs_pscan_info;
int cs_pcg_projected_scan(cs_kinship* K, cs_bed* bed, const int32_t* keep, int64_t n_keep,
                          const double* Y, int64_t ldy, int64_t T, const float* X, int64_t ldx,
                          int64_t c, double sigma_g, double sigma_e, const int32_t* cal_snps,
                          int64_t n_cal, const int32_t* snps, int64_t n_snps, double min_maf,
                          double rel_tol, int max_iter, int prec, double exact_log10p, double* freq,
                          int64_t* n_obs, int32_t* pass, double* beta, double* se, double* p,
                          double* log10p, cs_pscan_info* info);

/* ---- Randomized low-rank null model on the kinship operator (the matrix-free GWAS path).
 * LmmEngine::prepareLowRankMatrixFree (src/lmm_engine.cpp:897-1095): Omega from
 * std::mt19937_64(seed) + std::normal_distribution<double> (as the reference), K Omega, two power
 * iterations, B = Q^T K Q, eigendecomposition, the leading `rank` pairs, trace-fraction rank
 * selection (trace_threshold, min_rank), residual eigenvalue, REML (81-point grid on
 * log eta in [-12, 12] + golden section).  When the threshold is not met and no_fallback == 0 the
 * full rank is kept -- GwasScanner's fallback to MatrixFreeVce (src/matrix_free_vce.cpp:32-235),
 * the same computation with the threshold off.  y: n phenotypes over the operator's rows;
 * X: n x c float32 (intercept included).  Omega of the last call (n x sketch doubles) is kept in
 * host memory for the next call with the same seed and size (one call per LOCO chromosome). */
typedef struct cs_lowrank_diag {
    int requested_rank, effective_rank, sketch_size, total_rank;
    double residual_eigenvalue, trace, trace_fraction, relative_frobenius_error;
    int trace_threshold_met, fell_back;
    double sigma2_g, sigma2_e, delta, decomposition_seconds;
    int64_t estimated_peak_bytes;
} cs_lowrank_diag;
int cs_lowrank_prepare(cs_kinship* K, const double* y, const float* X, int64_t ldx, int64_t c, int rank,
                       int oversample, int64_t seed, double trace_threshold, int min_rank,
                       int no_fallback, cs_lowrank** out, cs_lowrank_diag* diag);
int cs_lowrank_destroy(cs_lowrank* lr);
/* Selected rank, eigenvalues (rank), U (n x rank, ldu), r0, sigma_g^2, sigma_e^2 (any may be NULL). */
int cs_lowrank_get(cs_lowrank* lr, int64_t* rank, double* lambda, double* U, int64_t ldu, double* r0,
                   double* sigma2_g, double* sigma2_e);
/* LmmEngine::scanLowRank (src/lmm_engine.cpp:1269-1361) over the keep list (the operator's rows):
 * outputs as cs_lmm_scan (NaN when not tested). */
int cs_lowrank_scan(cs_lowrank* lr, cs_bed* bed, const int32_t* keep, int64_t n_keep, const int32_t* snps,
                    int64_t n_snps, double min_maf, int prec, double* freq, int64_t* n_obs, int32_t* pass,
                    double* beta, double* se, double* p, double* log10p);

/* ---- Serving loop: scan requests over cohort subsets of one resident genotype matrix, run as
 * continuously batched multi-RHS PCG columns (one K.V apply per iteration for every admitted
 * request; the kinship matrix of a cohort is the restriction of the all-sample operator since the
 * standardisation uses all-sample statistics) and a PCG-projected fused scan per finished request
 * (cs_pcg_projected_scan's estimator, exact stage included: a request's selected markers come back
 * into the batch as columns of a second phase, in chunks that fit the free slots).  A request's
 * results do not depend on which other requests share the batch.  No reference counterpart: the
 * reference runs one phenotype per process. */
typedef struct cs_request {
    const int32_t* keep; int64_t n_keep;     /* cohort (distinct samples; NULL/0 = all) */
    const double* Y; int64_t ldy; int64_t n_traits;   /* n x T phenotypes over the cohort */
    const float* X; int64_t ldx; int64_t n_cov;       /* n x c covariates incl. intercept (1..32) */
    double sigma2_g, sigma2_e;               /* variance components of the null model */
    const int32_t* cal_snps; int64_t n_cal;  /* calibration SNPs */
    const int32_t* snps; int64_t n_snps;     /* SNPs to scan (NULL/0 = all) */
    double min_maf;
    double rel_tol;                          /* PCG relative residual (<= 0: 1e-6) */
    int max_iter;                            /* (<= 0: 200) */
    int prec;                                /* (0: CS_PREC_INT8_L3) */
    double exact_log10p;                     /* exact-stage threshold, as cs_pcg_projected_scan */
} cs_request;
typedef struct cs_request_stats {
    double t_submit, t_admit, t_done;        /* seconds since the server was created */
    int iterations;                          /* PCG iterations of the main solve */
    double gamma;
    int64_t n_calibration, n_columns;
    int64_t n_exact;                         /* SNPs recomputed by the exact stage */
    int exact_iterations;                    /* PCG iterations of the exact stage, all chunks */
    double calibration_rel_sd, dlog10p_bound_below;
    int64_t admit_bytes;                     /* device bytes charged at admission, results excluded */
    int compact_rows;                        /* 1: products on a compact copy of the cohort's rows */
} cs_request_stats;
/* max_columns: PCG columns held at once over all active requests; hbm_budget_bytes: bound on the
 * per-request device memory (<= 0: unbounded).  Each admitted request owns a kinship operator over
 * its cohort, whose products run on a compact copy of the cohort's genotype rows (2 roundup(n,
 * 256) roundup(M, 256) / 4 bytes, cs_kinship_compact_rows) when n <= 3/4 N and the budget can hold
 * it, else on the full matrix.  A request is admitted only when its columns fit and its device
 * memory fits the budget: the compact copy, operator lists, preconditioner and projection (held
 * to completion), the solve buffers and operator workspace (held while its columns are), and its
 * result arrays (n_snps x (4 T + 3) x 8 bytes, held until cs_server_release); exact-stage chunks
 * are sized to what is left.  A request that can never fit fails with CS_ERR_INVALID_ARGUMENT. */
int cs_server_create(cs_bed* bed, const int32_t* variants, int64_t n_variants, int64_t max_columns,
                     int64_t hbm_budget_bytes, cs_server** out);
int cs_server_destroy(cs_server* srv);
int cs_server_submit(cs_server* srv, const cs_request* req, int64_t* id);
/* One scheduler iteration: admit pending requests that fit, one PCG iteration (one K.V apply) of
 * every active request, finish the requests whose columns converged. */
int cs_server_step(cs_server* srv, int32_t* n_active, int32_t* n_pending, int32_t* n_completed);
/* status: 0 pending, 1 active, 2 done, -1 failed.  Asking for outputs of a failed request gives
 * CS_ERR_RUNTIME with its message.  When done, outputs as cs_pcg_projected_scan (any may be
 * NULL); a finished request keeps them in device memory (n_snps x (4 T + 3) x 8 bytes) until
 * cs_server_release, and each call copies the asked-for arrays to the host buffers given.
 * Threads: the server calls may come from several host threads.  cs_server_result copies a
 * finished request's arrays on its own stream without holding the scheduler, so a client thread
 * can fetch results while another thread runs cs_server_step; cs_server_release and
 * cs_server_destroy refuse (CS_ERR_INVALID_ARGUMENT) while a fetch of the request is running. */
int cs_server_result(cs_server* srv, int64_t id, int32_t* status, double* freq, int64_t* n_obs,
                     int32_t* pass, double* beta, double* se, double* p, double* log10p,
                     cs_request_stats* stats);
int cs_server_release(cs_server* srv, int64_t id);

/* ---- Dense null model and rotated scan (the default --make-grm path of
 * GwasScanner::runLocoScan, src/gwas_scanner.cpp:74-525).
 *
 * cs_lmm_create: y (n) and covariates X (n x c, float32 column-major with leading dimension ldx,
 * intercept included -- univariate_model.cpp:368-372) over the kept samples.  With c > 0 the
 * covariates are projected out (EigenVce with use_covariate_projection, the default): the
 * Householder QR of X is computed once, and every prepare applies its c reflectors to G.
 * Up to n - c = 32766 the eigendecomposition is rocSOLVER's dsyevd; beyond (rocSOLVER's 32-bit
 * interfaces) it is cs_sym_eigen_device's own 64-bit path, which needs about 6 n^2 doubles of
 * device memory in all. */
int cs_lmm_create(const double* y, const float* X, int64_t ldx, int64_t n, int64_t c, cs_lmm** out);
/* Dense symmetric eigendecomposition of the n x n device matrix A (lower triangle, leading
 * dimension lda): A is overwritten by the orthonormal eigenvectors, D (n doubles, device) by the
 * eigenvalues in ascending order -- the SelfAdjointEigenSolver step of EigenVce::prepare and
 * MvEigenVce::prepare (libCosmicCore/src/eigen_vce.cpp:34-38, 520-526).  leaf = 0: rocSOLVER's
 * dsyevd (n <= 32766).  leaf in 1..32766: the 64-bit path -- blocked Householder tridiagonal
 * reduction, divide and conquer over tridiagonal blocks of at most `leaf` rows (rocSOLVER's
 * dstedc on the blocks, rank-one merges with deflation and Gu-Eisenstat eigenvectors), compact-WY
 * back-transformation; any n that device memory holds (about 6 n^2 doubles).  Synchronous. */
int cs_sym_eigen_device(double* A, int64_t n, int64_t lda, double* D, int64_t leaf);
int cs_lmm_destroy(cs_lmm* lmm);
/* EigenVce::prepare (libCosmicCore/src/eigen_vce.cpp:19-49) for an n x n GRM G over the kept
 * samples (host memory, or device memory when g_on_device != 0): G_reml = A^T G A,
 * eigendecomposition (ascending, clamped at 0), W_proj = A U, Uty = U^T A^T y. */
int cs_lmm_prepare_dense(cs_lmm* lmm, const double* G, int64_t ldg, int g_on_device);
/* EigenVce::runNullModel with the default Brent solver = VCESolver::solveDiagonal
 * (libCosmicCore/src/vce.cpp:3073-3171): golden section on log10(delta) in [-5, 5], tol 1e-5.
 * Returns sigma_g^2, sigma_e^2 and delta = sigma_e^2 / sigma_g^2 (any may be NULL). */
int cs_lmm_null_brent(cs_lmm* lmm, double* sigma2_g, double* sigma2_e, double* delta);
/* Use given variance components instead of fitting them. */
int cs_lmm_set_variance(cs_lmm* lmm, double sigma2_g, double sigma2_e);
/* D (n-c), Uty (n-c) and W_proj (n x (n-c), leading dimension ldw); any may be NULL. */
int cs_lmm_get_eigen(cs_lmm* lmm, double* D, double* Uty, double* W_proj, int64_t ldw);
/* LmmEngine::scan (src/lmm_engine.cpp:1363-1525) with the fitted null model: per SNP in `snps`
 * (NULL/0 = all), over the keep list (the rows of y / X / G, in that order): the prepared marker
 * g, score = (W_proj^T g)^T V^-1 Uty, var = sum_l (W_proj^T g)_l^2 / (sigma_g D_l + sigma_e),
 * beta, se, p, log10p as cs_lmm_scan (NaN when not tested). */
int cs_lmm_scan_rotated(cs_lmm* lmm, cs_bed* bed, const int32_t* keep, int64_t n_keep,
                        const int32_t* snps, int64_t n_snps, double min_maf, int prec, double* freq,
                        int64_t* n_obs, int32_t* pass, double* beta, double* se, double* p,
                        double* log10p);
/* Dense LOCO GRM over the keep list from device raw GRMs (src/gwas_scanner.cpp:481-495):
 * G = raw / S_global restricted to keep x keep; when d_raw_chr is non-NULL, S_chr > 0 and
 * S_loco = S_global - S_chr > 1e-6: G * (S_global / S_loco) - raw_chr[keep, keep] / S_loco.
 * d_raw / d_raw_chr: N x N device; d_out: n_keep x n_keep device; keep NULL/0 = all samples. */
int cs_grm_loco_subset(const double* d_raw, const double* d_raw_chr, int64_t N, const int32_t* keep,
                       int64_t n_keep, double S_global, double S_chr, double* d_out);

/* validate_variance_components (libCosmicCore/src/validation.h:351-366), which both drivers run
 * after every null model (src/gwas_scanner.cpp:402-409): CS_OK when every component is finite and
 * non-negative, else CS_ERR_RUNTIME with the reference's report text ("[VARIANCE_NONFINITE]" or
 * "[VARIANCE_NEGATIVE]", then the context in parentheses). */
int cs_validate_variance_components(const double* v, int64_t n, const char* context);

/* ---- GWAS driver: the default dense PLINK path of GwasScanner::runLocoScan
 * (src/gwas_scanner.cpp:74-525) end to end on the GPU.  Reads bfile.{bed,bim,fam}; y has one
 * entry per .fam row (NaN = missing), cov is N x n_cov column-major (leading dimension ldc, NaN =
 * missing; an intercept is added).  Writes out_prefix.assoc.txt, .scan.meta.txt and
 * .cosmic.sumstats in the reference's formats.  LOCO over the .bim chromosome labels in
 * lexicographic order when there is more than one and no_loco is 0. */
typedef struct cs_gwas_options {
    double min_maf;        /* --maf, default 0.01 */
    int no_loco;           /* --no-loco */
    const char* scan_chr;  /* scan only this chromosome (NULL = all) */
    int prec;              /* scan precision, CS_PREC_INT8_L3 by default */
    int grm_alg;           /* --make-grm-alg: 1 (VanRaden, default) or 0 / 3 (standardised; the
                              LOCO scaling factors are then SNP counts, gwas_scanner.cpp:189,481) */
} cs_gwas_options;
#define CS_GWAS_MAX_MODELS 64
typedef struct cs_gwas_summary {
    int64_t n_samples, n_snps_grm, n_tested;
    double sum_2pq;
    int32_t loco, n_models;
    char model_label[CS_GWAS_MAX_MODELS][16];   /* "GLOBAL" or the chromosome label */
    double sigma2_g[CS_GWAS_MAX_MODELS], sigma2_e[CS_GWAS_MAX_MODELS];
    double ms_grm, ms_null, ms_scan;            /* wall time of the stages */
} cs_gwas_summary;
int cs_gwas_dense_run(const char* bfile, const double* y, const double* cov, int64_t ldc,
                      int64_t n_cov, const cs_gwas_options* opt, const char* out_prefix,
                      cs_gwas_summary* summary);
/* The matrix-free path (--fastlmm --fastlmm-matrix-free, src/gwas_scanner.cpp:194-198,257-372):
 * GRM variants = inv_std != 0 and all-sample MAF >= min_maf (and >= 1e-8); per LOCO chromosome a
 * kinship operator without that chromosome.  mode 0 (the reference path): randomized low-rank REML
 * (cs_lowrank_prepare with max_rank / oversample / seed / trace_threshold / min_rank /
 * no_fallback; a relative Frobenius error above error_tol also falls back to the full rank) and
 * scanLowRank; writes .low_rank.meta.tsv as well.  mode 1: PCG-projected scan
 * (cs_pcg_projected_scan) with sigma2_g / sigma2_e, or -- when sigma2_e <= 0 -- the variance
 * components of one global low-rank REML; n_calibration calibration SNPs per chromosome drawn with
 * std::mt19937_64(seed). */
typedef struct cs_gwas_mf_options {
    double min_maf;        /* 0.01 */
    int no_loco;
    const char* scan_chr;
    int prec;              /* CS_PREC_INT8_L3 */
    int mode;              /* 0 low-rank (reference), 1 PCG-projected */
    int max_rank;          /* --fastlmm-max-rank, 256 */
    int oversample;        /* --lmm-low-rank-oversample, 10 */
    int64_t seed;          /* --lmm-low-rank-seed, 42 */
    double trace_threshold;/* --fastlmm-trace-threshold, 0.95 */
    int min_rank;          /* --fastlmm-min-rank, 1 */
    int no_fallback;       /* --fastlmm-no-fallback */
    double error_tol;      /* --lmm-low-rank-error-tol, 1.0 */
    double sigma2_g, sigma2_e;   /* mode 1: fixed variance components (sigma2_e <= 0: estimate) */
    double pcg_tol;        /* mode 1: relative residual, 1e-6 */
    int pcg_max_iter;      /* mode 1: 200 */
    int n_calibration;     /* mode 1: calibration SNPs per chromosome, 32 */
    double exact_log10p;   /* mode 1: exact-stage threshold (cs_pcg_projected_scan; 0 = 5) */
} cs_gwas_mf_options;
int cs_gwas_matrix_free_run(const char* bfile, const double* y, const double* cov, int64_t ldc,
                            int64_t n_cov, const cs_gwas_mf_options* opt, const char* out_prefix,
                            cs_gwas_summary* summary);

/* ---- Multivariate LMM (d traits, 1 <= d <= 6): MvEigenVce + MvLmmEngine with covariate
 * projection (the defaults of MultivariateModel, src/models/multivariate_model.cpp:870-1110).
 *
 * cs_mvlmm_create: Y (n x d, column-major, ldy) and covariates X (n x c float32, intercept
 * included) over the kept samples; the Householder QR of X is computed once.  The
 * eigendecomposition takes the 64-bit path beyond n - c = 32766, as cs_lmm_create. */
int cs_mvlmm_create(const double* Y, int64_t ldy, int64_t n, int64_t d, const float* X, int64_t ldx,
                    int64_t c, cs_mvlmm** out);
int cs_mvlmm_destroy(cs_mvlmm* mv);
/* MvEigenVce::prepare (libCosmicCore/src/eigen_vce.cpp:496-549) for an n x n GRM over the kept
 * samples (host memory, or device memory when g_on_device != 0): A^T G A, eigendecomposition
 * (clamped at 0), W_proj = A U, Uty = U^T A^T Y, start values Vg = Ve = Cov(Y) / 2 (Cov with the
 * divisor n - c - 1 of the projected sample count, as the reference). */
int cs_mvlmm_prepare_dense(cs_mvlmm* mv, const double* G, int64_t ldg, int g_on_device);
/* MvEigenVce::runNullModel (eigen_vce.cpp:1031-1066): solver 0 = EM (updateEM until the
 * normalised squared change is below tol, at most max_iter iterations; converged = 0 otherwise,
 * which the reference turns into an error), solver 1 = runAireml (:631-910: EM warm-up of at
 * least em_min_iter iterations, eigenvalue clipping, AI-REML with step halving, ai_max_iter,
 * ai_tol; standard errors from the final AI matrix unless it stopped on ai_tol).  Reference
 * defaults: max_iter 1000 (--reml-maxit), tol 1e-8, em_min_iter 10, ai_max_iter 50, ai_tol 1e-8. */
int cs_mvlmm_null(cs_mvlmm* mv, int solver, int max_iter, double tol, int em_min_iter, int ai_max_iter,
                  double ai_tol, int* converged, int* em_iterations, int* ai_iterations);
/* Vg, Ve, their standard errors (NaN when not computed) as d x d row-major, and the REML
 * log-likelihood (calcLogLikelihood, eigen_vce.cpp:617-632); any may be NULL. */
int cs_mvlmm_get(cs_mvlmm* mv, double* Vg, double* Ve, double* Vg_se, double* Ve_se, double* logl);
int cs_mvlmm_set_variance(cs_mvlmm* mv, const double* Vg, const double* Ve);
/* D (n-c), Uty ((n-c) x d column-major), W_proj (n x (n-c), leading dimension ldw); any may be NULL. */
int cs_mvlmm_get_eigen(cs_mvlmm* mv, double* D, double* Uty, double* W_proj, int64_t ldw);
/* MvLmmEngine::scan(PlinkReader&) (src/mvlmm_engine.cpp:756-931, processBlock :210-394) with the
 * fitted (or set) null model, over the keep list (the rows of Y / X / G, in that order): per SNP
 * the centred marker (mean of the observed genotypes of the kept samples, missing -> 0; no MAF
 * filter), s = W_proj^T g, H = sum_c s_c^2 Sigma_c^-1, r = sum_c s_c Sigma_c^-1 y_c, beta = H^-1 r
 * (symmetric eigendecomposition, singular below 128 eps d max(1, |lambda|max)), Wald = beta^T H
 * beta, p = chi^2_d tail, vbeta = H^-1 conditioned as conditionCovariance
 * (src/mvlmm_numerics.h:27-66).  Outputs per requested SNP: freq (mean / 2), beta (n_snps x d),
 * vbeta (n_snps x d x d), wald, p; status 0 / 1 / 2 (OK / NUMERICAL_ROUNDOFF /
 * BOUNDARY_SOLUTION) when tested, -1 (NaN outputs) when not.  The reference reads the first n
 * rows of the .fam instead of the kept rows (mvlmm_engine.cpp:760-766, 858-866); both agree when
 * every sample is kept. */
int cs_mvlmm_scan(cs_mvlmm* mv, cs_bed* bed, const int32_t* keep, int64_t n_keep, const int32_t* snps,
                  int64_t n_snps, double* freq, double* beta, double* vbeta, double* wald, double* p,
                  int32_t* status);

/* ---- Multivariate GWAS driver: MultivariateModel::run's fixed-null Wald path
 * (src/models/multivariate_model.cpp:870-1110) end to end on the GPU.  Y is N x d (one row per
 * .fam sample, column-major, ldy; NaN = missing), cov as cs_gwas_dense_run.  GRM with grm_alg;
 * LOCO over the chromosomes in numeric order (non-numeric labels compared as strings) unless
 * no_loco or a single chromosome, with G_loco = G_norm (S / S_loco) - G_chr,norm (S_chr / S_loco)
 * and S, S_chr the sums of 2pq (:1040-1063).  Writes out_prefix.assoc.mv.txt (CHR SNP BP A1 A2
 * FREQ WALD_STAT P BETA_t SE_t ...; multivariate_output.cpp:42-70). */
typedef struct cs_gwas_mv_options {
    double min_maf;        /* GRM SNP filter, 0.01 */
    int no_loco;
    const char* scan_chr;  /* NULL = all */
    int grm_alg;           /* 1 */
    int solver;            /* 0 EM (default), 1 AI-REML (--mv-vce-solver aireml) */
    int max_iter;          /* --reml-maxit, 1000 */
    double tol;            /* --reml-tol, 1e-8 */
    int em_min_iter;       /* 10 */
    int ai_max_iter;       /* 50 */
    double ai_tol;         /* 1e-8 */
} cs_gwas_mv_options;
typedef struct cs_gwas_mv_summary {
    int64_t n_samples, n_traits, n_tested;
    int32_t loco, n_models;
    char model_label[CS_GWAS_MAX_MODELS][16];
    double vg[CS_GWAS_MAX_MODELS][36], ve[CS_GWAS_MAX_MODELS][36];   /* d x d row-major */
    double ms_grm, ms_null, ms_scan;
} cs_gwas_mv_summary;
int cs_gwas_mv_run(const char* bfile, const double* Y, int64_t ldy, int64_t n_traits, const double* cov,
                   int64_t ldc, int64_t n_cov, const cs_gwas_mv_options* opt, const char* out_prefix,
                   cs_gwas_mv_summary* summary);
/* LOCO GRM of the multivariate driver (multivariate_model.cpp:1040-1063) over the keep list from
 * device raw GRMs: G = raw / norm restricted to keep x keep; with d_raw_chr non-NULL, S_chr > 0 and
 * S_loco = S - S_chr > 1e-6: (raw / norm) (S / S_loco) - (raw_chr / norm_chr) (S_chr / S_loco). */
int cs_grm_loco_subset_scaled(const double* d_raw, const double* d_raw_chr, int64_t N, const int32_t* keep,
                              int64_t n_keep, double norm, double norm_chr, double S, double S_chr,
                              double* d_out);

#ifdef __cplusplus
}
#endif
#endif
