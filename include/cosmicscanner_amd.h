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
 * FP32:    plain FP32 VALU kernels (decode to FP32, FMA, no MFMA, no split) -- the comparator. */
#define CS_PREC_INT8_L3 3
#define CS_PREC_INT8_L2 2
#define CS_PREC_FP32 32

typedef struct cs_bed cs_bed;
typedef struct cs_kinship cs_kinship;

/* Thread-local message of the last failed call on this thread ("" if none). */
const char* cs_last_error(void);
/* Library version as major*10000 + minor*100 + patch. */
int cs_version(void);

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
/* apply / apply_block / applyTranspose (kinship_operator.cpp:211-304): Y = K X for X n x k
 * column-major with leading dimension ldx; Y may alias X; k == 0 is allowed.  ldx < n or
 * n_threads-like misuse -> CS_ERR_RUNTIME "dimension mismatch". */
int cs_kinship_apply(cs_kinship* op, const double* X, int64_t ldx, int64_t k, double* Y,
                     int64_t ldy);
/* Same with X, Y in device memory (no host transfer); for device-resident solvers. */
int cs_kinship_apply_device(cs_kinship* op, const double* dX, int64_t ldx, int64_t k, double* dY,
                            int64_t ldy);
/* trace() and getDiagonal() (kinship_operator.cpp:306-368). */
int cs_kinship_trace(cs_kinship* op, double* trace);
int cs_kinship_diag(cs_kinship* op, double* diag);
/* Time of the last apply's kernels in ms per phase (0: quantize+scatter, 1: Z^T product,
 * 2: middle epilogue, 3: Z product, 4: final epilogue), from HIP events on the operator's stream. */
int cs_kinship_last_timing(const cs_kinship* op, double* ms5);

#ifdef __cplusplus
}
#endif
#endif
