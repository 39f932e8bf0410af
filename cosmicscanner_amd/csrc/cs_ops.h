// Internal helpers shared by the translation units of the library (not part of the C-ABI).
#pragma once
#include "cs_common.h"

namespace csamd {

constexpr double kQmax3 = 127.0 * 65536.0;  // |q| bound for 3 balanced limbs (top limb <= 127)
constexpr double kQmax2 = 127.0 * 256.0;
constexpr double kQmax4 = 127.0 * 16777216.0;   // 4 limbs: 31-bit fixed point
inline int prec_limbs(int prec) { return prec == CS_PREC_INT8_L2 ? 2 : prec == CS_PREC_INT8_L4 ? 4 : 3; }
inline double limbs_qmax(int L) { return L == 2 ? kQmax2 : L == 4 ? kQmax4 : kQmax3; }
// Current operand precision of a kinship operator (CS_PREC_*).
int kinship_get_precision(const cs_kinship* op);

inline int grid_for(int64_t total, int block = 256) {
    int64_t g = (total + block - 1) / block;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (int)g;
}

// rocSOLVER's dorgqr / dormqr / dsyevd exist only with 32-bit (rocblas_int) interfaces and keep
// their n x n work arrays (LAPACK's 1 + 6 n + 2 n^2 elements for dsyevd) under 2^31 elements:
// n <= 32766.  Larger dense eigendecompositions take sym_eig_large.
constexpr int64_t kDenseEigenMaxN = 32766;

// Whether the genotype products run the LDS-shared kernels (CSAMD_GENO_LDSB unset or not 0).
bool geno_lds_enabled();

// Operand tiling of k right-hand sides with L limbs: G groups of 16 columns and NT = G L limb
// tiles, or -- packed, when k L <= 16 and allow_pack (LDS kernels only) -- one tile holding limb
// l of column c in tile column c L + l.
struct Tiling {
    int G, L, NT;
    bool packed;
};
inline Tiling tiling(int64_t k, int L, bool allow_pack = true) {
    if (allow_pack && k >= 1 && k * L <= 16 && geno_lds_enabled()) return Tiling{1, L, 1, true};
    const int G = (int)((k + 15) / 16);
    return Tiling{G, L, G * L, false};
}

// Quantise the columns of a row-major [rows][k] fp64 matrix (rows = padded k extent of the
// product) into the permuted int8 limb layout of the MFMA B operand; a common per-column scale
// over every matrix in W (all must have the same shape).  scl[0:k] = scales, scl[k:2k] = inverses.
void quantize_cols(const double* const* W, int nW, int64_t rows, int64_t k, const Tiling& tl,
                   hipStream_t st, DevBuf& mx, DevBuf& scl, int8_t* const* B);

// The same with caller-provided column maxima (k) and scales (2 k); allocates nothing.
void quantize_cols_into(const double* const* W, int nW, int64_t rows, int64_t k, const Tiling& tl,
                        hipStream_t st, unsigned long long* mx, double* scl, int8_t* const* B);

// Launch k_geno<L, TWO_B, EPI> (cs_geno_mfma.h) on the given grid.
struct GenoArgs;
void launch_geno(int L, bool two_b, int epi, dim3 grid, hipStream_t st, const GenoArgs& a,
                 bool only_x = false, bool packed = false);

void ensure_stats(cs_bed* b);
void ensure_sample_major(cs_bed* b);
// CSR lists of the missing genotypes (both orientations); idempotent.
void ensure_missing_lists(cs_bed* b);
// Frees a genotype handle (no release of the device-memory cache: for the row copies that
// kinship operators own, whose blocks the next cohort reuses).
void bed_free(cs_bed* b);
// A new resident matrix of the cohort keep[0..n) (host indices, caller's order): both tiled
// copies gathered from the parent's rows, the cohort's own class counts, and the parent's
// all-sample means and inverse standard deviations.  Freed with cs_bed_destroy.
cs_bed* bed_rows_copy(cs_bed* parent, const int32_t* keep, int64_t n);
// Device bytes of such a copy.
int64_t bed_rows_copy_bytes(const cs_bed* parent, int64_t n);
// Keep the kinship operator's genotype products on a compact copy of its cohort's rows (no-op
// for the identity cohort or when already compact); results are bitwise those of the
// full-sample products.  Needs a keep list without repeated samples.
void kinship_compact_rows(cs_kinship* op);
bool kinship_is_compact(const cs_kinship* op);
// The operator's compact copy of its cohort's rows (rows in the operator's order), or null.
cs_bed* kinship_rows_bed(cs_kinship* op);
// Upper bound of the device workspace one apply of k columns grows in an operator whose products
// read Npad sample rows and Mpad SNPs (both padded to 256) at precision prec.
int64_t kinship_work_bytes(int64_t Npad, int64_t Mpad, int prec, int64_t k);
// Free the operator's grown workspaces (the next apply grows them again).
void kinship_release_workspace(cs_kinship* op);
// K.V (dY = K dX, device buffers) enqueued on st without a host wait.
void kinship_apply_async(cs_kinship* op, const double* dX, int64_t ldx, int64_t k, double* dY, int64_t ldy,
                         hipStream_t st);
// Number of missing genotypes over all samples and SNPs (from the class counts; cached).
int64_t missing_total(cs_bed* b);
// Per-SNP class counts weighted by the host multiplicities w[0..n_w) (zero beyond); counts is
// M x 4 int64 on the device, by code 00, 01, 10, 11.  Synchronises st.
void weighted_counts(cs_bed* b, const int32_t* w, int64_t n_w, int64_t* d_counts, hipStream_t st);
// The same class counts for the SNPs `snps` only, into host memory ([ns][4]).
void weighted_counts_of(cs_bed* b, const int32_t* w, int64_t n_w, const int32_t* snps, int64_t ns, int64_t* h_counts,
                        hipStream_t st);

// Per-SNP scan over the keep list (cs_scan.hip): marker preparation, g^T R in int8 MFMA, and --
// with a model -- the fused statistics.  R is host memory, or device memory when device_R.
struct ScanArgs {
    const int32_t* keep = nullptr;
    int64_t n = 0;
    const double* R = nullptr;
    int64_t ldr = 0, nr = 0;
    double min_maf = 0.0;
    const int32_t* snps = nullptr;
    int64_t n_snps = 0;
    int prec = CS_PREC_INT8_L3;
    bool device_R = false;
};
struct ScanOut {
    double *freq = nullptr, *dot = nullptr, *gtg = nullptr;
    int64_t* n_obs = nullptr;
    int32_t* pass = nullptr;
    double *beta = nullptr, *se = nullptr, *p = nullptr, *log10p = nullptr;
    bool device = false;   // the output pointers are device memory (no host copy)
};
void run_scan(cs_bed* b, const ScanArgs& a, const cs_scan_model* model, const ScanOut& o);
// Projected scan stages (cs_pscan.hip); need rocblas.h in the including file for the handle.
#ifdef ROCBLAS_H
void pscan_prepare_rhs(cs_bed* bed, const int32_t* keep, int64_t n_keep, int64_t n, const double* Y, int64_t ldy,
                       int64_t T, const float* X, int64_t ldx, int64_t c, const int32_t* cal_snps, int64_t n_cal,
                       double min_maf, DevBuf& B, int64_t& L, hipStream_t st);
// Projection state of one solve: R = [P y_1..T, X] (n x (T + c), device), ZX = V^-1 X (n x c,
// device), C = (X^T V^-1 X)^-1, (X^T X)^-1 (c x c column-major), the calibration factor and the
// relative spread of the calibration markers' ratios (NaN with fewer than two).
struct PscanProj {
    int64_t n = 0, T = 0, c = 0;
    DevBuf R, ZX;
    std::vector<double> C, XtXi;
    double gamma = 0.0, rel_sd = 0.0;
    int64_t ncal = 0;
};
void pscan_project(int64_t n, int64_t T, int64_t c, int64_t L, const double* B, const double* Z, PscanProj& pj,
                   hipStream_t st, rocblas_handle blas);
void pscan_scan(cs_bed* bed, const int32_t* keep, const PscanProj& pj, const int32_t* snps, int64_t n_snps,
                double min_maf, int prec, double* freq, int64_t* n_obs, int32_t* pass, double* beta, double* se,
                double* p, double* log10p, bool device_out);
// Rows (indices into the scanned list, ascending) that pass and whose largest -log10 P over the
// T traits reaches thr (thr < 0: every tested row; +inf: none); log10p / pass on the device when
// `device`.
std::vector<int64_t> pscan_select(const double* log10p, const int32_t* pass, int64_t ns, int64_t T, double thr,
                                  bool device, hipStream_t st);
// Exact statistics of S selected rows from their prepared markers dG and an approximate solution
// dZg of V Z = dG (n x S, device), with dAux = V dZg (aux_is_residual false) or the residual
// dG - V dZg (true).  g^T V^-1 g comes from the variational form g^T z + z^T r, whose error
// -r^T V^-1 r is quadratic in the residual, so a loose solve tolerance suffices.  Overwrites
// rows[s] of the row-major ns x T outputs (device memory when device_out, else host).
void pscan_exact(const PscanProj& pj, int64_t S, const int64_t* rows, const double* dG, const double* dZg,
                 const double* dAux, bool aux_is_residual, double* beta, double* se, double* p, double* log10p,
                 bool device_out, hipStream_t st, rocblas_handle blas);
// Dense symmetric eigendecomposition on 64-bit interfaces (cs_eig.hip): A (n x n, ld lda, lower
// triangle, device) is overwritten by the eigenvectors, D (device) by the eigenvalues in ascending
// order; tridiagonal blocks of at most `leaf` rows go to rocSOLVER's dstedc.
void sym_eig_large(rocblas_handle h, hipStream_t st, int64_t n, double* A, int64_t lda, double* D, int64_t leaf);
// E (m x ncols) <- Q E or Q^T E for Q = H(0) ... H(c-1), H(i) = I - tau_i v_i v_i^T, v_i zero above
// row i, 1 on it, V(r, i) below (the geqrf layout); tau on the host.
void apply_reflectors(rocblas_handle h, hipStream_t st, const double* V, int64_t ldv, const double* tau_h, int64_t m,
                      int64_t c, bool transpose, double* E, int64_t lde, int64_t ncols);
#endif
// Leaf size of the 64-bit eigen path forced on every dense eigendecomposition by
// CSAMD_EIG_LARGE=<leaf> (0: unset -- the path runs only beyond rocSOLVER's limits).
int64_t eig_large_leaf();
// Whether a dense eigendecomposition of order n takes the 64-bit path.
inline bool eig_use_large(int64_t n) { return n > kDenseEigenMaxN || eig_large_leaf() > 0; }
// Exact-stage threshold of the projected scans on the largest -log10 P over T traits: the given
// level (0 -> the default 5) plus log10 T, i.e. the smallest P Bonferroni-adjusted over the traits
// reaches 10^-level; < 0 -> -1 (every tested SNP); +inf -> off; NaN is rejected.
constexpr double kPscanExactDefault = 5.0;
double pscan_threshold(double exact_log10p, int64_t T);
// Solve tolerance of the exact stage from the request's PCG tolerance (the variational variance
// is accurate to the square of the residual).
inline double pscan_exact_tol(double rel_tol) { return rel_tol * 100.0 < 1e-4 ? rel_tol * 100.0 : 1e-4; }
// Stated bound on |d(-log10 P)| of the calibrated statistics below threshold thr, from the
// calibration markers' relative spread (0 when thr < 0; NaN when off or unknown).
double pscan_bound_below(double thr, double rel_sd);
// Prepared markers of `snps` over the keep list (NULL = all samples) into dG (n x ns, device).
void markers_device(cs_bed* b, const int32_t* keep, int64_t n_keep, const int32_t* snps, int64_t ns,
                    double min_maf, double* dG, int64_t ldg, int32_t* pass_out, double* freq_out,
                    hipStream_t st);

__global__ void k_scatter(const double* __restrict__ V, int64_t ldv, int64_t k,
                          const int32_t* __restrict__ inv_off, const int32_t* __restrict__ inv_idx,
                          int64_t N, int64_t Npad, double* __restrict__ Vt);

}  // namespace csamd
