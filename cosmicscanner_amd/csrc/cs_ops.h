// Internal helpers shared by the translation units of the library (not part of the C-ABI).
#pragma once
#include "cs_common.h"

namespace csamd {

constexpr double kQmax3 = 127.0 * 65536.0;  // |q| bound for 3 balanced limbs (top limb <= 127)
constexpr double kQmax2 = 127.0 * 256.0;

inline int grid_for(int64_t total, int block = 256) {
    int64_t g = (total + block - 1) / block;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (int)g;
}

// Column tiling of an int8 product with k right-hand sides in L limbs: NT tiles of 16 columns,
// NTB tiles per block (grid.z = NT / NTB), NC = 16 NT columns in the output.
struct Tiling {
    int NT, NTB, NC;
};
Tiling tiling(int64_t k, int L);

void quantize_cols(const double* W, int64_t rows, int64_t k, int L, const Tiling& tl,
                   hipStream_t st, DevBuf& mx, DevBuf& scl, DevBuf& B);

template <bool TWO_B>
void launch_mfma(int NTB, dim3 grid, hipStream_t st, const uint8_t* A, int64_t lda,
                 int64_t rows_pad, int64_t kbps, int64_t kbt, int NT, const int8_t* Bx,
                 const int8_t* Bo, int32_t* out);

void ensure_stats(cs_bed* b);
void ensure_sample_major(cs_bed* b);
void weighted_counts(cs_bed* b, const int32_t* d_w_npad, int64_t* d_counts, hipStream_t st);

__global__ void k_scatter(const double* __restrict__ V, int64_t ldv, int64_t k,
                          const int32_t* __restrict__ inv_off, const int32_t* __restrict__ inv_idx,
                          int64_t N, int64_t Npad, double* __restrict__ Vt);

}  // namespace csamd
