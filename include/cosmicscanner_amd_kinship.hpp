// C++ host adapter over the C-ABI: the reference's matrix-free kinship interface
// (src/kinship_operator.h:19-47) served by the MI355X library.
//
// GpuKinshipOperator mirrors PlinkKinshipOperator's public surface -- rows / cols /
// numVariants / apply (vector and block) / applyTranspose / trace / getDiagonal / isDense -- with
// the same argument meaning and error behaviour: std::runtime_error for a null reader, indices
// out of range, an empty variant set or a dimension mismatch (kinship_operator.cpp:34-63,78,228),
// std::invalid_argument for a non-positive thread count (:76,141,230).  Vectors and matrices are
// column-major double buffers (Eigen's default layout), so the adapter binds to Eigen::VectorXd /
// MatrixXd via .data(); this header itself does not depend on Eigen.  Outputs may alias inputs.
#pragma once
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "cosmicscanner_amd.h"

namespace cosmic_amd {

inline void check(int rc) {
    if (rc == CS_OK) return;
    const std::string msg = cs_last_error();
    if (rc == CS_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

// Resident genotypes (PlinkReader's role for this path, libCosmicCore/src/plink_reader.h).
class GpuPlinkGenotypes {
public:
    GpuPlinkGenotypes(const uint8_t* bed_body, int64_t n_samples, int64_t n_snps) {
        check(cs_bed_create(bed_body, n_samples, n_snps, &h_));
        n_ = n_samples;
        m_ = n_snps;
    }
    ~GpuPlinkGenotypes() { if (h_) cs_bed_destroy(h_); }
    GpuPlinkGenotypes(const GpuPlinkGenotypes&) = delete;
    GpuPlinkGenotypes& operator=(const GpuPlinkGenotypes&) = delete;

    int64_t numSamples() const { return n_; }
    int64_t numSnps() const { return m_; }
    void computeStats(int = 1) { check(cs_bed_compute_stats(h_)); }
    void getStats(std::vector<double>& means, std::vector<double>& inv_stds) {
        means.resize((size_t)m_);
        inv_stds.resize((size_t)m_);
        check(cs_bed_get_stats(h_, means.data(), inv_stds.data()));
    }
    cs_bed* handle() const { return h_; }

private:
    cs_bed* h_ = nullptr;
    int64_t n_ = 0, m_ = 0;
};

class GpuKinshipOperator {
public:
    // createKinshipOperator(reader, keep, variants) (kinship_operator.h:98-100): empty lists mean
    // all samples / all SNPs; keep may be in any order and repeat samples.  The genotypes must
    // outlive the operator (kinship_operator.h:83).
    GpuKinshipOperator(GpuPlinkGenotypes* genotypes, const std::vector<int>& keep = {},
                       const std::vector<int>& variants = {}) {
        if (!genotypes) throw std::runtime_error("PlinkKinshipOperator: null reader");
        std::vector<int32_t> k(keep.begin(), keep.end()), v(variants.begin(), variants.end());
        check(cs_kinship_create(genotypes->handle(), k.empty() ? nullptr : k.data(), (int64_t)k.size(),
                                v.empty() ? nullptr : v.data(), (int64_t)v.size(), &h_));
        check(cs_kinship_rows(h_, &rows_));
        check(cs_kinship_num_variants(h_, &m_));
    }
    ~GpuKinshipOperator() { if (h_) cs_kinship_destroy(h_); }
    GpuKinshipOperator(const GpuKinshipOperator&) = delete;
    GpuKinshipOperator& operator=(const GpuKinshipOperator&) = delete;

    int rows() const { return (int)rows_; }
    int cols() const { return (int)rows_; }
    int numVariants() const { return (int)m_; }
    bool isDense() const { return false; }
    void setPrecision(int mode) { check(cs_kinship_set_precision(h_, mode)); }

    // out = K v ; v has rows() entries.  out may alias v.
    void apply(const double* v, int64_t v_size, std::vector<double>& out, int n_threads = 1) const {
        applyBlock(v, v_size, 1, out, n_threads);
    }
    // out = K X ; X is rows() x k column-major (k may be 0).  out may alias X.
    void applyBlock(const double* X, int64_t x_rows, int64_t k, std::vector<double>& out,
                    int n_threads = 1) const {
        if (n_threads < 1) throw std::invalid_argument("Kinship block threads must be positive");
        if (x_rows != rows_) throw std::runtime_error("PlinkKinshipOperator::apply: dimension mismatch");
        std::vector<double> result((size_t)(rows_ * k));
        if (k > 0) check(cs_kinship_apply(h_, X, rows_, k, result.data(), rows_));
        out.swap(result);
    }
    void applyTranspose(const double* v, int64_t v_size, std::vector<double>& out,
                        int n_threads = 1) const {
        apply(v, v_size, out, n_threads);   // K is symmetric
    }
    double trace() const {
        double t = 0.0;
        check(cs_kinship_trace(h_, &t));
        return t;
    }
    void getDiagonal(std::vector<double>& diag) const {
        diag.resize((size_t)rows_);
        check(cs_kinship_diag(h_, diag.data()));
    }

private:
    cs_kinship* h_ = nullptr;
    int64_t rows_ = 0, m_ = 0;
};

inline std::unique_ptr<GpuKinshipOperator> createKinshipOperator(
    GpuPlinkGenotypes* genotypes, const std::vector<int>& keep = {},
    const std::vector<int>& variants = {}) {
    return std::make_unique<GpuKinshipOperator>(genotypes, keep, variants);
}

}  // namespace cosmic_amd
