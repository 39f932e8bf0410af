// Host drivers of CosmicScanner's PLINK GWAS (GwasScanner::runLocoScan, src/gwas_scanner.cpp:74-525),
// written over the library's own C-ABI.
//
// Shared by both drivers:
//   * samples: DataFilter::filterSamples (libCosmicCore/src/data_filter.h:21-44) keeps rows with a
//     finite phenotype and finite covariates; fewer than 10 is an error (univariate_model.cpp:425);
//     X = [1, covariates as float32] (univariate_model.cpp:366-374);
//   * chromosome labels in std::set (lexicographic) order; LOCO when there is more than one and
//     --no-loco is off; --scan-chr restricts the scan;
//   * output: .assoc.txt (kUnifiedHeader, additive rows, NaN-P rows skipped, default ostream
//     formatting: gwas_output.cpp:123-184, gwas_output_schema.h:14-38), .scan.meta.txt
//     (gwas_scanner.cpp:93-116 and appendCommonMetadata, gwas_output_schema.h:57-71) and
//     .cosmic.sumstats (gwas_output.cpp:187-239).
//
// cs_gwas_dense_run: the default dense path -- exact global and per-chromosome raw GRMs on the GPU,
//   LOCO GRM (:481-495), EigenVce::prepare + Brent null model, rotated scan per chromosome.
// cs_gwas_matrix_free_run: the --fastlmm --fastlmm-matrix-free path (:194-198, 257-315, 322-372):
//   variant selection on all-sample stats, one kinship operator per LOCO chromosome, randomized
//   low-rank REML with the trace-threshold / error-tolerance fallbacks, scanLowRank, and
//   .low_rank.meta.tsv (:101-107, 410-428).  Mode 1 replaces the low-rank model by the
//   PCG-projected scan (variance components from one global low-rank REML, exact solves per LOCO
//   operator, calibrated variance).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <limits>
#include <memory>
#include <random>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "cs_common.h"

namespace {

using csamd::InvalidArgument;
using csamd::RuntimeError;

void chk(int rc) {
    if (rc == CS_OK) return;
    const std::string msg = cs_last_error();
    if (rc == CS_ERR_INVALID_ARGUMENT) throw InvalidArgument(msg);
    throw RuntimeError(msg);
}

struct BimRow {
    std::string chrom, id;
    double cm = 0.0;
    long long bp = 0;
    std::string alt, ref;   // .bim columns 5 (A1) and 6 (A2), plink_reader.cpp:61
};

std::vector<BimRow> read_bim(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw RuntimeError("Cannot open " + path);
    std::vector<BimRow> rows;
    std::string line;
    while (std::getline(f, line)) {
        if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
        std::istringstream ss(line);
        BimRow r;
        ss >> r.chrom >> r.id >> r.cm >> r.bp >> r.alt >> r.ref;
        rows.push_back(std::move(r));
    }
    return rows;
}

struct DevMem {
    void* p = nullptr;
    explicit DevMem(int64_t bytes) { chk(cs_dev_alloc(bytes, &p)); }
    ~DevMem() { if (p) cs_dev_free(p); }
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    double* d() const { return static_cast<double*>(p); }
};

template <class T, int (*Destroy)(T*)>
struct Handle {
    T* h = nullptr;
    Handle() = default;
    Handle(const Handle&) = delete;
    Handle& operator=(const Handle&) = delete;
    ~Handle() { reset(); }
    void reset() { if (h) Destroy(h); h = nullptr; }
};
using BedHandle = Handle<cs_bed, cs_bed_destroy>;
using LmmHandle = Handle<cs_lmm, cs_lmm_destroy>;
using KinHandle = Handle<cs_kinship, cs_kinship_destroy>;
using LowRankHandle = Handle<cs_lowrank, cs_lowrank_destroy>;

struct Result {   // the GwasResult fields written by the additive writer (lmm_engine.h:22-53)
    std::string chrom, snp_id, a1, a2;
    long long pos = 0;
    double freq = 0.0, beta = 0.0, se = 0.0;
    double p_value = std::numeric_limits<double>::quiet_NaN();
    double log10_p = 0.0;
    int n_obs = 0;
};

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Genotypes, sample filter and design matrix.
struct Inputs {
    std::vector<BimRow> bim;
    int64_t N = 0, M = 0, n = 0, c = 0;
    BedHandle bed;
    std::vector<int32_t> keep;
    std::vector<double> y;         // n x d, column-major
    std::vector<float> X;          // n x c, column-major, intercept first
    std::set<std::string> labels;
};

// y: N x d phenotypes (column-major, ldy); a sample is kept when every trait and covariate is
// finite (the multivariate filter of MultivariateModel::loadCovariates,
// src/models/multivariate_model.cpp:830-867, is the same rule over d traits).
void load_inputs(Inputs& in, const char* bfile, const double* y, const double* cov, int64_t ldc, int64_t n_cov,
                 int64_t d = 1, int64_t ldy = 0) {
    const std::string prefix(bfile);
    in.bim = read_bim(prefix + ".bim");
    std::ifstream fb(prefix + ".bed", std::ios::binary);
    if (!fb) throw RuntimeError("Cannot open " + prefix + ".bed");
    std::vector<uint8_t> image((std::istreambuf_iterator<char>(fb)), std::istreambuf_iterator<char>());
    {
        std::ifstream ff(prefix + ".fam");
        if (!ff) throw RuntimeError("Cannot open " + prefix + ".fam");
        std::string line;
        while (std::getline(ff, line))
            if (line.find_first_not_of(" \t\r") != std::string::npos) ++in.N;
    }
    in.M = (int64_t)in.bim.size();
    chk(cs_bed_create_from_file_image(image.data(), (int64_t)image.size(), in.N, in.M, &in.bed.h));
    if (n_cov > 0 && ldc < in.N) throw RuntimeError("DataFilter: X and y dimension mismatch");
    if (d > 1 && ldy < in.N) throw RuntimeError("multivariate: Y dimension mismatch");
    for (int64_t i = 0; i < in.N; ++i) {
        bool k = true;
        for (int64_t t = 0; t < d; ++t)
            if (!std::isfinite(y[t * ldy + i])) k = false;
        if (d == 1) k = !std::isnan(y[i]);
        for (int64_t j = 0; k && j < n_cov; ++j)
            if (std::isnan((float)cov[j * ldc + i])) k = false;
        if (k) in.keep.push_back((int32_t)i);
    }
    in.n = (int64_t)in.keep.size();
    if (in.n < 10) throw RuntimeError("Too few valid samples.");
    in.c = n_cov + 1;
    in.y.resize((size_t)(in.n * d));
    in.X.resize((size_t)(in.n * in.c));
    for (int64_t r = 0; r < in.n; ++r) {
        for (int64_t t = 0; t < d; ++t) in.y[(size_t)(t * in.n + r)] = y[t * ldy + in.keep[(size_t)r]];
        in.X[(size_t)r] = 1.0f;
        for (int64_t j = 0; j < n_cov; ++j)
            in.X[(size_t)((j + 1) * in.n + r)] = (float)cov[j * ldc + in.keep[(size_t)r]];
    }
    for (const auto& b : in.bim) in.labels.insert(b.chrom);
}

std::vector<int32_t> chrom_snps(const Inputs& in, const std::string& chrom) {
    std::vector<int32_t> s;
    for (int64_t j = 0; j < in.M; ++j)
        if (in.bim[(size_t)j].chrom == chrom) s.push_back((int32_t)j);
    return s;
}

void append_results(const Inputs& in, const std::vector<int32_t>& snps, const std::vector<double>& freq,
                    const std::vector<double>& beta, const std::vector<double>& se, const std::vector<double>& p,
                    const std::vector<double>& l10, std::vector<Result>& out) {
    for (size_t i = 0; i < freq.size(); ++i) {
        const int64_t j = snps.empty() ? (int64_t)i : snps[i];
        Result r;
        r.chrom = in.bim[(size_t)j].chrom;
        r.snp_id = in.bim[(size_t)j].id;
        r.pos = in.bim[(size_t)j].bp;
        r.a1 = in.bim[(size_t)j].alt;
        r.a2 = in.bim[(size_t)j].ref;
        r.freq = freq[i];
        r.n_obs = (int)in.n;
        r.beta = beta[i];
        r.se = se[i];
        r.p_value = p[i];
        r.log10_p = l10[i];
        out.push_back(std::move(r));
    }
}

void write_scan_meta_header(const std::string& out, const char* approx, int requested_rank, double threshold,
                            const char* selection) {
    std::ofstream meta(out + ".scan.meta.txt");
    if (!meta) throw RuntimeError("Cannot create " + out + ".scan.meta.txt");
    meta << "Item\tValue\n"
         << "analysis_model\tunivariate_lmm\n"
         << "requested_scan_tier\tlegacy_default\n"
         << "resolved_scan_tier\tapprox\n"
         << "snp_vc_strategy\tfixed_null\n"
         << "SNP_MODEL\tadditive\n"
         << "HWE_APPLIED\tfalse\n"
         << "HWE_METHOD\tNONE\n"
         << "HWE_REASON_IF_SKIPPED\tFILTER_DISABLED\n"
         << "matrix_approximation\t" << approx << "\n"
         << "requested_rank\t" << requested_rank << "\n"
         << "trace_threshold\t" << threshold << "\n"
         << "RANK_SELECTION\t" << selection << "\n";
}

int64_t write_outputs(const std::string& out, const std::vector<Result>& results, int64_t n, bool loco) {
    int64_t n_tested = 0;
    {
        std::ofstream f(out + ".assoc.txt");
        if (!f) throw RuntimeError("Cannot create " + out + ".assoc.txt");
        f << "CHR\tSNP\tBP\tA1\tA2\tFREQ\tN"
             "\tBETA_A\tSE_A\tP_A\tLOG10P_A"
             "\tBETA_D\tSE_D\tP_D\tLOG10P_D"
             "\tP_JOINT\tLOG10P_JOINT\n";
        for (const Result& r : results) {
            if (std::isnan(r.p_value)) continue;
            f << r.chrom << '\t' << r.snp_id << '\t' << r.pos << '\t' << r.a1 << '\t' << r.a2 << '\t' << r.freq
              << '\t' << r.n_obs << '\t' << r.beta << '\t' << r.se << '\t' << r.p_value << '\t' << r.log10_p
              << "\tNA\tNA\tNA\tNA\tNA\tNA\n";
            ++n_tested;
        }
    }
    {
        std::ofstream f(out + ".cosmic.sumstats");
        if (!f) throw RuntimeError("Cannot create " + out + ".cosmic.sumstats");
        const std::time_t now = std::time(nullptr);
        std::ostringstream date;
        date << std::put_time(std::gmtime(&now), "%Y-%m-%d %H:%M:%S UTC");
        f << "# CosmicScanner Summary Statistics v1.0\n# Created: " << date.str() << "\n# Model: LMM\n#\n"
          << "CHR\tSNP\tPOS\tA1\tA2\tBETA\tSE\tP\tN\tAF\tLOG10P\tMODEL\tZ\tWEIGHT\tINFO\tWARNING\n";
        for (const Result& r : results) {
            if (std::isnan(r.p_value)) continue;
            f << r.chrom << '\t' << r.snp_id << '\t' << r.pos << '\t' << r.a1 << '\t' << r.a2 << '\t'
              << std::setprecision(10) << r.beta << '\t' << std::setprecision(10) << r.se << '\t'
              << std::setprecision(10) << r.p_value << '\t' << r.n_obs << '\t' << std::setprecision(6) << r.freq
              << '\t' << std::setprecision(6) << r.log10_p << "\tLMM\t" << std::setprecision(6)
              << (r.se > 0 ? r.beta / r.se : 0.0) << '\t' << std::setprecision(6)
              << (r.se > 0 ? 1.0 / (r.se * r.se) : 0.0) << '\t' << std::setprecision(4) << 1.0 << "\t.\n";
        }
    }
    {
        std::ofstream meta(out + ".scan.meta.txt", std::ios::app);
        meta << "MODEL\tadditive\n"
             << "N\t" << n << "\n"
             << "N_SEMANTICS\tanalysis_samples_after_phenotype_and_covariate_filtering;"
                "per-SNP_missing_genotypes_mean-imputed\n"
             << "FREQ_SEMANTICS\tA1_allele_frequency_(not_MAF)\n"
             << "LOCO\t" << (loco ? "true" : "false") << "\n"
             << "COVARIATE_PROJECTION\ttrue\n"
             << "VCE_CONVERGED\ttrue\n";
    }
    return n_tested;
}

void fill_summary(cs_gwas_summary* s, const Inputs& in, int64_t nv, double S, bool loco, int64_t n_tested,
                  const std::vector<std::string>& labels, const std::vector<double>& vg,
                  const std::vector<double>& ve, double ms_grm, double ms_null, double ms_scan) {
    if (!s) return;
    std::memset(s, 0, sizeof(*s));
    s->n_samples = in.n;
    s->n_snps_grm = nv;
    s->sum_2pq = S;
    s->loco = loco ? 1 : 0;
    s->n_tested = n_tested;
    s->n_models = (int32_t)labels.size();
    for (size_t i = 0; i < labels.size() && i < CS_GWAS_MAX_MODELS; ++i) {
        std::strncpy(s->model_label[i], labels[i].c_str(), 15);
        s->sigma2_g[i] = vg[i];
        s->sigma2_e[i] = ve[i];
    }
    s->ms_grm = ms_grm;
    s->ms_null = ms_null;
    s->ms_scan = ms_scan;
}

}  // namespace

// validate_variance_components (libCosmicCore/src/validation.h:351-366), run after every null
// model (src/gwas_scanner.cpp:402-409): the first NaN / Inf or negative component fails with the
// reference's report text (ValidationReport::summary, validation.h:103-116).
extern "C" int cs_validate_variance_components(const double* v, int64_t n, const char* context) {
    return csamd::guarded([&] {
        if (n > 0 && !v) throw csamd::InvalidArgument("cs_validate_variance_components: null argument");
        const char* code = nullptr;
        const char* msg = nullptr;
        for (int64_t i = 0; i < n && !code; ++i) {
            if (!std::isfinite(v[i])) {
                code = "VARIANCE_NONFINITE";
                msg = "variance component is NaN or Inf";
            } else if (v[i] < 0.0) {
                code = "VARIANCE_NEGATIVE";
                msg = "variance component is negative";
            }
        }
        if (!code) return;
        std::ostringstream os;
        os << "Validation : FAIL\nChecks     : 0 passed, 0 warnings, 1 failures\n  [" << code << "] " << msg;
        if (context && context[0]) os << " (" << context << ")";
        os << "\n";
        throw csamd::RuntimeError(os.str());
    });
}

extern "C" int cs_gwas_dense_run(const char* bfile, const double* y, const double* cov, int64_t ldc,
                                 int64_t n_cov, const cs_gwas_options* opt, const char* out_prefix,
                                 cs_gwas_summary* summary) {
    return csamd::guarded([&] {
        if (!bfile || !y || !out_prefix || (n_cov > 0 && !cov))
            throw InvalidArgument("cs_gwas_dense_run: null argument");
        cs_gwas_options o{0.01, 0, nullptr, CS_PREC_INT8_L3, 1};
        if (opt) o = *opt;
        const std::string out(out_prefix);
        Inputs in;
        load_inputs(in, bfile, y, cov, ldc, n_cov);
        const int64_t N = in.N, M = in.M, n = in.n;
        const bool use_loco = in.labels.size() > 1 && !o.no_loco;
        const std::string scan_chr = o.scan_chr ? std::string(o.scan_chr) : std::string();
        write_scan_meta_header(out, "full_eigen", 0, 0.0, "FULL");

        auto t0 = std::chrono::steady_clock::now();
        double ms_grm = 0, ms_null = 0, ms_scan = 0;
        DevMem raw(N * N * 8), gsub(n * n * 8);
        double S_global = 0.0;
        int64_t nv = 0;
        const bool std_alg = o.grm_alg == 0 || o.grm_alg == 3;
        chk(cs_grm_build_raw(in.bed.h, o.grm_alg, (float)o.min_maf, 0.0f, nullptr, 0, raw.d(), 1, &S_global, nullptr, &nv));
        if (std_alg) S_global = (double)nv;
        ms_grm += ms_since(t0);
        LmmHandle lmm;
        chk(cs_lmm_create(in.y.data(), in.X.data(), n, n, in.c, &lmm.h));

        std::vector<Result> results;
        std::vector<std::string> model_labels;
        std::vector<double> vg_all, ve_all;
        auto run_scan_with_grm = [&](const std::vector<int32_t>& snps, const std::string& label) {
            auto t1 = std::chrono::steady_clock::now();
            chk(cs_lmm_prepare_dense(lmm.h, gsub.d(), n, 1));
            double vg = 0, ve = 0;
            chk(cs_lmm_null_brent(lmm.h, &vg, &ve, nullptr));
            const double vc[2] = {vg, ve};
            chk(cs_validate_variance_components(vc, 2, "CosmicScanner null model"));
            ms_null += ms_since(t1);
            model_labels.push_back(label.empty() ? "GLOBAL" : label);
            vg_all.push_back(vg);
            ve_all.push_back(ve);
            t1 = std::chrono::steady_clock::now();
            const int64_t ns = snps.empty() ? M : (int64_t)snps.size();
            std::vector<double> freq((size_t)ns), beta((size_t)ns), se((size_t)ns), p((size_t)ns), l10((size_t)ns);
            chk(cs_lmm_scan_rotated(lmm.h, in.bed.h, in.keep.data(), n, snps.empty() ? nullptr : snps.data(), ns,
                                    o.min_maf, o.prec, freq.data(), nullptr, nullptr, beta.data(), se.data(),
                                    p.data(), l10.data()));
            ms_scan += ms_since(t1);
            append_results(in, snps, freq, beta, se, p, l10, results);
        };

        if (!use_loco) {
            chk(cs_grm_loco_subset(raw.d(), nullptr, N, in.keep.data(), n, S_global, 0.0, gsub.d()));
            std::vector<int32_t> snps;
            if (!scan_chr.empty()) {
                snps = chrom_snps(in, scan_chr);
                if (snps.empty()) throw RuntimeError("No chromosome matched the requested additive scan filters");
            }
            run_scan_with_grm(snps, scan_chr);
        } else {
            DevMem raw_chr(N * N * 8);
            for (const std::string& chrom : in.labels) {
                if (!scan_chr.empty() && chrom != scan_chr) continue;
                const std::vector<int32_t> snps = chrom_snps(in, chrom);
                auto t1 = std::chrono::steady_clock::now();
                double S_chr = 0.0;
                int64_t nv_chr = 0;
                chk(cs_grm_build_raw(in.bed.h, o.grm_alg, (float)o.min_maf, 0.0f, snps.data(), (int64_t)snps.size(),
                                     raw_chr.d(), 1, nullptr, &S_chr, &nv_chr));
                if (std_alg) S_chr = (double)nv_chr;
                chk(cs_grm_loco_subset(raw.d(), raw_chr.d(), N, in.keep.data(), n, S_global, S_chr, gsub.d()));
                ms_grm += ms_since(t1);
                run_scan_with_grm(snps, chrom);
            }
            if (model_labels.empty()) throw RuntimeError("No chromosome matched the requested additive scan filters");
        }
        const int64_t n_tested = write_outputs(out, results, n, use_loco);
        fill_summary(summary, in, nv, S_global, use_loco, n_tested, model_labels, vg_all, ve_all, ms_grm, ms_null,
                     ms_scan);
    });
}

extern "C" int cs_gwas_matrix_free_run(const char* bfile, const double* y, const double* cov, int64_t ldc,
                                       int64_t n_cov, const cs_gwas_mf_options* opt, const char* out_prefix,
                                       cs_gwas_summary* summary) {
    return csamd::guarded([&] {
        if (!bfile || !y || !out_prefix || (n_cov > 0 && !cov) || !opt)
            throw InvalidArgument("cs_gwas_matrix_free_run: null argument");
        const cs_gwas_mf_options o = *opt;
        if (o.mode != 0 && o.mode != 1) throw InvalidArgument("cs_gwas_matrix_free_run: mode must be 0 or 1");
        const std::string out(out_prefix);
        Inputs in;
        load_inputs(in, bfile, y, cov, ldc, n_cov);
        const int64_t M = in.M, n = in.n;
        const bool use_loco = in.labels.size() > 1 && !o.no_loco;
        const std::string scan_chr = o.scan_chr ? std::string(o.scan_chr) : std::string();
        write_scan_meta_header(out, o.mode == 0 ? "fastlmm_auto_trace" : "pcg_projected", o.max_rank,
                               o.trace_threshold, o.mode == 0 ? "AUTO_TRACE" : "NONE");
        {
            std::ofstream meta(out + ".low_rank.meta.tsv");
            meta << "SCOPE\tALGORITHM\tREQUESTED_RANK\tEFFECTIVE_RANK\tSKETCH_SIZE"
                 << "\tRESIDUAL_EIGENVALUE\tRELATIVE_FROBENIUS_ERROR\tERROR_TOLERANCE"
                 << "\tDECOMPOSITION_SECONDS\tESTIMATED_PEAK_BYTES"
                 << "\tFULL_EIGEN_ESTIMATED_PEAK_BYTES\tSELECTED_RANK\tTOTAL_RANK"
                 << "\tTRACE_FRACTION\tRANK_SELECTION\tFALLBACK\n";
        }
        // variant selection on all-sample statistics (gwas_scanner.cpp:257-277)
        chk(cs_bed_compute_stats(in.bed.h));
        std::vector<double> means((size_t)M), inv((size_t)M);
        chk(cs_bed_get_stats(in.bed.h, means.data(), inv.data()));
        std::vector<int32_t> grm_vars;
        for (int64_t j = 0; j < M; ++j) {
            if (inv[(size_t)j] == 0.0) continue;
            const double f = means[(size_t)j] / 2.0;
            const double maf = std::min(f, 1.0 - f);
            if (maf < o.min_maf || std::abs(maf) < 1e-8) continue;
            grm_vars.push_back((int32_t)j);
        }
        if (grm_vars.empty()) throw RuntimeError("Matrix-free GRM selection produced no variants");
        const int rank_req = o.max_rank;
        const int eff_rank = std::max(1, std::min(rank_req, (int)n - 1));

        std::vector<Result> results;
        std::vector<std::string> model_labels;
        std::vector<double> vg_all, ve_all;
        double ms_null = 0, ms_scan = 0;
        double pcg_sg = o.sigma2_g, pcg_se = o.sigma2_e;
        if (o.mode == 1 && !(pcg_sg >= 0.0 && pcg_se > 0.0)) {
            // variance components from one global low-rank REML
            auto t1 = std::chrono::steady_clock::now();
            KinHandle K;
            chk(cs_kinship_create(in.bed.h, in.keep.data(), n, grm_vars.data(), (int64_t)grm_vars.size(), &K.h));
            LowRankHandle lr;
            chk(cs_lowrank_prepare(K.h, in.y.data(), in.X.data(), n, in.c, eff_rank, o.oversample, o.seed,
                                   o.trace_threshold, o.min_rank, o.no_fallback, &lr.h, nullptr));
            chk(cs_lowrank_get(lr.h, nullptr, nullptr, nullptr, 0, nullptr, &pcg_sg, &pcg_se));
            ms_null += ms_since(t1);
        }
        if (o.mode == 1) {
            const double vc[2] = {pcg_sg, pcg_se};
            chk(cs_validate_variance_components(vc, 2, "CosmicScanner null model"));
        }
        auto run_chrom = [&](const std::string& chrom) {
            std::vector<int32_t> loco_vars;
            if (use_loco && !chrom.empty()) {
                for (int32_t j : grm_vars)
                    if (in.bim[(size_t)j].chrom != chrom) loco_vars.push_back(j);
            } else {
                loco_vars = grm_vars;
            }
            if (loco_vars.empty())
                throw RuntimeError("Matrix-free LOCO left no GRM variants for chromosome " + chrom);
            const std::vector<int32_t> snps = chrom.empty() ? std::vector<int32_t>() : chrom_snps(in, chrom);
            const int64_t ns = snps.empty() ? M : (int64_t)snps.size();
            std::vector<double> freq((size_t)ns), beta((size_t)ns), se((size_t)ns), p((size_t)ns), l10((size_t)ns);
            KinHandle K;
            chk(cs_kinship_create(in.bed.h, in.keep.data(), n, loco_vars.data(), (int64_t)loco_vars.size(), &K.h));
            auto t1 = std::chrono::steady_clock::now();
            if (o.mode == 0) {
                LowRankHandle lr;
                cs_lowrank_diag d{};
                std::string fallback = "none";
                chk(cs_lowrank_prepare(K.h, in.y.data(), in.X.data(), n, in.c, eff_rank, o.oversample, o.seed,
                                       o.trace_threshold, o.min_rank, o.no_fallback, &lr.h, &d));
                if (d.fell_back) {
                    fallback = "trace_threshold";
                } else if (d.relative_frobenius_error > o.error_tol) {
                    // error-tolerance fallback: MatrixFreeVce with the threshold off (full rank)
                    fallback = "error_tolerance";
                    cs_lowrank_diag d2{};
                    lr.reset();
                    chk(cs_lowrank_prepare(K.h, in.y.data(), in.X.data(), n, in.c, eff_rank, o.oversample, o.seed,
                                           0.0, o.min_rank, 1, &lr.h, &d2));
                }
                double vg = 0, ve = 0;
                chk(cs_lowrank_get(lr.h, nullptr, nullptr, nullptr, 0, nullptr, &vg, &ve));
                ms_null += ms_since(t1);
                const double vc[2] = {vg, ve};
                chk(cs_validate_variance_components(vc, 2, "CosmicScanner null model"));
                {
                    std::ofstream meta(out + ".low_rank.meta.tsv", std::ios::app);
                    meta << (chrom.empty() ? "GLOBAL" : chrom) << "\trandomized_subspace_iteration_tail_mean_woodbury"
                         << "\t" << d.effective_rank << "\t" << d.effective_rank << "\t" << d.sketch_size << "\t"
                         << std::setprecision(12) << d.residual_eigenvalue << "\t" << d.relative_frobenius_error
                         << "\t" << o.error_tol << "\t" << d.decomposition_seconds << "\t" << d.estimated_peak_bytes
                         << "\t" << (24LL * n * n) << "\t" << d.effective_rank << "\t" << d.total_rank << "\t"
                         << d.trace_fraction << "\tMATRIX_FREE_AUTO_TRACE\t" << fallback << "\n";
                }
                model_labels.push_back(chrom.empty() ? "GLOBAL" : chrom);
                vg_all.push_back(vg);
                ve_all.push_back(ve);
                t1 = std::chrono::steady_clock::now();
                chk(cs_lowrank_scan(lr.h, in.bed.h, in.keep.data(), n, snps.empty() ? nullptr : snps.data(), ns,
                                    o.min_maf, o.prec, freq.data(), nullptr, nullptr, beta.data(), se.data(), p.data(),
                                    l10.data()));
                ms_scan += ms_since(t1);
            } else {
                // calibration SNPs: a seeded sample of the scanned SNPs
                std::vector<int32_t> pool = snps;
                if (pool.empty()) {
                    pool.resize((size_t)M);
                    for (int64_t j = 0; j < M; ++j) pool[(size_t)j] = (int32_t)j;
                }
                std::mt19937_64 gen((uint64_t)o.seed);
                std::shuffle(pool.begin(), pool.end(), gen);
                pool.resize(std::min<size_t>(pool.size(), (size_t)std::max(1, o.n_calibration)));
                std::vector<int64_t> nobs((size_t)ns);
                std::vector<int32_t> ps((size_t)ns);
                cs_pscan_info info{};
                chk(cs_pcg_projected_scan(K.h, in.bed.h, in.keep.data(), n, in.y.data(), n, 1, in.X.data(), n, in.c,
                                          pcg_sg, pcg_se, pool.data(), (int64_t)pool.size(),
                                          snps.empty() ? nullptr : snps.data(), ns, o.min_maf, o.pcg_tol, o.pcg_max_iter,
                                          o.prec, o.exact_log10p, freq.data(), nobs.data(), ps.data(), beta.data(), se.data(), p.data(),
                                          l10.data(), &info));
                ms_scan += ms_since(t1);
                model_labels.push_back(chrom.empty() ? "GLOBAL" : chrom);
                vg_all.push_back(pcg_sg);
                ve_all.push_back(pcg_se);
            }
            append_results(in, snps, freq, beta, se, p, l10, results);
        };
        if (!use_loco) {
            if (!scan_chr.empty() && chrom_snps(in, scan_chr).empty())
                throw RuntimeError("No chromosome matched the requested additive scan filters");
            run_chrom(scan_chr);
        } else {
            for (const std::string& chrom : in.labels) {
                if (!scan_chr.empty() && chrom != scan_chr) continue;
                run_chrom(chrom);
            }
            if (model_labels.empty()) throw RuntimeError("No chromosome matched the requested additive scan filters");
        }
        const int64_t n_tested = write_outputs(out, results, n, use_loco);
        fill_summary(summary, in, (int64_t)grm_vars.size(), (double)M, use_loco, n_tested, model_labels, vg_all,
                     ve_all, 0.0, ms_null, ms_scan);
    });
}

// ---------------------------------------------------------------------------------------------
// Multivariate driver: MultivariateModel::run's fixed-null Wald path
// (src/models/multivariate_model.cpp:870-1110) over the library's C-ABI.
namespace {

using MvHandle = Handle<cs_mvlmm, cs_mvlmm_destroy>;

// std::sort comparator of the LOCO chromosome order (multivariate_model.cpp:1022-1024).
bool chrom_less(const std::string& a, const std::string& b) {
    try {
        return std::stoi(a) < std::stoi(b);
    } catch (...) {
        return a < b;
    }
}

void write_mv_header(const std::string& path, int64_t d) {
    std::ofstream f(path);
    if (!f) throw RuntimeError("Cannot create " + path);
    f << "CHR\tSNP\tBP\tA1\tA2\tFREQ\tWALD_STAT\tP";
    for (int64_t t = 0; t < d; ++t) f << "\tBETA_" << (t + 1) << "\tSE_" << (t + 1);
    f << "\n";
}

// MultivariateModel::writeResults (multivariate_output.cpp:42-70): rows with a finite P.
int64_t append_mv_rows(const std::string& path, const Inputs& in, const std::vector<int32_t>& snps, int64_t d,
                       const std::vector<double>& freq, const std::vector<double>& beta,
                       const std::vector<double>& vbeta, const std::vector<double>& wald,
                       const std::vector<double>& p) {
    std::ofstream f(path, std::ios::app);
    int64_t rows = 0;
    for (size_t i = 0; i < snps.size(); ++i) {
        if (std::isnan(p[i])) continue;
        const BimRow& b = in.bim[(size_t)snps[i]];
        f << b.chrom << '\t' << b.id << '\t' << b.bp << '\t' << b.alt << '\t' << b.ref << '\t' << freq[i] << '\t'
          << wald[i] << '\t' << p[i];
        for (int64_t t = 0; t < d; ++t) {
            const double var = vbeta[(i * d + t) * d + t];
            if (!std::isfinite(var) || var < 0.0)
                throw RuntimeError("mvLMM covariance TRUE_SINGULARITY for SNP " + b.id + ": invalid trait variance");
            f << '\t' << beta[i * d + t] << '\t' << std::sqrt(var);
        }
        f << '\n';
        ++rows;
    }
    return rows;
}

}  // namespace

extern "C" int cs_gwas_mv_run(const char* bfile, const double* Y, int64_t ldy, int64_t n_traits, const double* cov,
                              int64_t ldc, int64_t n_cov, const cs_gwas_mv_options* opt, const char* out_prefix,
                              cs_gwas_mv_summary* summary) {
    return csamd::guarded([&] {
        if (!bfile || !Y || !out_prefix || (n_cov > 0 && !cov))
            throw InvalidArgument("cs_gwas_mv_run: null argument");
        if (n_traits < 1 || n_traits > 6) throw InvalidArgument("multivariate LMM: 1 to 6 traits are supported");
        cs_gwas_mv_options o{0.01, 0, nullptr, 1, 0, 1000, 1e-8, 10, 50, 1e-8};
        if (opt) o = *opt;
        const int64_t d = n_traits;
        Inputs in;
        load_inputs(in, bfile, Y, cov, ldc, n_cov, d, ldy);
        const int64_t N = in.N, M = in.M, n = in.n;
        const std::string scan_chr = o.scan_chr ? std::string(o.scan_chr) : std::string();
        const std::string path = std::string(out_prefix) + ".assoc.mv.txt";
        const bool std_alg = o.grm_alg == 0 || o.grm_alg == 3;
        auto t0 = std::chrono::steady_clock::now();
        double ms_grm = 0, ms_null = 0, ms_scan = 0;
        DevMem raw(N * N * 8), gsub(n * n * 8);
        double S_global = 0.0;
        int64_t nv = 0;
        chk(cs_grm_build_raw(in.bed.h, o.grm_alg, (float)o.min_maf, 0.0f, nullptr, 0, raw.d(), 1, &S_global, nullptr, &nv));
        const double norm_global = std_alg ? (double)nv : S_global;
        ms_grm += ms_since(t0);
        write_mv_header(path, d);
        std::vector<std::string> labels;
        std::vector<std::vector<double>> vgs, ves;
        int64_t n_tested = 0;
        auto fit_and_scan = [&](const std::vector<int32_t>& snps, const std::string& label, bool loco_model) {
            auto t1 = std::chrono::steady_clock::now();
            MvHandle mv;
            chk(cs_mvlmm_create(in.y.data(), n, n, d, in.X.data(), n, in.c, &mv.h));
            chk(cs_mvlmm_prepare_dense(mv.h, gsub.d(), n, 1));
            int conv = 0;
            chk(cs_mvlmm_null(mv.h, o.solver, o.max_iter, o.tol, o.em_min_iter, o.ai_max_iter, o.ai_tol, &conv,
                              nullptr, nullptr));
            if (!conv) {
                if (loco_model)
                    throw RuntimeError("Multivariate LOCO null model did not converge for chromosome " + label +
                                       "; stopping association scan. Check the model/data or increase --reml-maxit.");
                throw RuntimeError("Multivariate null model did not converge; no association scan was performed. "
                                   "Check the model/data or increase --reml-maxit.");
            }
            std::vector<double> vg((size_t)(d * d)), ve((size_t)(d * d));
            chk(cs_mvlmm_get(mv.h, vg.data(), ve.data(), nullptr, nullptr, nullptr));
            ms_null += ms_since(t1);
            labels.push_back(label);
            vgs.push_back(vg);
            ves.push_back(ve);
            t1 = std::chrono::steady_clock::now();
            const size_t ns = snps.size();
            std::vector<double> freq(ns), beta(ns * d), vbeta(ns * d * d), wald(ns), p(ns);
            std::vector<int32_t> status(ns);
            if (ns > 0)
                chk(cs_mvlmm_scan(mv.h, in.bed.h, in.keep.data(), n, snps.data(), (int64_t)ns, freq.data(),
                                  beta.data(), vbeta.data(), wald.data(), p.data(), status.data()));
            ms_scan += ms_since(t1);
            n_tested += append_mv_rows(path, in, snps, d, freq, beta, vbeta, wald, p);
        };
        const bool use_loco = !(o.no_loco || in.labels.size() <= 1);
        if (!use_loco) {
            chk(cs_grm_loco_subset_scaled(raw.d(), nullptr, N, in.keep.data(), n, norm_global, 1.0, S_global, 0.0,
                                          gsub.d()));
            std::vector<int32_t> snps;
            for (int64_t j = 0; j < M; ++j)
                if (scan_chr.empty() || in.bim[(size_t)j].chrom == scan_chr) snps.push_back((int32_t)j);
            fit_and_scan(snps, "GLOBAL", false);
        } else {
            std::vector<std::string> order(in.labels.begin(), in.labels.end());
            std::sort(order.begin(), order.end(), chrom_less);
            DevMem raw_chr(N * N * 8);
            for (const std::string& chrom : order) {
                if (!scan_chr.empty() && chrom != scan_chr) continue;
                const std::vector<int32_t> snps = chrom_snps(in, chrom);
                auto t1 = std::chrono::steady_clock::now();
                double S_chr = 0.0;
                int64_t nv_chr = 0;
                chk(cs_grm_build_raw(in.bed.h, o.grm_alg, (float)o.min_maf, 0.0f, snps.data(), (int64_t)snps.size(),
                                     raw_chr.d(), 1, &S_chr, nullptr, &nv_chr));
                const double norm_chr = std_alg ? (double)nv_chr : S_chr;
                chk(cs_grm_loco_subset_scaled(raw.d(), raw_chr.d(), N, in.keep.data(), n, norm_global, norm_chr,
                                              S_global, S_chr, gsub.d()));
                ms_grm += ms_since(t1);
                fit_and_scan(snps, chrom, true);
            }
        }
        if (summary) {
            std::memset(summary, 0, sizeof(*summary));
            summary->n_samples = n;
            summary->n_traits = d;
            summary->n_tested = n_tested;
            summary->loco = use_loco ? 1 : 0;
            summary->n_models = (int32_t)labels.size();
            for (size_t i = 0; i < labels.size() && i < CS_GWAS_MAX_MODELS; ++i) {
                std::strncpy(summary->model_label[i], labels[i].c_str(), 15);
                for (int64_t k = 0; k < d * d; ++k) {
                    summary->vg[i][k] = vgs[i][(size_t)k];
                    summary->ve[i][k] = ves[i][(size_t)k];
                }
            }
            summary->ms_grm = ms_grm;
            summary->ms_null = ms_null;
            summary->ms_scan = ms_scan;
        }
    });
}
