// Host driver of the default (dense, LOCO) PLINK GWAS of GwasScanner::runLocoScan
// (src/gwas_scanner.cpp:74-525), written over the library's own C-ABI:
//
//   * samples: DataFilter::filterSamples (libCosmicCore/src/data_filter.h:21-44) keeps rows with a
//     finite phenotype and finite covariates; fewer than 10 is an error (univariate_model.cpp:425);
//     X = [1, covariates as float32] (univariate_model.cpp:366-374);
//   * GRM: exact global raw Z Z^T and, with more than one chromosome, one raw GRM per chromosome
//     (GrmBuilder::computeSplit, grm_builder.cpp:401-785) on the GPU; chromosome labels in
//     std::set (lexicographic) order;
//   * per chromosome: LOCO GRM (gwas_scanner.cpp:481-495), EigenVce::prepare + Brent null model,
//     rotated scan of that chromosome's SNPs (LmmEngine::scan with chrom_filter);
//   * output: .assoc.txt (header kUnifiedHeader, additive rows, NaN-P rows skipped, default
//     ostream formatting: gwas_output.cpp:123-184, gwas_output_schema.h:14-38), .scan.meta.txt
//     (gwas_scanner.cpp:93-116 and appendCommonMetadata, gwas_output_schema.h:57-71) and
//     .cosmic.sumstats (gwas_output.cpp:187-239).
#include <chrono>
#include <cmath>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <limits>
#include <memory>
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

struct LmmHandle {
    cs_lmm* h = nullptr;
    ~LmmHandle() { if (h) cs_lmm_destroy(h); }
};

struct BedHandle {
    cs_bed* h = nullptr;
    ~BedHandle() { if (h) cs_bed_destroy(h); }
};

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

}  // namespace

extern "C" int cs_gwas_dense_run(const char* bfile, const double* y, const double* cov, int64_t ldc,
                                 int64_t n_cov, const cs_gwas_options* opt, const char* out_prefix,
                                 cs_gwas_summary* summary) {
    return csamd::guarded([&] {
        if (!bfile || !y || !out_prefix || (n_cov > 0 && !cov))
            throw InvalidArgument("cs_gwas_dense_run: null argument");
        cs_gwas_options o{0.01, 0, nullptr, CS_PREC_INT8_L3};
        if (opt) o = *opt;
        const std::string prefix(bfile), out(out_prefix);
        const std::vector<BimRow> bim = read_bim(prefix + ".bim");
        // .bed image -> resident genotypes
        std::ifstream fb(prefix + ".bed", std::ios::binary);
        if (!fb) throw RuntimeError("Cannot open " + prefix + ".bed");
        std::vector<uint8_t> image((std::istreambuf_iterator<char>(fb)), std::istreambuf_iterator<char>());
        int64_t N = 0;
        {
            std::ifstream ff(prefix + ".fam");
            if (!ff) throw RuntimeError("Cannot open " + prefix + ".fam");
            std::string line;
            while (std::getline(ff, line))
                if (line.find_first_not_of(" \t\r") != std::string::npos) ++N;
        }
        const int64_t M = (int64_t)bim.size();
        BedHandle bed;
        chk(cs_bed_create_from_file_image(image.data(), (int64_t)image.size(), N, M, &bed.h));
        image.clear();
        image.shrink_to_fit();
        if (n_cov > 0 && ldc < N) throw RuntimeError("DataFilter: X and y dimension mismatch");

        // sample filter and design matrix (float32 covariates; NaN tested after the cast)
        std::vector<int32_t> keep;
        for (int64_t i = 0; i < N; ++i) {
            bool k = !std::isnan(y[i]);
            for (int64_t j = 0; k && j < n_cov; ++j)
                if (std::isnan((float)cov[j * ldc + i])) k = false;
            if (k) keep.push_back((int32_t)i);
        }
        const int64_t n = (int64_t)keep.size();
        if (n < 10) throw RuntimeError("Too few valid samples.");
        const int64_t c = n_cov + 1;
        std::vector<double> ys((size_t)n);
        std::vector<float> X((size_t)(n * c));
        for (int64_t r = 0; r < n; ++r) {
            ys[(size_t)r] = y[keep[(size_t)r]];
            X[(size_t)r] = 1.0f;
            for (int64_t j = 0; j < n_cov; ++j)
                X[(size_t)((j + 1) * n + r)] = (float)cov[j * ldc + keep[(size_t)r]];
        }

        std::set<std::string> labels;
        for (const auto& b : bim) labels.insert(b.chrom);
        const bool use_loco = labels.size() > 1 && !o.no_loco;
        const std::string scan_chr = o.scan_chr ? std::string(o.scan_chr) : std::string();

        {
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
                 << "matrix_approximation\tfull_eigen\n"
                 << "requested_rank\t0\n"
                 << "trace_threshold\t0\n"
                 << "RANK_SELECTION\tFULL\n";
        }

        auto t0 = std::chrono::steady_clock::now();
        double ms_grm = 0, ms_null = 0, ms_scan = 0;
        DevMem raw(N * N * 8), gsub(n * n * 8);
        double S_global = 0.0;
        int64_t nv = 0;
        chk(cs_grm_build_raw(bed.h, 1, (float)o.min_maf, nullptr, 0, raw.d(), 1, &S_global, nullptr, &nv));
        ms_grm += ms_since(t0);

        LmmHandle lmm;
        chk(cs_lmm_create(ys.data(), X.data(), n, n, c, &lmm.h));

        std::vector<Result> results;
        std::vector<std::string> model_labels;
        std::vector<double> vg_all, ve_all;
        auto run_scan_with_grm = [&](const std::vector<int32_t>& snps, const std::string& label) {
            auto t1 = std::chrono::steady_clock::now();
            chk(cs_lmm_prepare_dense(lmm.h, gsub.d(), n, 1));
            double vg = 0, ve = 0;
            chk(cs_lmm_null_brent(lmm.h, &vg, &ve, nullptr));
            if (!(vg >= 0.0) || !(ve >= 0.0) || !std::isfinite(vg) || !std::isfinite(ve))
                throw RuntimeError("CosmicScanner null model: invalid variance components");
            ms_null += ms_since(t1);
            model_labels.push_back(label.empty() ? "GLOBAL" : label);
            vg_all.push_back(vg);
            ve_all.push_back(ve);
            t1 = std::chrono::steady_clock::now();
            const int64_t ns = snps.empty() ? M : (int64_t)snps.size();
            std::vector<double> freq((size_t)ns), beta((size_t)ns), se((size_t)ns), p((size_t)ns),
                l10((size_t)ns);
            chk(cs_lmm_scan_rotated(lmm.h, bed.h, keep.data(), n, snps.empty() ? nullptr : snps.data(), ns,
                                    o.min_maf, o.prec, freq.data(), nullptr, nullptr, beta.data(), se.data(),
                                    p.data(), l10.data()));
            ms_scan += ms_since(t1);
            for (int64_t i = 0; i < ns; ++i) {
                const int64_t j = snps.empty() ? i : snps[(size_t)i];
                Result r;
                r.chrom = bim[(size_t)j].chrom;
                r.snp_id = bim[(size_t)j].id;
                r.pos = bim[(size_t)j].bp;
                r.a1 = bim[(size_t)j].alt;
                r.a2 = bim[(size_t)j].ref;
                r.freq = freq[(size_t)i];
                r.n_obs = (int)n;
                r.beta = beta[(size_t)i];
                r.se = se[(size_t)i];
                r.p_value = p[(size_t)i];
                r.log10_p = l10[(size_t)i];
                results.push_back(std::move(r));
            }
        };
        auto chrom_snps = [&](const std::string& chrom) {
            std::vector<int32_t> s;
            for (int64_t j = 0; j < M; ++j)
                if (bim[(size_t)j].chrom == chrom) s.push_back((int32_t)j);
            return s;
        };

        if (!use_loco) {
            chk(cs_grm_loco_subset(raw.d(), nullptr, N, keep.data(), n, S_global, 0.0, gsub.d()));
            std::vector<int32_t> snps;
            if (!scan_chr.empty()) {
                snps = chrom_snps(scan_chr);
                if (snps.empty()) throw RuntimeError("No chromosome matched the requested additive scan filters");
            }
            run_scan_with_grm(snps, scan_chr);
        } else {
            DevMem raw_chr(N * N * 8);
            for (const std::string& chrom : labels) {
                if (!scan_chr.empty() && chrom != scan_chr) continue;
                const std::vector<int32_t> snps = chrom_snps(chrom);
                auto t1 = std::chrono::steady_clock::now();
                double S_chr = 0.0;
                chk(cs_grm_build_raw(bed.h, 1, (float)o.min_maf, snps.data(), (int64_t)snps.size(),
                                     raw_chr.d(), 1, nullptr, &S_chr, nullptr));
                chk(cs_grm_loco_subset(raw.d(), raw_chr.d(), N, keep.data(), n, S_global, S_chr, gsub.d()));
                ms_grm += ms_since(t1);
                run_scan_with_grm(snps, chrom);
            }
            if (model_labels.empty()) throw RuntimeError("No chromosome matched the requested additive scan filters");
        }

        // .assoc.txt, one pass, NaN-P rows skipped
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
                f << r.chrom << '\t' << r.snp_id << '\t' << r.pos << '\t' << r.a1 << '\t' << r.a2 << '\t'
                  << r.freq << '\t' << r.n_obs << '\t' << r.beta << '\t' << r.se << '\t' << r.p_value
                  << '\t' << r.log10_p << "\tNA\tNA\tNA\tNA\tNA\tNA\n";
                ++n_tested;
            }
        }
        {
            std::ofstream f(out + ".cosmic.sumstats");
            if (!f) throw RuntimeError("Cannot create " + out + ".cosmic.sumstats");
            const std::time_t now = std::time(nullptr);
            std::ostringstream date;
            date << std::put_time(std::gmtime(&now), "%Y-%m-%d %H:%M:%S UTC");
            f << "# CosmicScanner Summary Statistics v1.0\n# Created: " << date.str()
              << "\n# Model: LMM\n#\n"
              << "CHR\tSNP\tPOS\tA1\tA2\tBETA\tSE\tP\tN\tAF\tLOG10P\tMODEL\tZ\tWEIGHT\tINFO\tWARNING\n";
            for (const Result& r : results) {
                if (std::isnan(r.p_value)) continue;
                f << r.chrom << '\t' << r.snp_id << '\t' << r.pos << '\t' << r.a1 << '\t' << r.a2 << '\t'
                  << std::setprecision(10) << r.beta << '\t' << std::setprecision(10) << r.se << '\t'
                  << std::setprecision(10) << r.p_value << '\t' << r.n_obs << '\t'
                  << std::setprecision(6) << r.freq << '\t' << std::setprecision(6) << r.log10_p
                  << "\tLMM\t" << std::setprecision(6) << (r.se > 0 ? r.beta / r.se : 0.0) << '\t'
                  << std::setprecision(6) << (r.se > 0 ? 1.0 / (r.se * r.se) : 0.0) << '\t'
                  << std::setprecision(4) << 1.0 << "\t.\n";
            }
        }
        {
            std::ofstream meta(out + ".scan.meta.txt", std::ios::app);
            meta << "MODEL\tadditive\n"
                 << "N\t" << n << "\n"
                 << "N_SEMANTICS\tanalysis_samples_after_phenotype_and_covariate_filtering;"
                    "per-SNP_missing_genotypes_mean-imputed\n"
                 << "FREQ_SEMANTICS\tA1_allele_frequency_(not_MAF)\n"
                 << "LOCO\t" << (use_loco ? "true" : "false") << "\n"
                 << "COVARIATE_PROJECTION\ttrue\n"
                 << "VCE_CONVERGED\ttrue\n";
        }
        if (summary) {
            std::memset(summary, 0, sizeof(*summary));
            summary->n_samples = n;
            summary->n_snps_grm = nv;
            summary->sum_2pq = S_global;
            summary->loco = use_loco ? 1 : 0;
            summary->n_tested = n_tested;
            summary->n_models = (int32_t)model_labels.size();
            for (size_t i = 0; i < model_labels.size() && i < CS_GWAS_MAX_MODELS; ++i) {
                std::strncpy(summary->model_label[i], model_labels[i].c_str(), 15);
                summary->sigma2_g[i] = vg_all[i];
                summary->sigma2_e[i] = ve_all[i];
            }
            summary->ms_grm = ms_grm;
            summary->ms_null = ms_null;
            summary->ms_scan = ms_scan;
        }
    });
}
