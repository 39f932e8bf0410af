// Kernel (c): the per-SNP LMM association scan on the resident packed genotypes.
//
// For every tested SNP j the reference prepares the marker on the keep subset
// (genetics::prepareAssociationMarker, src/genetics/genotype_encoding.h:33-81: observed count,
// A1 frequency, MAF filter with n_obs >= 3, mean-impute then centre) and then forms GLS sums of
// the prepared marker g_j against projected vectors: score = g^T P y and a variance g^T P g
// (lmm_engine.cpp:1363-1525 rotated form, :1269-1361 low-rank form).  Every one of those sums is
// a dot product of g_j with a column of some n x nr matrix R, and
//     g_j^T R_c = sum_s x~_sj R~_sc - mean_j sum_s o_sj R~_sc
// (R~ = R scattered into the full sample space with keep multiplicities), i.e. the same exact
// int8-MFMA Z^T V product as the kinship operator; g^T g comes from weighted class counts.
// The FP64 epilogue assembles score, variance, beta = score/var, SE = var^-1/2 and the chi^2_1
// tail with the reference's own series / continued fraction (lmm_engine.cpp:22-64).
#include <algorithm>
#include <cmath>
#include <vector>

#include "cs_common.h"
#include "cs_geno_mfma.h"
#include "cs_ops.h"
#include "cs_stats.h"

namespace csamd {

struct ModelDev {
    int64_t n_score, quad_begin, quad_end, cov_begin, cov_end;
    double alpha;
    const double* quad_w;  // device
    const double* cov_C;   // device
};

// One thread per requested SNP whose row falls in [row_lo, row_lo + rows): prepared-marker
// statistics, the nr dot products, and (stats mode) the association statistics.  Columns below
// `split` come from PQ (rows of 2 x G16) with scales `scale`, the others from PQ2 (a packed-limb
// product of at most 16 columns: rows of 2 x 16) with scales `scale2`.
__global__ void k_scan_epilogue(const int32_t* __restrict__ snps, int64_t n_snps, int64_t row_lo,
                                int64_t rows, const int64_t* __restrict__ counts,
                                const long long* __restrict__ PQ, int64_t G16, int64_t nr,
                                const double* __restrict__ scale, int64_t split,
                                const long long* __restrict__ PQ2, const double* __restrict__ scale2,
                                double min_maf,
                                double* __restrict__ freq, int64_t* __restrict__ n_obs,
                                int32_t* __restrict__ pass, double* __restrict__ dot,
                                double* __restrict__ gtg, ModelDev md, double* __restrict__ beta,
                                double* __restrict__ se, double* __restrict__ p,
                                double* __restrict__ log10p) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n_snps) return;
    const int64_t j = snps ? snps[i] : i;   // no list: every SNP in order
    if (j < row_lo || j >= row_lo + rows) return;
    const int64_t* c = counts + j * 4;
    const int64_t obs = c[0] + c[2] + c[3];
    const double sx = 2.0 * (double)c[0] + (double)c[2];
    double f = 0.0;
    int ps = 0;
    double mean = 0.0, g2 = 0.0;
    if (obs > 0) {
        f = sx / (2.0 * (double)obs);
        const double maf = f < 1.0 - f ? f : 1.0 - f;
        ps = (obs >= 3 && maf >= min_maf) ? 1 : 0;
        mean = sx / (double)obs;
        g2 = (double)c[0] * (2.0 - mean) * (2.0 - mean) + (double)c[2] * (1.0 - mean) * (1.0 - mean) +
             (double)c[3] * mean * mean;
    }
    freq[i] = f;
    n_obs[i] = obs;
    pass[i] = ps;
    if (gtg) gtg[i] = g2;
    const long long* row = PQ + (j - row_lo) * 2 * G16;
    const long long* row2 = PQ2 ? PQ2 + (j - row_lo) * 32 : nullptr;
    auto dotc = [&](int64_t col) -> double {
        if (col < split) return ((double)row[col] - mean * (double)row[G16 + col]) * scale[col];
        const int64_t c2 = col - split;
        return ((double)row2[c2] - mean * (double)row2[16 + c2]) * scale2[c2];
    };
    if (dot) {
        for (int64_t col = 0; col < nr; ++col) dot[i * nr + col] = obs > 0 ? dotc(col) : 0.0;
    }
    if (!beta) return;
    // variance shared by all score columns
    double var = md.alpha * g2;
    for (int64_t col = md.quad_begin; col < md.quad_end; ++col) {
        const double v = dotc(col);
        var += md.quad_w[col - md.quad_begin] * v * v;
    }
    const int64_t nc = md.cov_end - md.cov_begin;
    if (nc > 0) {
        double q[32];
        for (int64_t a = 0; a < nc; ++a) q[a] = dotc(md.cov_begin + a);
        double corr = 0.0;
        for (int64_t a = 0; a < nc; ++a) {
            double t = 0.0;
            for (int64_t b2 = 0; b2 < nc; ++b2) t += md.cov_C[a * nc + b2] * q[b2];
            corr += q[a] * t;
        }
        var -= corr;
    }
    const bool tested = ps && obs > 0 && var > 1e-12 && isfinite(var);
    for (int64_t t = 0; t < md.n_score; ++t) {
        const int64_t o = i * md.n_score + t;
        if (!tested) {
            beta[o] = se[o] = p[o] = log10p[o] = __builtin_nan("");
            continue;
        }
        const double score = dotc(t);
        const double sd = sqrt(var);
        const double stat = score / sd;
        const double pv = d_chisq_prob(stat * stat, 1);
        beta[o] = score / var;
        se[o] = 1.0 / sd;
        p[o] = pv;
        log10p[o] = pv > 0.0 ? -log10(pv) : 300.0;
    }
}

// Event times of the last scan pass (any thread): class counts, genotype products, epilogues.
static struct {
    double ms_counts = 0.0, ms_products = 0.0, ms_epilogue = 0.0;
    int64_t packed_bytes = 0;
} g_scan_timing;

void run_scan(cs_bed* b, const ScanArgs& a, const cs_scan_model* model, const ScanOut& o) {
    if (a.nr < 1) throw InvalidArgument("scan: need at least one projected column");
    ensure_stats(b);
    const int64_t N = b->N, M = b->M;
    const int64_t n = a.keep ? a.n : N;
    if (a.ldr < n) throw RuntimeError("scan: dimension mismatch");
    if (a.keep)
        for (int64_t r = 0; r < n; ++r)
            if (a.keep[r] < 0 || a.keep[r] >= N) throw RuntimeError("scan: sample index out of range");
    // no SNP list: every SNP in order, with no index list built, checked or uploaded
    const int32_t* snps = a.snps;
    const int64_t ns = snps ? a.n_snps : M;
    int64_t lo = snps ? M : 0, hi = snps ? -1 : M - 1;
    for (int64_t i = 0; snps && i < ns; ++i) {
        if (snps[i] < 0 || snps[i] >= M) throw RuntimeError("scan: variant index out of range");
        lo = std::min<int64_t>(lo, snps[i]);
        hi = std::max<int64_t>(hi, snps[i]);
    }
    if (ns == 0) return;
    const int L = prec_limbs(a.prec);
    hipStream_t st;
    CS_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    try {
        // multiplicity weights and keep inverse
        std::vector<int32_t> w((size_t)b->Npad, 0);
        std::vector<int32_t> off((size_t)N + 1, 0), idx((size_t)n);
        for (int64_t r = 0; r < n; ++r) w[(size_t)(a.keep ? a.keep[r] : r)] += 1;
        for (int64_t s = 0; s < N; ++s) off[(size_t)s + 1] = off[(size_t)s] + w[(size_t)s];
        {
            std::vector<int32_t> fill(off.begin(), off.end() - 1);
            for (int64_t r = 0; r < n; ++r) idx[(size_t)fill[(size_t)(a.keep ? a.keep[r] : r)]++] = (int32_t)r;
        }
        DevBuf doff(off.size() * 4), didx(std::max<size_t>(1, idx.size()) * 4);
        CS_HIP(hipMemcpyAsync(doff.p, off.data(), off.size() * 4, hipMemcpyHostToDevice, st));
        if (!idx.empty()) CS_HIP(hipMemcpyAsync(didx.p, idx.data(), idx.size() * 4, hipMemcpyHostToDevice, st));
        DevBuf counts((size_t)M * 4 * sizeof(int64_t));
        hipEvent_t evc[2];
        CS_HIP(hipEventCreate(&evc[0]));
        CS_HIP(hipEventCreate(&evc[1]));
        CS_HIP(hipEventRecord(evc[0], st));
        weighted_counts(b, w.data(), (int64_t)w.size(), counts.as<int64_t>(), st);
        CS_HIP(hipEventRecord(evc[1], st));
        // R into device (compact n x nr) then scatter to the full sample space
        DevBuf dR;
        const double* Rd = a.R;
        int64_t ldr = a.ldr;
        if (!a.device_R) {
            dR.alloc((size_t)(n * a.nr) * sizeof(double));
            CS_HIP(hipMemcpy2DAsync(dR.p, (size_t)n * sizeof(double), a.R, (size_t)a.ldr * sizeof(double),
                                    (size_t)n * sizeof(double), (size_t)a.nr, hipMemcpyHostToDevice, st));
            Rd = dR.as<double>();
            ldr = n;
        }
        // A last partial group of r columns with r L <= 16 runs as a packed-limb product of its own
        // (2 MFMAs per genotype fragment instead of 6): the columns' scales and integer products
        // do not depend on their neighbours, so the results are the same bit for bit.  Each part
        // is scattered to the full sample space as its own row-major [Npad][columns] matrix.
        const int64_t rr = a.nr % 16;
        const bool split = a.nr > 16 && rr > 0 && tiling(rr, L).packed;   // rr L <= 16, LDS kernels on
        const int64_t k1 = split ? a.nr - rr : a.nr;
        DevBuf Rt((size_t)(b->Npad * k1) * sizeof(double)), Rt2;
        k_scatter<<<grid_for(b->Npad * k1), 256, 0, st>>>(Rd, ldr, k1, doff.as<int32_t>(), didx.as<int32_t>(), N,
                                                          b->Npad, Rt.as<double>());
        if (split) {
            Rt2.alloc((size_t)(b->Npad * rr) * sizeof(double));
            k_scatter<<<grid_for(b->Npad * rr), 256, 0, st>>>(Rd + k1 * ldr, ldr, rr, doff.as<int32_t>(),
                                                              didx.as<int32_t>(), N, b->Npad, Rt2.as<double>());
        }
        CS_HIP(hipGetLastError());
        const Tiling tl = tiling(k1, L);
        DevBuf mx, scl, B((size_t)(b->Npad * tl.NT * 16));
        {
            const double* Ws[1] = {Rt.as<double>()};
            int8_t* Bs[1] = {B.as<int8_t>()};
            quantize_cols(Ws, 1, b->Npad, k1, tl, st, mx, scl, Bs);
        }
        const Tiling tl2 = tiling(split ? rr : 1, L);
        DevBuf mx2, scl2, B2;
        if (split) {
            B2.alloc((size_t)(b->Npad * tl2.NT * 16));
            const double* Ws[1] = {Rt2.as<double>()};
            int8_t* Bs[1] = {B2.as<int8_t>()};
            quantize_cols(Ws, 1, b->Npad, rr, tl2, st, mx2, scl2, Bs);
        }
        // model to device
        ModelDev md{};
        DevBuf dqw, dC;
        if (model) {
            md.n_score = model->n_score;
            md.alpha = model->alpha;
            md.quad_begin = model->quad_begin;
            md.quad_end = model->quad_end;
            md.cov_begin = model->cov_begin;
            md.cov_end = model->cov_end;
            if (md.n_score < 1 || md.n_score > a.nr || md.quad_begin < 0 || md.quad_end > a.nr ||
                md.quad_begin > md.quad_end || md.cov_begin < 0 || md.cov_end > a.nr ||
                md.cov_begin > md.cov_end || md.cov_end - md.cov_begin > 32)
                throw InvalidArgument("scan: bad model column ranges");
            const int64_t nq = md.quad_end - md.quad_begin, nc = md.cov_end - md.cov_begin;
            dqw.alloc((size_t)std::max<int64_t>(1, nq) * 8);
            dC.alloc((size_t)std::max<int64_t>(1, nc * nc) * 8);
            if (nq) CS_HIP(hipMemcpyAsync(dqw.p, model->quad_w, (size_t)nq * 8, hipMemcpyHostToDevice, st));
            if (nc) CS_HIP(hipMemcpyAsync(dC.p, model->cov_C, (size_t)(nc * nc) * 8, hipMemcpyHostToDevice, st));
            md.quad_w = dqw.as<double>();
            md.cov_C = dC.as<double>();
        }
        // outputs on the device: the caller's arrays when they are device memory, else staging
        DevBuf dsn, sfreq, snobs, spass, sdot, sgtg, sbeta, sse, sp, sl;
        if (snps) {
            dsn.alloc((size_t)ns * 4);
            CS_HIP(hipMemcpyAsync(dsn.p, snps, (size_t)ns * 4, hipMemcpyHostToDevice, st));
        }
        const int64_t nsc = model ? model->n_score : 0;
        auto dev_out = [&](void* user, DevBuf& stage, size_t bytes, bool wanted) -> void* {
            if (!wanted) return nullptr;
            if (o.device && user) return user;
            stage.alloc(bytes);
            return stage.p;
        };
        double* pfreq = (double*)dev_out(o.freq, sfreq, (size_t)ns * 8, true);
        int64_t* pnobs = (int64_t*)dev_out(o.n_obs, snobs, (size_t)ns * 8, true);
        int32_t* ppass = (int32_t*)dev_out(o.pass, spass, (size_t)ns * 4, true);
        double* pdot = (double*)dev_out(o.dot, sdot, (size_t)(ns * a.nr) * 8, o.dot != nullptr);
        double* pgtg = (double*)dev_out(o.gtg, sgtg, (size_t)ns * 8, o.gtg != nullptr);
        double* pbeta = (double*)dev_out(o.beta, sbeta, (size_t)(ns * nsc) * 8, model != nullptr);
        double* pse = (double*)dev_out(o.se, sse, (size_t)(ns * nsc) * 8, model != nullptr);
        double* pp = (double*)dev_out(o.p, sp, (size_t)(ns * nsc) * 8, model != nullptr);
        double* pl = (double*)dev_out(o.log10p, sl, (size_t)(ns * nsc) * 8, model != nullptr);
        // SNP-row chunks of the Z^T R product
        const int64_t row_lo0 = lo / kPad * kPad;
        const int64_t row_hi = round_up(hi + 1, kPad);
        const int64_t G16 = 16 * (int64_t)tl.G;
        int64_t chunk = std::max<int64_t>(kPad, ((int64_t)(256ll << 20) / (16ll * G16)) / kPad * kPad);
        DevBuf PQ((size_t)std::min<int64_t>(chunk, row_hi - row_lo0) * 2 * G16 * sizeof(long long)), PQ2;
        if (split) PQ2.alloc((size_t)std::min<int64_t>(chunk, row_hi - row_lo0) * 2 * 16 * sizeof(long long));
        std::vector<hipEvent_t> ev;
        for (int64_t r0 = row_lo0; r0 < row_hi; r0 += chunk) {
            for (int e = 0; e < 3; ++e) {
                ev.push_back(nullptr);
                CS_HIP(hipEventCreate(&ev.back()));
            }
            CS_HIP(hipEventRecord(ev[ev.size() - 3], st));
            const int64_t rows = std::min<int64_t>(chunk, row_hi - r0);
            GenoArgs ga{};
            ga.A = b->snp.as<uint8_t>() + (((r0 / 64) * b->KBs) << 12);
            ga.KB = b->KBs;
            ga.rows = std::min<int64_t>(rows, M - r0);
            ga.kb_total = ga.kb_per_split = b->KBs;
            ga.G = tl.G;
            ga.Bx = ga.Bo = B.as<int8_t>();
            ga.k = k1;
            ga.pq = PQ.as<long long>();
            launch_geno(L, false, EPI_PQ, dim3((unsigned)(rows / 256), 1, (unsigned)tl.G), st, ga, false, tl.packed);
            if (split) {
                GenoArgs g2 = ga;
                g2.G = tl2.G;
                g2.Bx = g2.Bo = B2.as<int8_t>();
                g2.k = rr;
                g2.pq = PQ2.as<long long>();
                launch_geno(L, false, EPI_PQ, dim3((unsigned)(rows / 256), 1, 1), st, g2, false, true);
            }
            CS_HIP(hipEventRecord(ev[ev.size() - 2], st));
            k_scan_epilogue<<<grid_for(ns), 256, 0, st>>>(
                dsn.as<int32_t>(), ns, r0, rows, counts.as<int64_t>(), PQ.as<long long>(), G16, a.nr,
                scl.as<double>(), k1, split ? PQ2.as<long long>() : nullptr, split ? scl2.as<double>() : nullptr,
                a.min_maf, pfreq, pnobs, ppass, pdot, pgtg, md, pbeta, pse, pp, pl);
            CS_HIP(hipGetLastError());
            CS_HIP(hipEventRecord(ev[ev.size() - 1], st));
        }
        // staged outputs to the caller (host copies, or device copies when the caller passed
        // device memory only for some)
        auto d2h = [&](void* dst, const DevBuf& stage, size_t bytes) {
            if (dst && stage.p)
                CS_HIP(hipMemcpyAsync(dst, stage.p, bytes, o.device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                      st));
        };
        d2h(o.freq, sfreq, (size_t)ns * 8);
        d2h(o.n_obs, snobs, (size_t)ns * 8);
        d2h(o.pass, spass, (size_t)ns * 4);
        d2h(o.dot, sdot, (size_t)(ns * a.nr) * 8);
        d2h(o.gtg, sgtg, (size_t)ns * 8);
        if (model) {
            d2h(o.beta, sbeta, (size_t)(ns * nsc) * 8);
            d2h(o.se, sse, (size_t)(ns * nsc) * 8);
            d2h(o.p, sp, (size_t)(ns * nsc) * 8);
            d2h(o.log10p, sl, (size_t)(ns * nsc) * 8);
        }
        CS_HIP(hipStreamSynchronize(st));
        float ms = 0.f;
        double mp = 0.0, me = 0.0;
        for (size_t e = 0; e + 2 < ev.size(); e += 3) {
            if (hipEventElapsedTime(&ms, ev[e], ev[e + 1]) == hipSuccess) mp += ms;
            if (hipEventElapsedTime(&ms, ev[e + 1], ev[e + 2]) == hipSuccess) me += ms;
        }
        g_scan_timing.ms_counts = hipEventElapsedTime(&ms, evc[0], evc[1]) == hipSuccess ? ms : 0.0;
        g_scan_timing.ms_products = mp;
        g_scan_timing.ms_epilogue = me;
        g_scan_timing.packed_bytes = (row_hi - row_lo0) * b->KBs * 64;
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        (void)hipEventDestroy(evc[0]);
        (void)hipEventDestroy(evc[1]);
    } catch (...) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
        throw;
    }
    CS_HIP(hipStreamDestroy(st));
}

// Prepared markers (genotype_encoding.h:61-79) as dense columns over the keep list: x - mean for
// observed genotypes, 0 for missing (mean imputation then centring); columns of SNPs that fail
// the filter are zero.
__global__ void k_markers(const uint8_t* __restrict__ snp, int64_t KBs, const int32_t* __restrict__ keep,
                          int64_t n, const int32_t* __restrict__ snps, int64_t ns,
                          const double* __restrict__ mean, const int32_t* __restrict__ pass,
                          double* __restrict__ G, int64_t ldg) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * ns;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = e / n, r = e - c * n;
        double v = 0.0;
        if (pass[c]) {
            const int64_t s = keep ? keep[r] : r;
            const int code = (snp[tiled_offset(snps[c], s >> 2, KBs)] >> (2 * (s & 3))) & 3;
            if (code != 1) v = (code == 0 ? 2.0 : code == 2 ? 1.0 : 0.0) - mean[c];
        }
        G[c * ldg + r] = v;
    }
}

void markers_device(cs_bed* b, const int32_t* keep, int64_t n_keep, const int32_t* snps, int64_t ns,
                    double min_maf, double* dG, int64_t ldg, int32_t* pass_out, double* freq_out,
                    hipStream_t st) {
    ensure_stats(b);
    const int64_t N = b->N, M = b->M;
    const int64_t n = keep ? n_keep : N;
    if (ldg < n) throw RuntimeError("markers: dimension mismatch");
    for (int64_t r = 0; keep && r < n; ++r)
        if (keep[r] < 0 || keep[r] >= N) throw RuntimeError("markers: sample index out of range");
    for (int64_t i = 0; i < ns; ++i)
        if (snps[i] < 0 || snps[i] >= M) throw RuntimeError("markers: variant index out of range");
    if (ns == 0) return;
    std::vector<int32_t> w((size_t)b->Npad, 0);
    for (int64_t r = 0; r < n; ++r) w[(size_t)(keep ? keep[r] : r)] += 1;
    DevBuf dk((size_t)std::max<int64_t>(1, keep ? n : 1) * 4), dsn((size_t)ns * 4), dmean((size_t)ns * 8),
        dpass((size_t)ns * 4);
    // class counts of the selected SNPs only (their 64-SNP row tiles)
    std::vector<int64_t> cnt((size_t)ns * 4);
    weighted_counts_of(b, w.data(), (int64_t)w.size(), snps, ns, cnt.data(), st);
    std::vector<double> mean((size_t)ns, 0.0);
    std::vector<int32_t> ps((size_t)ns, 0);
    for (int64_t i = 0; i < ns; ++i) {
        const int64_t* c = &cnt[(size_t)i * 4];
        const int64_t obs = c[0] + c[2] + c[3];
        const double sx = 2.0 * (double)c[0] + (double)c[2];
        double f = 0.0;
        if (obs > 0) {
            f = sx / (2.0 * (double)obs);
            const double maf = f < 1.0 - f ? f : 1.0 - f;
            ps[(size_t)i] = (obs >= 3 && maf >= min_maf) ? 1 : 0;
            mean[(size_t)i] = sx / (double)obs;
        }
        if (freq_out) freq_out[i] = f;
        if (pass_out) pass_out[i] = ps[(size_t)i];
    }
    if (keep) CS_HIP(hipMemcpyAsync(dk.p, keep, (size_t)n * 4, hipMemcpyHostToDevice, st));
    CS_HIP(hipMemcpyAsync(dsn.p, snps, (size_t)ns * 4, hipMemcpyHostToDevice, st));
    CS_HIP(hipMemcpyAsync(dmean.p, mean.data(), (size_t)ns * 8, hipMemcpyHostToDevice, st));
    CS_HIP(hipMemcpyAsync(dpass.p, ps.data(), (size_t)ns * 4, hipMemcpyHostToDevice, st));
    k_markers<<<grid_for(n * ns), 256, 0, st>>>(b->snp.as<uint8_t>(), b->KBs, keep ? dk.as<int32_t>() : nullptr,
                                                n, dsn.as<int32_t>(), ns, dmean.as<double>(),
                                                dpass.as<int32_t>(), dG, ldg);
    CS_HIP(hipGetLastError());
    CS_HIP(hipStreamSynchronize(st));
}

}  // namespace csamd

using namespace csamd;

extern "C" {

int cs_markers_device(cs_bed* bed, const int32_t* keep, int64_t n_keep, const int32_t* snps,
                      int64_t n_snps, double min_maf, double* d_G, int64_t ldg, int32_t* pass,
                      double* freq) {
    return guarded([&] {
        if (!bed || (n_snps > 0 && (!snps || !d_G))) throw InvalidArgument("cs_markers_device: null argument");
        hipStream_t st;
        CS_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        try {
            markers_device(bed, (keep && n_keep > 0) ? keep : nullptr, n_keep, snps, n_snps, min_maf, d_G, ldg,
                           pass, freq, st);
        } catch (...) {
            (void)hipStreamDestroy(st);
            throw;
        }
        CS_HIP(hipStreamDestroy(st));
    });
}

int cs_scan_sums(cs_bed* bed, const int32_t* keep, int64_t n_keep, const double* R, int64_t ldr,
                 int64_t nr, double min_maf, const int32_t* snps, int64_t n_snps, int prec,
                 double* freq, int64_t* n_obs, int32_t* pass, double* dot, double* gtg) {
    return guarded([&] {
        if (!bed || !R) throw InvalidArgument("cs_scan_sums: null argument");
        ScanArgs a;
        a.keep = (keep && n_keep > 0) ? keep : nullptr;
        a.n = a.keep ? n_keep : bed->N;
        a.R = R; a.ldr = ldr; a.nr = nr; a.min_maf = min_maf;
        a.snps = (snps && n_snps > 0) ? snps : nullptr;
        a.n_snps = a.snps ? n_snps : bed->M;
        a.prec = prec;
        ScanOut o;
        o.freq = freq; o.n_obs = n_obs; o.pass = pass; o.dot = dot; o.gtg = gtg;
        run_scan(bed, a, nullptr, o);
    });
}

int cs_lmm_scan(cs_bed* bed, const int32_t* keep, int64_t n_keep, const double* R, int64_t ldr,
                int64_t nr, double min_maf, const int32_t* snps, int64_t n_snps, int prec,
                const cs_scan_model* model, double* freq, int64_t* n_obs, int32_t* pass,
                double* beta, double* se, double* p, double* log10p) {
    return guarded([&] {
        if (!bed || !R || !model) throw InvalidArgument("cs_lmm_scan: null argument");
        ScanArgs a;
        a.keep = (keep && n_keep > 0) ? keep : nullptr;
        a.n = a.keep ? n_keep : bed->N;
        a.R = R; a.ldr = ldr; a.nr = nr; a.min_maf = min_maf;
        a.snps = (snps && n_snps > 0) ? snps : nullptr;
        a.n_snps = a.snps ? n_snps : bed->M;
        a.prec = prec;
        ScanOut o;
        o.freq = freq; o.n_obs = n_obs; o.pass = pass;
        o.beta = beta; o.se = se; o.p = p; o.log10p = log10p;
        run_scan(bed, a, model, o);
    });
}

int cs_lmm_scan_device(cs_bed* bed, const int32_t* keep, int64_t n_keep, const double* dR, int64_t ldr,
                       int64_t nr, double min_maf, const int32_t* snps, int64_t n_snps, int prec,
                       const cs_scan_model* model, double* d_freq, int64_t* d_n_obs, int32_t* d_pass,
                       double* d_beta, double* d_se, double* d_p, double* d_log10p) {
    return guarded([&] {
        if (!bed || !dR || !model) throw InvalidArgument("cs_lmm_scan_device: null argument");
        ScanArgs a;
        a.keep = (keep && n_keep > 0) ? keep : nullptr;
        a.n = a.keep ? n_keep : bed->N;
        a.R = dR; a.ldr = ldr; a.nr = nr; a.min_maf = min_maf;
        a.snps = (snps && n_snps > 0) ? snps : nullptr;
        a.n_snps = a.snps ? n_snps : bed->M;
        a.prec = prec;
        a.device_R = true;
        ScanOut o;
        o.freq = d_freq; o.n_obs = d_n_obs; o.pass = d_pass;
        o.beta = d_beta; o.se = d_se; o.p = d_p; o.log10p = d_log10p;
        o.device = true;
        run_scan(bed, a, model, o);
    });
}

int cs_scan_last_timing(double* ms_counts, double* ms_products, double* ms_epilogue, int64_t* packed_bytes) {
    return guarded([&] {
        if (ms_counts) *ms_counts = g_scan_timing.ms_counts;
        if (ms_products) *ms_products = g_scan_timing.ms_products;
        if (ms_epilogue) *ms_epilogue = g_scan_timing.ms_epilogue;
        if (packed_bytes) *packed_bytes = g_scan_timing.packed_bytes;
    });
}

}  // extern "C"
