"""GPU parity of the dense LMM path against the CPU oracle, through the C-ABI.

GwasScanner::runLocoScan's default PLINK path (src/gwas_scanner.cpp:74-525): exact GRM and
per-chromosome raw GRMs, the LOCO formula (:481-495), EigenVce::prepare + solveDiagonal
(libCosmicCore/src/eigen_vce.cpp:19-113, vce.cpp:3073-3171), the rotated scan
(src/lmm_engine.cpp:1363-1525), the .assoc.txt / .scan.meta.txt writers.

Tolerances.
* LOCO GRM from the exact raw GRMs: bit-exact (same exact inputs, same FP64 operation order).
* Eigenvalues (rocSOLVER vs LAPACK): 1e-9 relative to the largest; variance components: 1e-6.
* Scan statistics (int8 MFMA products of 23-bit fixed-point projections, FP64 epilogue):
  |d log10 P| <= 1e-4, |d beta| <= 1e-5 SE, SE relative 1e-5.  Printed .assoc.txt values carry the
  reference's 6 significant digits on top (5e-6 relative); FREQ prints identically.
* With the standardised GRM (--make-grm-alg 0: FP16-split MFMA products, FP32 accumulation, GRM
  entries within ~8e-6 of FP64) the bounds are the path's stated ones: variance components 1e-4
  relative, |d log10 P| <= 1e-3, |d beta| <= 1e-3 SE, SE relative 1e-4.
"""
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def csa():
    import cosmicscanner_amd as csa
    return csa


def write_plink(prefix, body, N, M, chroms):
    with open(prefix + ".bed", "wb") as f:
        f.write(bytes([0x6C, 0x1B, 0x01]))
        f.write(np.asarray(body, dtype=np.uint8).tobytes())
    with open(prefix + ".fam", "w") as f:
        for i in range(N):
            f.write(f"F{i:04d} I{i:04d} 0 0 1 -9\n")
    with open(prefix + ".bim", "w") as f:
        for j in range(M):
            f.write(f"{chroms[j]}\trs{j:06d}\t0\t{1000 + 10 * j}\tA\tC\n")


@pytest.fixture(scope="module")
def synth_gwas(tmp_path_factory):
    """n = 500, m = 2400 over chromosomes 1, 2, 10 (lexicographic order 1, 10, 2), 1 % missing,
    intercept + 2 covariates (some missing), 20 causal SNPs, h2 ~ 0.5."""
    N, M = 500, 2400
    body = oracle.gen_hwe_bed(20261203, N, M, 0.05, 0.5, 0.01)
    chroms = ["1"] * 900 + ["2"] * 700 + ["10"] * 800
    rng = np.random.default_rng(4)
    D = oracle.decode_numpy(body, N, M)
    Dm = np.where(np.isnan(D), np.nanmean(D, axis=0), D)
    causal = rng.choice(M, 20, replace=False)
    g = (Dm[:, causal] - Dm[:, causal].mean(0)) @ rng.normal(0, np.sqrt(0.5 / 20), 20)
    cov = rng.standard_normal((N, 2))
    y = 1.0 + cov @ np.array([0.3, -0.2]) + g + rng.normal(0, np.sqrt(0.5), N)
    y[rng.choice(N, 7, replace=False)] = np.nan
    cov[rng.choice(N, 5, replace=False), 1] = np.nan
    prefix = str(tmp_path_factory.mktemp("gwas") / "syn")
    write_plink(prefix, body, N, M, chroms)
    return dict(prefix=prefix, body=body, N=N, M=M, chroms=chroms, y=y, cov=cov)


def read_assoc(path):
    lines = open(path).read().splitlines()
    header = lines[0].split("\t")
    return header, [ln.split("\t") for ln in lines[1:]]


def compare_rows(bim_ids, rows_gpu, rows_orc):
    assert [r[1] for r in rows_gpu] == [bim_ids[r[0]] for r in rows_orc]
    for rg, ro in zip(rows_gpu, rows_orc):
        _, freq, beta, se, p, l10 = ro
        assert rg[5] == f"{freq:.6g}"
        b, s, lp = float(rg[7]), float(rg[8]), float(rg[10])
        assert abs(b - beta) <= 1e-5 * se + 2e-5 * abs(beta)
        assert abs(s - se) <= 2e-5 * se
        assert abs(lp - l10) <= 1e-4 + 1e-5 * abs(l10)
        assert rg[11:] == ["NA"] * 6


def test_loco_grm_bitwise(csa, synth_gwas):
    d = synth_gwas
    N, M, body, chroms = d["N"], d["M"], d["body"], d["chroms"]
    r = csa.PlinkReader.from_body(body, N, M)
    keep = np.nonzero(~np.isnan(d["y"]))[0].astype(np.int32)
    n = keep.size
    raw = csa.DeviceArray(N, N)
    raw_c = csa.DeviceArray(N, N)
    out = csa.DeviceArray(n, n)
    S, _, _ = csa.grm_build_raw(r, raw)
    on = [j for j in range(M) if chroms[j] == "10"]
    _, S_c, _ = csa.grm_build_raw(r, raw_c, snps=on)
    assert S_c == oracle.sequential_sum_2pq(body, N, M, 0.01, on)
    csa.grm_loco_subset(raw, raw_c, N, keep, S, S_c, out)
    raw_o, _, S_o, _ = oracle.grm_exact_centred(body, N, M)
    raw_co = oracle.grm_exact_centred(oracle.select_snps(body, N, on), N, len(on))[0]
    assert S == S_o
    S_l = S - S_c
    ref = (raw_o / S)[np.ix_(keep, keep)] * (S / S_l) - raw_co[np.ix_(keep, keep)] / S_l
    assert np.array_equal(out.download(), ref)
    # no LOCO: plain subset of raw / S
    csa.grm_loco_subset(raw, None, N, keep, S, 0.0, out)
    assert np.array_equal(out.download(), (raw_o / S)[np.ix_(keep, keep)])


def test_dense_lmm_components(csa, synth_gwas):
    d = synth_gwas
    N, M, body = d["N"], d["M"], d["body"]
    keep = np.nonzero(~np.isnan(d["y"]) & ~np.isnan(d["cov"]).any(1))[0].astype(np.int32)
    ys = d["y"][keep]
    X = np.column_stack([np.ones(keep.size), d["cov"][keep]]).astype(np.float32)
    raw, _, S, _ = oracle.grm_exact_centred(body, N, M)
    G = (raw / S)[np.ix_(keep, keep)]
    lmm = csa.DenseLmm(ys, X)
    for dev in (False, True):
        lmm.prepare(csa.DeviceArray.from_host(G) if dev else G)
        D, Uty = lmm.eigen()
        Do, Wo, Utyo = oracle.eigen_vce_prepare(G, ys, X.astype(np.float64))
        assert np.max(np.abs(D - Do)) <= 1e-9 * np.max(Do)
        # Uty is basis dependent; its energy per (distinct) eigenvalue is not
        assert abs(np.sum(Uty ** 2) - np.sum(Utyo ** 2)) <= 1e-9 * np.sum(Utyo ** 2)
        lmm.runNullModel()
        vg, ve, delta = oracle.solve_diagonal(Do, Utyo)
        assert abs(lmm.sigma2_g - vg) <= 1e-6 * vg and abs(lmm.sigma2_e - ve) <= 1e-6 * ve
        r = csa.PlinkReader.from_body(body, N, M)
        res = lmm.scan(r, keep)
        Vinv = 1.0 / (vg * Do + ve)
        R = np.column_stack([Wo @ (Utyo * Vinv), Wo])
        freq, _, ps, dot, _ = oracle.scan_sums(body, N, M, keep, R)
        st = oracle._stats_from(dot[:, 0], (dot[:, 1:] ** 2) @ Vinv)
        t = ps & st["tested"]
        assert np.array_equal(np.isnan(res["p"]), ~t)
        assert np.array_equal(res["freq"], freq)
        assert np.max(np.abs(res["log10p"][t] - st["log10p"][t])) <= 1e-4
        assert np.max(np.abs(res["beta"][t] - st["beta"][t]) / st["se"][t]) <= 1e-5
        assert np.max(np.abs(res["se"][t] / st["se"][t] - 1)) <= 1e-5


def test_gwas_end_to_end_synthetic(csa, synth_gwas, tmp_path):
    d = synth_gwas
    out = str(tmp_path / "run")
    sm = csa.gwas_dense_run(d["prefix"], d["y"], out, cov=d["cov"])
    models, rows = oracle.gwas_dense(d["body"], d["N"], d["M"], d["chroms"], d["y"], d["cov"])
    assert sm["loco"] and [m[0] for m in sm["models"]] == ["1", "10", "2"]
    for (lg, vg, ve), (lo, vgo, veo) in zip(sm["models"], models):
        assert lg == lo and abs(vg - vgo) <= 1e-6 * vgo and abs(ve - veo) <= 1e-6 * veo
    header, rg = read_assoc(out + ".assoc.txt")
    assert header == ("CHR SNP BP A1 A2 FREQ N BETA_A SE_A P_A LOG10P_A BETA_D SE_D P_D "
                      "LOG10P_D P_JOINT LOG10P_JOINT").split()
    assert sm["n_tested"] == len(rg) == len(rows)
    assert all(r[6] == str(sm["n_samples"]) for r in rg)
    compare_rows([f"rs{j:06d}" for j in range(d["M"])], rg, rows)
    meta = dict(ln.split("\t", 1) for ln in open(out + ".scan.meta.txt").read().splitlines())
    assert meta["LOCO"] == "true" and meta["N"] == str(sm["n_samples"])
    assert meta["COVARIATE_PROJECTION"] == "true" and meta["matrix_approximation"] == "full_eigen"
    check_sumstats(out + ".cosmic.sumstats", rg, rows, [f"rs{j:06d}" for j in range(d["M"])], sm["n_samples"])


def check_sumstats(path, assoc_rows, oracle_rows, bim_ids, n):
    """.cosmic.sumstats (gwas_output.cpp:187-239): the v1.0 comment header, the canonical columns,
    one row per .assoc.txt row in the same order with BETA / SE / P at 10 significant digits
    (against the oracle: |d log10 P| <= 1e-4, |d beta| <= 1e-5 SE), AF and LOG10P at 6, MODEL LMM,
    Z = BETA / SE, WEIGHT = 1 / SE^2 at 6 digits, INFO 1 and WARNING '.'."""
    lines = open(path).read().splitlines()
    assert lines[0] == "# CosmicScanner Summary Statistics v1.0"
    assert lines[1].startswith("# Created: ") and lines[1].endswith(" UTC")
    assert lines[2] == "# Model: LMM" and lines[3] == "#"
    assert lines[4].split("\t") == ("CHR SNP POS A1 A2 BETA SE P N AF LOG10P MODEL Z WEIGHT INFO "
                                    "WARNING").split()
    body = [ln.split("\t") for ln in lines[5:]]
    assert len(body) == len(assoc_rows) == len(oracle_rows)
    for sr, ar, orow in zip(body, assoc_rows, oracle_rows):
        j, freq, beta, se, p, l10 = orow
        assert sr[:5] == ar[:5] and sr[1] == bim_ids[j]
        b, s, pv = float(sr[5]), float(sr[6]), float(sr[7])
        assert sr[8] == str(n) and sr[9] == f"{freq:.6g}" and sr[11] == "LMM"
        assert abs(b - beta) <= 1e-5 * se and abs(s / se - 1) <= 1e-5
        lp = -np.log10(pv) if pv > 0 else 300.0
        assert abs(lp - l10) <= 1e-4 and abs(float(sr[10]) - l10) <= 1e-4 + 5e-6 * abs(l10)
        # the .assoc.txt values are the same numbers at 6 significant digits
        assert ar[7] == f"{b:.6g}" and ar[8] == f"{s:.6g}" and ar[9] == f"{pv:.6g}"
        assert abs(float(sr[12]) - b / s) <= 1e-5 * abs(b / s) + 1e-12
        assert abs(float(sr[13]) * s * s - 1.0) <= 1e-5
        assert sr[14] == "1" and sr[15] == "."


def test_gwas_options(csa, synth_gwas, tmp_path):
    d = synth_gwas
    out = str(tmp_path / "nl")
    sm = csa.gwas_dense_run(d["prefix"], d["y"], out, cov=d["cov"], no_loco=True)
    models, rows = oracle.gwas_dense(d["body"], d["N"], d["M"], d["chroms"], d["y"], d["cov"],
                                     no_loco=True)
    assert not sm["loco"] and len(sm["models"]) == 1 and sm["models"][0][0] == "GLOBAL"
    compare_rows([f"rs{j:06d}" for j in range(d["M"])], read_assoc(out + ".assoc.txt")[1], rows)
    out2 = str(tmp_path / "c2")
    csa.gwas_dense_run(d["prefix"], d["y"], out2, cov=d["cov"], scan_chr="2")
    _, rows2 = oracle.gwas_dense(d["body"], d["N"], d["M"], d["chroms"], d["y"], d["cov"], scan_chr="2")
    rg2 = read_assoc(out2 + ".assoc.txt")[1]
    assert {r[0] for r in rg2} == {"2"}
    compare_rows([f"rs{j:06d}" for j in range(d["M"])], rg2, rows2)
    with pytest.raises(RuntimeError, match="Too few valid samples"):
        csa.gwas_dense_run(d["prefix"], np.full(d["N"], np.nan), str(tmp_path / "x"))


def test_gwas_standardised_grm(csa, synth_gwas, tmp_path):
    d = synth_gwas
    out = str(tmp_path / "alg0")
    sm = csa.gwas_dense_run(d["prefix"], d["y"], out, cov=d["cov"], grm_alg=0)
    models, rows = oracle.gwas_dense(d["body"], d["N"], d["M"], d["chroms"], d["y"], d["cov"],
                                     grm_alg=0)
    assert sm["loco"] and [m[0] for m in sm["models"]] == ["1", "10", "2"]
    for (lg, vg, ve), (lo, vgo, veo) in zip(sm["models"], models):
        assert lg == lo and abs(vg - vgo) <= 1e-4 * vgo and abs(ve - veo) <= 1e-4 * veo
    _, rg = read_assoc(out + ".assoc.txt")
    ids = [f"rs{j:06d}" for j in range(d["M"])]
    assert [r[1] for r in rg] == [ids[r[0]] for r in rows]
    worst = 0.0
    for r, (_, freq, beta, se, p, l10) in zip(rg, rows):
        assert r[5] == f"{freq:.6g}"
        b, s_, lp = float(r[7]), float(r[8]), float(r[10])
        assert abs(b - beta) <= 1e-3 * se and abs(s_ - se) <= 1e-4 * se
        assert abs(lp - l10) <= 1e-3
        worst = max(worst, abs(lp - l10))
    print(f"  standardised-GRM scan: max |d log10 P| = {worst:.2e}", flush=True)


@pytest.mark.parametrize("name,pheno", [("synth", "ref_synth/quantitative.pheno"),
                                        ("gemma", "ref_gemma_accuracy/gemma_accuracy.pheno")])
def test_gwas_reference_fixtures(csa, ref_datasets, tmp_path, name, pheno):
    body, N, M, bim, fam = ref_datasets[name]
    g = os.path.join(os.path.dirname(__file__), "golden")
    prefix = {"synth": "ref_synth/synth", "gemma": "ref_gemma_accuracy/gemma_accuracy"}[name]
    y = csa.read_pheno(os.path.join(g, pheno), fam)
    out = str(tmp_path / name)
    csa.gwas_dense_run(os.path.join(g, prefix), y, out)
    _, rows = oracle.gwas_dense(body, N, M, [b[0] for b in bim], y)
    _, rg = read_assoc(out + ".assoc.txt")
    compare_rows([b[1] for b in bim], rg, rows)
    if name == "synth":
        assert len(rg) == 180          # README.md:63
    else:
        rank = [r[1] for r in sorted(rg, key=lambda r: float(r[9]))]
        assert rank.index("rs000013") < 5 and rank.index("rs000212") < 5


@pytest.mark.parametrize("max_rank,threshold", [(64, 0.95), (64, 0.05)])
def test_gwas_matrix_free_lowrank(csa, synth_gwas, tmp_path, max_rank, threshold):
    d = synth_gwas
    out = str(tmp_path / "mf")
    sm = csa.gwas_matrix_free_run(d["prefix"], d["y"], out, cov=d["cov"], max_rank=max_rank,
                                  trace_threshold=threshold)
    models, rows = oracle.gwas_matrix_free(d["body"], d["N"], d["M"], d["chroms"], d["y"], d["cov"],
                                           max_rank=max_rank, trace_threshold=threshold)
    assert [m[0] for m in sm["models"]] == ["1", "10", "2"]
    for (lg, vg, ve), (lo, vgo, veo) in zip(sm["models"], models):
        assert lg == lo and abs(vg / vgo - 1) <= 1e-4 and abs(ve / veo - 1) <= 1e-4
    _, rg = read_assoc(out + ".assoc.txt")
    assert len(rg) == len(rows)
    ids = [f"rs{j:06d}" for j in range(d["M"])]
    assert [r[1] for r in rg] == [ids[r[0]] for r in rows]
    for g, o in zip(rg, rows):
        assert g[5] == f"{o[1]:.6g}"
        assert abs(float(g[10]) - o[5]) <= 1e-3
        assert abs(float(g[7]) - o[2]) <= 1e-3 * o[3] + 2e-5 * abs(o[2])
    meta = open(out + ".low_rank.meta.tsv").read().splitlines()
    assert meta[0].split("\t")[0] == "SCOPE" and [m.split("\t")[0] for m in meta[1:]] == ["1", "10", "2"]
    fb = {m.split("\t")[-1] for m in meta[1:]}
    assert fb == ({"trace_threshold"} if threshold > 0.5 else {"none"})


def test_gwas_matrix_free_pcg(csa, synth_gwas, tmp_path):
    d = synth_gwas
    out = str(tmp_path / "pcg")
    sm = csa.gwas_matrix_free_run(d["prefix"], d["y"], out, cov=d["cov"], mode="pcg", sigma2_g=0.4, sigma2_e=0.6,
                                  pcg_tol=1e-8, n_calibration=24)
    assert sm["loco"] and len(sm["models"]) == 3
    _, rg = read_assoc(out + ".assoc.txt")
    # each chromosome against the oracle's exact-solve restatement of the projected scan
    ok = ~np.isnan(d["y"]) & ~np.isnan(d["cov"]).any(1)
    keep = np.nonzero(ok)[0].astype(np.int32)
    X = np.column_stack([np.ones(keep.size), d["cov"][keep]]).astype(np.float32)
    means, inv = oracle.stats(d["body"], d["N"], d["M"])
    f = means / 2
    grm = [j for j in range(d["M"]) if inv[j] != 0 and min(f[j], 1 - f[j]) >= 0.01]
    got = {r[1]: float(r[10]) for r in rg}
    got_beta = {r[1]: float(r[7]) for r in rg}
    for chrom in ("1", "10", "2"):
        snps = np.array([j for j in range(d["M"]) if d["chroms"][j] == chrom], dtype=np.int32)
        var = np.array([j for j in grm if d["chroms"][j] != chrom], dtype=np.int32)
        # the driver's calibration sample: std::shuffle of the chromosome's SNPs with
        # std::mt19937_64(seed = 42), the first n_calibration
        cal = np.array(oracle.std_shuffle(snps.tolist(), 42)[:24], dtype=np.int32)
        st, gamma, _ = oracle.pcg_projected_scan(d["body"], d["N"], d["M"], keep, var, d["y"][keep], X, 0.4, 0.6,
                                                 cal, exact_log10p=0.0)
        lo = st[0]["log10p"][snps]
        ids = [f"rs{j:06d}" for j in snps]
        lg = np.array([got.get(i, np.nan) for i in ids])
        assert np.array_equal(np.isnan(lo), np.isnan(lg))
        m = ~np.isnan(lo)
        # .assoc.txt holds 6 significant digits
        assert np.max(np.abs(lo[m] - lg[m]) / np.maximum(1.0, np.abs(lo[m]))) <= 1e-3
        bo = st[0]["beta"][snps][m]
        bg = np.array([got_beta[i] for i, mm in zip(ids, m) if mm])
        assert np.max(np.abs(bo - bg) / st[0]["se"][snps][m]) <= 1e-3 + 1e-5
