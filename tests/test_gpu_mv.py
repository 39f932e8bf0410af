"""GPU parity of the multivariate LMM (MvEigenVce, libCosmicCore/src/eigen_vce.cpp:496-910;
MvLmmEngine::scan, src/mvlmm_engine.cpp:210-394, 756-931; MultivariateModel::run,
src/models/multivariate_model.cpp:870-1110) against the CPU oracle, through the C-ABI.

Tolerances.
* Eigenvalues (rocSOLVER vs LAPACK): 1e-9 relative to the largest; |Uty| row-wise (eigenvector
  signs are arbitrary) 1e-8 of the largest.
* EM with a fixed iteration count: Vg, Ve within 1e-9 relative (same FP64 recursion, per-sample
  sums in a different order).  AI-REML at convergence: 1e-6 relative (both reach the REML optimum
  to ai_tol).
* Scan: the rotation W^T g runs as FP16 hi/lo split products with FP32 accumulation over the
  samples, so the statistics carry ~1e-6 relative error: |d log10 P| <= 1e-3, |d beta| <= 1e-3 SE,
  SE relative 1e-4, |d Wald| <= 1e-4 (Wald + 1).
* Driver (LOCO models, EM stopping on a squared-change threshold): Vg, Ve 1e-4 relative.
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


def make_case(N, M, d, seed, miss=0.01):
    bed = oracle.gen_hwe_bed(seed, N, M, 0.05, 0.5, miss)
    raw, _, S, _ = oracle.grm_exact_centred(bed, N, M)
    rng = np.random.default_rng(seed)
    X = np.column_stack([np.ones(N), rng.standard_normal((N, 2))]).astype(np.float32)
    Dm = oracle.decode_numpy(bed, N, M)
    Dm = np.where(np.isnan(Dm), np.nanmean(Dm, 0), Dm)
    L = np.linalg.cholesky(np.full((d, d), 0.3) + 0.7 * np.eye(d))
    g = (Dm[:, :40] - Dm[:, :40].mean(0)) @ rng.normal(0, 0.2, (40, d)) @ L.T
    Y = X[:, 1:2].astype(np.float64) * 0.3 + g + rng.normal(0, 1, (N, d)) @ L.T
    return bed, raw / S, X, Y


def test_mv_prepare_and_em(csa):
    N, M, d = 300, 1200, 3
    bed, G, X, Y = make_case(N, M, d, 11)
    print(f"mv prepare N={N} M={M} d={d}", flush=True)
    mv = csa.MvLmm(Y, X)
    mv.prepare(G)
    D, Uty = mv.eigen()
    Do, Wo, Uo, Vg0, Ve0 = oracle.mv_prepare(G, Y, X.astype(np.float64))
    assert np.max(np.abs(D - Do)) <= 1e-9 * np.max(Do)
    assert np.max(np.abs(np.abs(Uty) - np.abs(Uo))) <= 1e-8 * np.max(np.abs(Uo))
    Vg, Ve, _, _, _ = mv.variance()
    assert np.allclose(Vg, Vg0, rtol=1e-12, atol=0) and np.allclose(Ve, Ve0, rtol=1e-12, atol=0)
    conv, em_it, ai_it = mv.runNullModel("em", max_iter=40, tol=0.0)
    assert not conv and em_it == 40 and ai_it == 0
    Vg, Ve, _, _, ll = mv.variance()
    Vgo, Veo, _, _, _ = oracle.mv_null(Do, Uo, Vg0, Ve0, "em", max_iter=40, tol=0.0)
    assert np.max(np.abs(Vg - Vgo)) <= 1e-9 * np.max(np.abs(Vgo))
    assert np.max(np.abs(Ve - Veo)) <= 1e-9 * np.max(np.abs(Veo))
    assert abs(ll - oracle.mv_logl(Do, Uo, Vgo, Veo)) <= 1e-9 * abs(ll)
    mv.close()


def test_mv_aireml(csa):
    N, M, d = 400, 1500, 2
    bed, G, X, Y = make_case(N, M, d, 12)
    mv = csa.MvLmm(Y, X)
    mv.prepare(G)
    conv, em_it, ai_it = mv.runNullModel("aireml")
    Vg, Ve, Vg_se, Ve_se, ll = mv.variance()
    Do, Wo, Uo, Vg0, Ve0 = oracle.mv_prepare(G, Y, X.astype(np.float64))
    Vgo, Veo, convo, em_o, ai_o = oracle.mv_null(Do, Uo, Vg0, Ve0, "aireml")
    print(f"  aireml: gpu em {em_it} ai {ai_it}; oracle em {em_o} ai {ai_o}", flush=True)
    assert conv and convo and ai_it > 0
    assert np.max(np.abs(Vg - Vgo)) <= 1e-6 * np.max(np.abs(Vgo))
    assert np.max(np.abs(Ve - Veo)) <= 1e-6 * np.max(np.abs(Veo))
    assert abs(ll - oracle.mv_logl(Do, Uo, Vgo, Veo)) <= 1e-8 * abs(ll)
    mv.close()


def compare_scan(o_gpu, o_orc, tag):
    tested = o_orc["status"] >= 0
    assert np.array_equal(o_gpu["status"] >= 0, tested)
    assert np.allclose(o_gpu["freq"], o_orc["freq"], rtol=1e-14, atol=0)
    t = np.flatnonzero(tested)
    se_o = np.sqrt(np.diagonal(o_orc["vbeta"][t], axis1=1, axis2=2))
    se_g = np.sqrt(np.diagonal(o_gpu["vbeta"][t], axis1=1, axis2=2))
    dl = np.max(np.abs(np.log10(o_gpu["p"][t]) - np.log10(o_orc["p"][t])))
    db = np.max(np.abs(o_gpu["beta"][t] - o_orc["beta"][t]) / se_o)
    ds = np.max(np.abs(se_g - se_o) / se_o)
    dw = np.max(np.abs(o_gpu["wald"][t] - o_orc["wald"][t]) / (o_orc["wald"][t] + 1.0))
    print(f"  {tag}: {t.size} tested, |dlog10P| {dl:.2e}, |dbeta|/SE {db:.2e}, dSE {ds:.2e}, dWald {dw:.2e}",
          flush=True)
    assert dl <= 1e-3 and db <= 1e-3 and ds <= 1e-4 and dw <= 1e-4


@pytest.mark.parametrize("d", [1, 2, 4])
def test_mv_scan(csa, d):
    N, M = 333, 1100
    bed, G, X, Y = make_case(N, M, d, 20 + d)
    print(f"mv scan N={N} M={M} d={d}", flush=True)
    r = csa.PlinkReader.from_body(bed, N, M)
    mv = csa.MvLmm(Y, X)
    mv.prepare(G)
    Do, Wo, Uo, _, _ = oracle.mv_prepare(G, Y, X.astype(np.float64))
    rng = np.random.default_rng(d)
    A = rng.normal(0, 0.3, (d, d))
    Vg = A @ A.T + 0.2 * np.eye(d)
    B = rng.normal(0, 0.3, (d, d))
    Ve = B @ B.T + 0.6 * np.eye(d)
    mv.set_variance(Vg, Ve)
    snps = np.arange(0, M, 3)
    o = mv.scan(r, None, snps)
    oo = oracle.mv_scan(bed, N, M, None, Wo, Do, Uo, Vg, Ve, snps)
    compare_scan(o, oo, f"d={d}")
    mv.close()


def test_mv_scan_keep_subset(csa):
    N, M, d = 420, 900, 2
    bed, Gfull, X, Y = make_case(N, M, d, 31)
    rng = np.random.default_rng(31)
    keep = np.sort(rng.choice(N, 350, replace=False)).astype(np.int32)[::-1].copy()
    G = Gfull[np.ix_(keep, keep)]
    r = csa.PlinkReader.from_body(bed, N, M)
    mv = csa.MvLmm(Y[keep], X[keep])
    mv.prepare(G)
    conv, _, _ = mv.runNullModel("aireml")
    Vg, Ve, _, _, _ = mv.variance()
    Do, Wo, Uo, _, _ = oracle.mv_prepare(G, Y[keep], X[keep].astype(np.float64))
    o = mv.scan(r, keep, None)
    oo = oracle.mv_scan(bed, N, M, keep, Wo, Do, Uo, Vg, Ve, None)
    compare_scan(o, oo, "keep subset")
    mv.close()


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


@pytest.mark.parametrize("solver,grm_alg", [("em", 1), ("aireml", 0)])
def test_gwas_mv_driver(csa, tmp_path, solver, grm_alg):
    N, M, d = 360, 1500, 3
    bed, _, X, Y = make_case(N, M, d, 41)
    chroms = ["1"] * 600 + ["2"] * 400 + ["10"] * 500
    prefix = str(tmp_path / "mv")
    write_plink(prefix, bed, N, M, chroms)
    Yn = Y.copy()
    Yn[5, 1] = np.nan
    cov = X[:, 1:].astype(np.float64)
    out = str(tmp_path / ("run_" + solver))
    print(f"mv driver solver={solver} alg={grm_alg}", flush=True)
    sm = csa.gwas_mv_run(prefix, Yn, out, cov=cov, solver=solver, grm_alg=grm_alg)
    models, rows = oracle.gwas_mv(bed, N, M, chroms, Yn, cov, solver=solver, grm_alg=grm_alg)
    assert sm["loco"] and [m[0] for m in sm["models"]] == ["1", "2", "10"]
    tol = 1e-4
    for (lg, vg, ve), (lo, vgo, veo) in zip(sm["models"], models):
        assert lg == lo
        assert np.max(np.abs(vg - vgo)) <= tol * np.max(np.abs(vgo))
        assert np.max(np.abs(ve - veo)) <= tol * np.max(np.abs(veo))
    lines = open(out + ".assoc.mv.txt").read().splitlines()
    assert lines[0].split("\t") == ["CHR", "SNP", "BP", "A1", "A2", "FREQ", "WALD_STAT", "P",
                                    "BETA_1", "SE_1", "BETA_2", "SE_2", "BETA_3", "SE_3"]
    body = [ln.split("\t") for ln in lines[1:]]
    assert sm["n_tested"] == len(body) == len(rows)
    worst = 0.0
    for rg, (j, freq, wald, p, beta, se) in zip(body, rows):
        assert rg[1] == f"rs{j:06d}" and rg[0] == chroms[j] and rg[3] == "A" and rg[4] == "C"
        assert rg[5] == f"{freq:.6g}"
        assert abs(np.log10(float(rg[7])) - np.log10(p)) <= 1e-3
        for t in range(d):
            assert abs(float(rg[8 + 2 * t]) - beta[t]) <= 1e-3 * se[t] + 5e-6 * abs(beta[t])
            assert abs(float(rg[9 + 2 * t]) - se[t]) <= 1e-4 * se[t]
        worst = max(worst, abs(np.log10(float(rg[7])) - np.log10(p)))
    print(f"  driver: {len(body)} rows, max |dlog10P| {worst:.2e}", flush=True)


def test_mv_errors(csa):
    N = 50
    Y = np.random.default_rng(0).normal(size=(N, 7))
    X = np.ones((N, 1), dtype=np.float32)
    with pytest.raises(ValueError, match="1 to 6 traits"):
        csa.MvLmm(Y, X)
    mv = csa.MvLmm(Y[:, :2], X)
    with pytest.raises(RuntimeError, match="prepare"):
        mv.runNullModel()
    mv.close()
