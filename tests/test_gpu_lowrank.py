"""GPU parity of the randomized low-rank matrix-free null model and scanLowRank against the oracle.

The sketch uses the reference's own generator stream (std::mt19937_64 + std::normal_distribution,
restated bit-exactly in the oracle and pinned by known answers), so GPU and oracle factor the same
sketch; they differ only through the int8 operator's 23-bit RHS quantisation and LAPACK vs
rocSOLVER rounding.  Bounds: eigenvalues 1e-5 of the largest, r0 1e-5 relative, variance
components 1e-4 relative, |d log10 P| <= 1e-3, |d beta| <= 1e-3 SE.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def csa():
    import cosmicscanner_amd as csa
    return csa


@pytest.mark.parametrize("rank,threshold,no_fallback", [(48, 0.95, False), (40, 0.10, False), (40, 0.99, True)])
def test_lowrank_vs_oracle(csa, rank, threshold, no_fallback):
    N, M = 460, 2600
    body = oracle.gen_hwe_bed(31, N, M, 0.05, 0.5, 0.01)
    rng = np.random.default_rng(5)
    keep = np.sort(rng.choice(N, 420, replace=False)).astype(np.int32)
    variants = np.arange(0, 2100, dtype=np.int32)
    n = keep.size
    X = np.column_stack([np.ones(n), rng.standard_normal((n, 2))]).astype(np.float32)
    K = oracle.dense_kinship(body, N, M, keep, variants)
    g = np.linalg.cholesky(K + 1e-6 * np.eye(n)) @ rng.standard_normal(n)
    y = 0.3 * X[:, 1] + g * 0.8 + rng.standard_normal(n)
    r = csa.PlinkReader.from_body(body, N, M)
    op = csa.PlinkKinshipOperator(r, keep, variants)
    lr = csa.LowRankLmm(op, y, X, rank=rank, oversample=10, seed=42, trace_threshold=threshold,
                        no_fallback=no_fallback)
    m = oracle.lowrank_prepare(K, y, X, rank, 10, 42, threshold, 1, no_fallback)
    got = lr.get()
    print(f"rank {got['rank']} (oracle {m['rank']}), fell back {lr.diag['fell_back']}, "
          f"sg {got['sg']:.6f}/{m['sg']:.6f} se {got['se']:.6f}/{m['se']:.6f}")
    assert got["rank"] == m["rank"] and bool(lr.diag["fell_back"]) == m["fell_back"]
    assert np.max(np.abs(got["lam"] - m["lam"])) <= 1e-5 * m["lam"][0]
    assert abs(got["r0"] - m["r0"]) <= 1e-5 * m["r0"]
    assert abs(got["sg"] / m["sg"] - 1) <= 1e-4 and abs(got["se"] / m["se"] - 1) <= 1e-4
    assert abs(lr.diag["trace"] - m["trace"]) <= 1e-10 * m["trace"]
    snps = np.arange(2100, 2600, dtype=np.int32)
    out = lr.scan(r, keep, snps=snps)
    st = oracle.lowrank_scan(m, body, N, M, keep)
    lp, lo = out["log10p"], st["log10p"][snps]
    assert np.array_equal(np.isnan(lp), np.isnan(lo))
    t = ~np.isnan(lo)
    assert np.max(np.abs(lp[t] - lo[t])) <= 1e-3
    assert np.max(np.abs(out["beta"][t] - st["beta"][snps][t]) / st["se"][snps][t]) <= 1e-3


def test_lowrank_rank_deficient_sketch(csa):
    """Fewer selected SNPs (30) than sketch columns (40 + 10): K Omega is rank-deficient, so the
    sketch orthonormalisation meets a singular (or 1e6-ill-conditioned) Gram matrix.  The leading
    30 eigenvalues still match the oracle's and the rest vanish."""
    N, M = 460, 2600
    body = oracle.gen_hwe_bed(31, N, M, 0.05, 0.5, 0.01)
    rng = np.random.default_rng(6)
    keep = np.sort(rng.choice(N, 420, replace=False)).astype(np.int32)
    variants = np.arange(100, 130, dtype=np.int32)
    n = keep.size
    X = np.column_stack([np.ones(n), rng.standard_normal((n, 2))]).astype(np.float32)
    K = oracle.dense_kinship(body, N, M, keep, variants)
    y = 0.3 * X[:, 1] + rng.standard_normal(n)
    r = csa.PlinkReader.from_body(body, N, M)
    op = csa.PlinkKinshipOperator(r, keep, variants)
    lr = csa.LowRankLmm(op, y, X, rank=40, oversample=10, seed=42, trace_threshold=0.0)
    m = oracle.lowrank_prepare(K, y, X, 40, 10, 42, 0.0, 1, False)
    got = lr.get()
    top = np.linalg.eigvalsh(K)[::-1][:30]
    print(f"rank {got['rank']}: lam[29] {got['lam'][29]:.3e}, lam[30] {got['lam'][30]:.3e}, "
          f"oracle lam[30] {m['lam'][30]:.3e}", flush=True)
    assert got["rank"] == 40
    assert np.max(np.abs(got["lam"][:30] - m["lam"][:30])) <= 1e-5 * m["lam"][0]
    assert np.max(np.abs(got["lam"][:30] - top)) <= 1e-5 * top[0]
    assert np.max(np.abs(got["lam"][30:])) <= 1e-5 * top[0]
    assert np.isfinite(got["sg"]) and np.isfinite(got["se"])
    lr.close()
    op.close()
