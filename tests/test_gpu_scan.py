"""GPU parity of kernel (c), the per-SNP scan, against the CPU oracle through the C-ABI.

Tolerances: freq / n_obs / MAF pass are bit-exact (integer counts, same FP64 formula); g^T g is
FP64 (1e-12 relative); dot products carry the 23-bit fixed-point quantisation of each R column
(<= 2e-6 of max_c |dot| per column).  Association statistics against the FP64 oracle:
|d(-log10 P)| <= 1e-4 (target 1e-3), |d beta| <= 1e-5 SE (beta near 0 makes a relative bound
meaningless), relative SE <= 1e-5.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def csa():
    import cosmicscanner_amd as csa
    return csa


def test_scan_sums_keep_subset_and_snp_list(csa, ref_datasets):
    for name in ("synth", "gemma"):
        body, N, M, _, _ = ref_datasets[name]
        reader = csa.PlinkReader.from_body(body, N, M)
        rng = np.random.default_rng(11)
        keep = rng.choice(N, int(N * 0.7), replace=False).astype(np.int32)   # unsorted subset
        keep = np.concatenate([keep, keep[:5]])                              # with repeats
        R = rng.standard_normal((keep.size, 6))
        snps = np.sort(rng.choice(M, M // 2, replace=False)).astype(np.int32)[::-1].copy()
        freq, nobs, ps, dot, gtg = csa.scan_sums(reader, keep, R, 0.05, snps)
        f0, n0, p0, d0, g0 = oracle.scan_sums(body, N, M, keep, R, min_maf=0.05)
        assert np.array_equal(freq, f0[snps])
        assert np.array_equal(nobs, n0[snps])
        assert np.array_equal(ps, p0[snps])
        assert np.max(np.abs(gtg - g0[snps]) / np.maximum(g0[snps], 1e-300)) < 1e-12
        den = np.max(np.abs(d0[snps]), axis=0)
        assert np.max(np.max(np.abs(dot - d0[snps]), axis=0) / den) < 2e-6


@pytest.mark.parametrize("nr", [20, 21, 36])
def test_scan_partial_last_group_bitwise(csa, nr):
    """A last partial group of at most five columns (three limbs) runs as a packed-limb product of
    its own: every column's dot products equal those of the same column scanned in a block of its
    own, bit for bit, and match the oracle.  (21 columns: five past 16, the largest split; 36: four
    past 32, after two full groups.)"""
    N, M = 1100, 2600
    reader = csa.PlinkReader.synthetic(77, N, M, 0.02, 0.5, 0.01)
    body = reader.download()
    rng = np.random.default_rng(5)
    keep = rng.choice(N, 900, replace=False).astype(np.int32)
    R = rng.standard_normal((keep.size, nr))
    freq, nobs, ps, dot, gtg = csa.scan_sums(reader, keep, R, 0.05, None)
    k1 = nr - nr % 16
    _, _, _, d1, _ = csa.scan_sums(reader, keep, np.ascontiguousarray(R[:, :k1]), 0.05, None)
    _, _, _, d2, _ = csa.scan_sums(reader, keep, np.ascontiguousarray(R[:, k1:]), 0.05, None)
    assert np.array_equal(dot[:, :k1], d1)
    assert np.array_equal(dot[:, k1:], d2)
    f0, n0, p0, d0, g0 = oracle.scan_sums(body, N, M, keep, R, min_maf=0.05)
    assert np.array_equal(freq, f0) and np.array_equal(nobs, n0) and np.array_equal(ps, p0)
    den = np.max(np.abs(d0), axis=0)
    assert np.max(np.max(np.abs(dot - d0), axis=0) / den) < 2e-6
    reader.close()


def _small_problem(seed=21, N=400, M=1500):
    body = oracle.gen_hwe_bed(seed, N, M, 0.02, 0.5, 0.01)
    G, _, _ = oracle.grm(body, N, M, alg=1, min_maf=0.01)
    rng = np.random.default_rng(seed)
    X = np.column_stack([np.ones(N), rng.standard_normal((N, 3)).astype(np.float32)])
    L = np.linalg.cholesky(G + 1e-6 * np.eye(N))
    y = X @ np.array([0.3, 0.5, -0.2, 0.1]) + L @ rng.standard_normal(N) + rng.standard_normal(N)
    return body, N, M, G, X, y


def test_rotated_dense_scan(csa):
    body, N, M, G, X, y = _small_problem()
    D, W, Uty = oracle.eigen_vce_prepare(G, y, X)
    vg, ve, delta = oracle.solve_diagonal(D, Uty)
    Vinv = 1.0 / (vg * D + ve)
    Py = W @ (Uty * Vinv)                       # score column: g^T W_proj (V^-1 Uty)
    Rrot = W * np.sqrt(Vinv)[None, :]           # var = sum_i (W_i^T g)^2 V^-1_i
    R = np.column_stack([Py, Rrot])
    reader = csa.PlinkReader.from_body(body, N, M)
    out = csa.lmm_scan(reader, None, R, n_score=1, alpha=0.0, quad=(1, R.shape[1]),
                       quad_w=np.ones(R.shape[1] - 1), min_maf=0.01)
    Dd = oracle.decode_numpy(body, N, M)
    mu = np.nanmean(Dd, 0)
    Gm = np.where(np.isnan(Dd), mu, Dd) - mu
    ref = oracle.rotated_scan(W, D, Uty, vg, ve, Gm)
    f0, n0, p0, _, _ = oracle.scan_sums(body, N, M, None, np.zeros((N, 1)), min_maf=0.01)
    assert np.array_equal(out["passes"], p0)
    t = p0 & ref["tested"]
    assert np.array_equal(np.isfinite(out["p"][:, 0]), t)
    assert np.max(np.abs(out["log10p"][t, 0] - ref["log10p"][t])) < 1e-4
    assert np.max(np.abs(out["beta"][t, 0] - ref["beta"][t]) / ref["se"][t]) < 1e-5
    assert np.max(np.abs(out["se"][t, 0] / ref["se"][t] - 1)) < 1e-5


def test_low_rank_woodbury_scan_multi_trait(csa):
    body, N, M, G, X, y = _small_problem(seed=5, N=350, M=1200)
    rng = np.random.default_rng(3)
    Y = np.column_stack([y, y + rng.standard_normal(N), rng.standard_normal(N)])
    lam, U = np.linalg.eigh(G)
    r = 40
    lam, U = lam[::-1][:r], U[:, ::-1][:, :r]
    r0 = (np.trace(G) - lam.sum()) / (N - r)
    vg, ve = 0.6, 0.9
    base = 1.0 / (vg * r0 + ve)
    corr = 1.0 / (vg * lam + ve) - base
    Vinv = base * np.eye(N) + (U * corr) @ U.T
    info = X.T @ Vinv @ X
    C = np.linalg.inv(info)
    beta0 = C @ (X.T @ Vinv @ Y)
    resid = Y - X @ beta0
    py = Vinv @ resid
    R = np.column_stack([py, U, Vinv @ X])
    nt = Y.shape[1]
    reader = csa.PlinkReader.from_body(body, N, M)
    out = csa.lmm_scan(reader, None, R, n_score=nt, alpha=base, quad=(nt, nt + r), quad_w=corr,
                       cov=(nt + r, nt + r + X.shape[1]), cov_C=C, min_maf=0.01)
    Dd = oracle.decode_numpy(body, N, M)
    mu = np.nanmean(Dd, 0)
    Gm = np.where(np.isnan(Dd), mu, Dd) - mu
    V = np.linalg.inv(Vinv)
    for t in range(nt):
        ref = oracle.gls_scan(V, X, Y[:, t], Gm)
        ok = ref["tested"] & out["passes"]
        assert np.max(np.abs(out["log10p"][ok, t] - ref["log10p"][ok])) < 1e-4
        assert np.max(np.abs(out["beta"][ok, t] - ref["beta"][ok]) / ref["se"][ok]) < 1e-5
        assert np.max(np.abs(out["se"][ok, t] / ref["se"][ok] - 1)) < 1e-5


def test_scan_device_outputs_equal_host(csa, ref_datasets):
    """cs_lmm_scan_device (R and every statistic in HBM, written by the epilogue directly) gives
    the same bits as cs_lmm_scan with host buffers, over a keep subset with repeats and a SNP list;
    the pass timings are reported."""
    body, N, M, _, _ = ref_datasets["gemma"]
    reader = csa.PlinkReader.from_body(body, N, M)
    rng = np.random.default_rng(5)
    keep = rng.choice(N, int(N * 0.8), replace=False).astype(np.int32)
    keep = np.concatenate([keep, keep[:3]])
    T, c = 3, 2
    R = rng.standard_normal((keep.size, T + c))
    snps = np.sort(rng.choice(M, M // 2, replace=False)).astype(np.int32)
    C = np.array([[0.02, 0.001], [0.001, 0.03]])
    kw = dict(n_score=T, alpha=0.9, cov=(T, T + c), cov_C=C, min_maf=0.02, snps=snps)
    host = csa.lmm_scan(reader, keep, R, **kw)
    dR = csa.DeviceArray.from_host(R)
    out, timing = csa.lmm_scan_device(reader, keep, dR, **kw)
    ns = snps.size

    def rows(d, width):
        return d.download().reshape(-1, order="F")[: ns * width].reshape(ns, width)

    for key in ("beta", "se", "p", "log10p"):
        got = rows(out[key], T)
        assert np.array_equal(got, host[key], equal_nan=True), key
    assert np.array_equal(rows(out["freq"], 1)[:, 0], host["freq"])
    assert np.array_equal(rows(out["n_obs"], 1)[:, 0].view(np.int64), host["n_obs"])
    passes = out["passes"].download().reshape(-1, order="F").view(np.int32)[:ns]
    assert np.array_equal(passes.astype(bool), host["passes"])
    ms_counts, ms_prod, ms_epi, nbytes = timing
    assert ms_prod > 0 and ms_epi > 0 and ms_counts > 0 and nbytes > 0
    for a in [dR] + list(out.values()):
        a.close()
    reader.close()
