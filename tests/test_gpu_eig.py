"""GPU checks of the dense symmetric eigendecomposition behind EigenVce::prepare and
MvEigenVce::prepare (libCosmicCore/src/eigen_vce.cpp:34-38, 520-526): rocSOLVER's dsyevd up to
n = 32766 and the 64-bit path beyond (blocked Householder tridiagonal reduction, divide and
conquer with rank-one merges, compact-WY back-transformation).  The leaf size of the divide and
conquer is a parameter, so small matrices exercise the merges (deflation of small weights and of
close pole pairs, secular roots, Gu-Eisenstat eigenvectors) that large ones need.

Checks against LAPACK (numpy.linalg.eigvalsh) and the definition of an eigendecomposition, for
spectra that stress each step -- random, repeated eigenvalues (rotation deflation), low rank
(many zero eigenvalues: weight deflation), diagonal (a split tridiagonal: rho = 0), a graded
spectrum over 12 decades, and a GRM from the HWE genotype generator:
  |lambda - lambda_LAPACK| <= 1e-12 n^0.5 ||A||,  ||A V - V diag(lambda)||_max <= 1e-12 n^0.5 ||A||,
  ||V^T V - I||_max <= 1e-12 n^0.5.
Through the LMM front ends (CSAMD_EIG_LARGE=<leaf> forces the 64-bit path): the rotated scan and
the multivariate scan against the oracle at the tolerances of their own tests.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def csa():
    import cosmicscanner_amd as csa
    return csa


def make_matrix(kind, n, seed):
    rng = np.random.default_rng(seed)
    if kind == "random":
        A = rng.standard_normal((n, n))
        return (A + A.T) / 2
    if kind == "diagonal":
        return np.diag(rng.standard_normal(n))
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    if kind == "repeated":
        lam = rng.choice([-1.0, 0.5, 2.0, 3.0], n)
    elif kind == "lowrank":
        lam = np.zeros(n)
        lam[: max(1, n // 5)] = rng.uniform(0.5, 2.0, max(1, n // 5))
    elif kind == "graded":
        lam = 10.0 ** rng.uniform(-12, 0, n)
    else:
        raise ValueError(kind)
    A = (Q * lam) @ Q.T
    return (A + A.T) / 2


def grm_matrix(n, seed):
    body = oracle.gen_hwe_bed(seed, n, 3 * n // 2, 0.05, 0.5, 0.01)
    G, _, _ = oracle.grm(body, n, 3 * n // 2, alg=1, min_maf=0.01)
    return G


def run_eig(csa, A, leaf):
    n = A.shape[0]
    dA = csa.DeviceArray.from_host(A)
    dD = csa.DeviceArray(n, 1)
    csa.sym_eigen_device(dA, dD, leaf=leaf)
    V, lam = dA.download(), dD.download()[:, 0]
    dA.close()
    dD.close()
    return lam, V


def check(A, lam, V, tag):
    n = A.shape[0]
    ref = np.linalg.eigvalsh(A)
    scale = max(np.max(np.abs(ref)), 1e-300)
    tol = 1e-12 * max(1.0, np.sqrt(n))
    dl = np.max(np.abs(lam - ref)) / scale
    res = np.max(np.abs(A @ V - V * lam[None, :])) / scale
    orth = np.max(np.abs(V.T @ V - np.eye(n)))
    print(f"  {tag}: |dlambda| {dl:.1e}, residual {res:.1e}, orthogonality {orth:.1e}", flush=True)
    assert np.all(np.diff(lam) >= 0)
    assert dl <= tol and res <= tol and orth <= tol


@pytest.mark.parametrize("kind", ["random", "repeated", "lowrank", "diagonal", "graded"])
@pytest.mark.parametrize("n,leaf", [(1, 16), (2, 1), (7, 2), (129, 16), (700, 64), (700, 32766), (1500, 200)])
def test_sym_eigen_large_path(csa, kind, n, leaf):
    print(f"eig kind={kind} n={n} leaf={leaf}", flush=True)
    A = make_matrix(kind, n, 1000 + n)
    lam, V = run_eig(csa, A, leaf)
    check(A, lam, V, f"{kind} n={n} leaf={leaf}")


@pytest.mark.parametrize("leaf", [0, 100])
def test_sym_eigen_grm(csa, leaf):
    n = 900
    print(f"eig GRM n={n} leaf={leaf}", flush=True)
    G = grm_matrix(n, 77)
    lam, V = run_eig(csa, G, leaf)
    check(G, lam, V, f"GRM leaf={leaf}")


def test_sym_eigen_paths_agree(csa):
    # the two routes on one matrix: same spectrum, same invariant subspaces (projector of the
    # 50 largest eigenvalues; eigenvector signs are arbitrary)
    n = 1100
    G = grm_matrix(n, 78)
    l0, V0 = run_eig(csa, G, 0)
    l1, V1 = run_eig(csa, G, 128)
    assert np.max(np.abs(l0 - l1)) <= 1e-12 * np.sqrt(n) * np.max(np.abs(l0))
    P0 = V0[:, -50:] @ V0[:, -50:].T
    P1 = V1[:, -50:] @ V1[:, -50:].T
    assert np.max(np.abs(P0 - P1)) <= 1e-9


def test_rotated_scan_through_large_path(csa, monkeypatch):
    monkeypatch.setenv("CSAMD_EIG_LARGE", "64")
    seed, N, M = 21, 400, 1500
    body = oracle.gen_hwe_bed(seed, N, M, 0.02, 0.5, 0.01)
    G, _, _ = oracle.grm(body, N, M, alg=1, min_maf=0.01)
    rng = np.random.default_rng(seed)
    X = np.column_stack([np.ones(N), rng.standard_normal((N, 3))]).astype(np.float32)
    L = np.linalg.cholesky(G + 1e-6 * np.eye(N))
    y = X.astype(np.float64) @ np.array([0.3, 0.5, -0.2, 0.1]) + L @ rng.standard_normal(N) + rng.standard_normal(N)
    lmm = csa.DenseLmm(y, X)
    lmm.prepare(G)
    D, Uty = lmm.eigen()
    Do, Wo, Uo = oracle.eigen_vce_prepare(G, y, X)
    assert np.max(np.abs(D - Do)) <= 1e-9 * np.max(Do)
    assert np.max(np.abs(np.abs(Uty) - np.abs(Uo))) <= 1e-8 * np.max(np.abs(Uo))
    lmm.runNullModel()
    vg, ve, _ = oracle.solve_diagonal(Do, Uo)
    assert abs(lmm.sigma2_g / vg - 1) <= 1e-6 and abs(lmm.sigma2_e / ve - 1) <= 1e-6
    reader = csa.PlinkReader.from_body(body, N, M)
    out = lmm.scan(reader, None, min_maf=0.01)
    Dd = oracle.decode_numpy(body, N, M)
    mu = np.nanmean(Dd, 0)
    Gm = np.where(np.isnan(Dd), mu, Dd) - mu
    ref = oracle.rotated_scan(Wo, Do, Uo, lmm.sigma2_g, lmm.sigma2_e, Gm)
    t = out["passes"] & ref["tested"]
    assert np.array_equal(np.isfinite(out["p"]), t)
    assert np.max(np.abs(out["log10p"][t] - ref["log10p"][t])) < 1e-4
    assert np.max(np.abs(out["beta"][t] - ref["beta"][t]) / ref["se"][t]) < 1e-5
    lmm.close()
    reader.close()


def test_mv_scan_through_large_path(csa, monkeypatch):
    monkeypatch.setenv("CSAMD_EIG_LARGE", "50")
    N, M, d = 333, 1100, 3
    bed = oracle.gen_hwe_bed(23, N, M, 0.05, 0.5, 0.01)
    raw, _, S, _ = oracle.grm_exact_centred(bed, N, M)
    G = raw / S
    rng = np.random.default_rng(23)
    X = np.column_stack([np.ones(N), rng.standard_normal((N, 2))]).astype(np.float32)
    Y = rng.standard_normal((N, d)) + X[:, 1:2].astype(np.float64) * 0.3
    mv = csa.MvLmm(Y, X)
    mv.prepare(G)
    D, Uty = mv.eigen()
    Do, Wo, Uo, _, _ = oracle.mv_prepare(G, Y, X.astype(np.float64))
    assert np.max(np.abs(D - Do)) <= 1e-9 * np.max(Do)
    assert np.max(np.abs(np.abs(Uty) - np.abs(Uo))) <= 1e-8 * np.max(np.abs(Uo))
    A = rng.normal(0, 0.3, (d, d))
    Vg = A @ A.T + 0.2 * np.eye(d)
    B = rng.normal(0, 0.3, (d, d))
    Ve = B @ B.T + 0.6 * np.eye(d)
    mv.set_variance(Vg, Ve)
    r = csa.PlinkReader.from_body(bed, N, M)
    o = mv.scan(r, None, None)
    oo = oracle.mv_scan(bed, N, M, None, Wo, Do, Uo, Vg, Ve, None)
    tested = oo["status"] >= 0
    assert np.array_equal(o["status"] >= 0, tested)
    t = np.flatnonzero(tested)
    assert np.max(np.abs(np.log10(o["p"][t]) - np.log10(oo["p"][t]))) <= 1e-3
    mv.close()
    r.close()
