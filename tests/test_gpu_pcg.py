"""GPU parity of the matrix-free solves and the PCG-projected scan against the CPU oracle.

* PCG (MatrixFreePCG::solve, libCosmicCore/src/matrix_free_pcg.h:20-66, multi-RHS, Jacobi):
  solutions of (sigma_g K + sigma_e I) X = B against a dense FP64 solve of the oracle's K,
  relative error <= 2e-5 (the operator's 23-bit RHS quantisation, amplified by cond(V) < 10);
  the reported true residual meets the tolerance.
* Projected scan, calibrated statistics (exact stage off): against the oracle's dense-solve
  restatement of the same estimator (exact V^-1, same calibration markers): |d log10 P| <= 1e-3,
  gamma relative 1e-4.
* Projected scan with the exact stage: every SNP whose largest -log10 P reaches the threshold
  (level 5 by default, 2 and -1 here, plus log10 of the number of traits) against textbook GLS with an explicit dense V^-1 (the definition of LmmEngine::scan,
  src/lmm_engine.cpp:1363-1525): |d log10 P| <= 1e-3 and |d beta| <= 1e-3 SE, at n = 400 and at
  n = 2000; below the threshold the distance stays within the reported bound.  With threshold -1
  every tested SNP meets the 1e-3 bound.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def csa():
    import cosmicscanner_amd as csa
    return csa


@pytest.fixture(scope="module")
def data():
    N, M = 430, 2600
    body = oracle.gen_hwe_bed(777, N, M, 0.05, 0.5, 0.01)
    rng = np.random.default_rng(3)
    keep = np.sort(rng.choice(N, 400, replace=False)).astype(np.int32)
    loco = np.arange(0, 2000, dtype=np.int32)         # "chromosome" 2000..2599 left out
    return dict(N=N, M=M, body=body, keep=keep, loco=loco, rng=rng)


def test_pcg_solve_vs_dense(csa, data):
    N, M, body, keep, loco = data["N"], data["M"], data["body"], data["keep"], data["loco"]
    r = csa.PlinkReader.from_body(body, N, M)
    op = csa.PlinkKinshipOperator(r, keep, loco)
    n = keep.size
    K = oracle.dense_kinship(body, N, M, keep, loco)
    B = data["rng"].standard_normal((n, 9))
    for sg, se, pre in ((0.4, 0.6, 1), (1.5, 0.2, 0), (0.0, 1.0, 1)):
        dB = csa.DeviceArray.from_host(B)
        dX = csa.DeviceArray(n, 9)
        it, res = csa.pcg_solve(op, sg, se, dB, dX, rel_tol=1e-9, max_iter=500, precond=pre)
        Xr = np.linalg.solve(sg * K + se * np.eye(n), B)
        err = np.max(np.abs(dX.download() - Xr)) / np.max(np.abs(Xr))
        print(f"sg={sg} se={se} precond={pre}: {it} iterations, true residual {res.max():.2e}, err {err:.2e}")
        assert err <= 2e-5 and res.max() <= 2e-5
    # per-column variance components and a warm start that is already converged
    sgs, ses = np.array([0.3, 0.9, 0.5]), np.array([0.7, 0.1, 0.5])
    Bc = B[:, :3]
    dB = csa.DeviceArray.from_host(Bc)
    dX = csa.DeviceArray(n, 3)
    import ctypes
    it = ctypes.c_int()
    res = np.empty(3)
    csa._check(csa.lib().cs_pcg_solve_cols(op._h, sgs.ctypes.data_as(csa._dp), ses.ctypes.data_as(csa._dp),
                                           dB.ptr, n, 3, dX.ptr, n, 1e-9, 500, 1, 0, ctypes.byref(it),
                                           res.ctypes.data_as(csa._dp)))
    for c in range(3):
        xr = np.linalg.solve(sgs[c] * K + ses[c] * np.eye(n), Bc[:, c])
        assert np.max(np.abs(dX.download()[:, c] - xr)) <= 2e-5 * np.max(np.abs(xr))
    it2, _ = csa.pcg_solve(op, 0.4, 0.6, csa.DeviceArray.from_host(B), csa.DeviceArray.from_host(
        np.linalg.solve(0.4 * K + 0.6 * np.eye(n), B)), rel_tol=1e-4, warm_start=True)
    assert it2 <= 1
    with pytest.raises(ValueError):
        csa.pcg_solve(op, 0.4, 0.0, dB, dX)


def test_markers_device(csa, data):
    N, M, body, keep = data["N"], data["M"], data["body"], data["keep"]
    r = csa.PlinkReader.from_body(body, N, M)
    snps = np.array([5, 17, 2599, 1000, 3], dtype=np.int32)
    dG = csa.DeviceArray(keep.size, snps.size)
    ps, _ = csa.markers_device(r, keep, snps, dG)
    G, ok = oracle.prepared_markers(body, N, M, keep, snps)
    assert np.array_equal(ps, ok)
    assert np.max(np.abs(dG.download() - G)) <= 1e-12


def test_projected_scan_vs_oracle(csa, data):
    N, M, body, keep, loco, rng = data["N"], data["M"], data["body"], data["keep"], data["loco"], data["rng"]
    n = keep.size
    r = csa.PlinkReader.from_body(body, N, M)
    op = csa.PlinkKinshipOperator(r, keep, loco)
    X = np.column_stack([np.ones(n), rng.standard_normal((n, 2))]).astype(np.float32)
    Y = rng.standard_normal((n, 3)) + X[:, 1:2]
    cal = rng.choice(np.arange(2000, 2600), 24, replace=False).astype(np.int32)
    snps = np.arange(2000, 2600, dtype=np.int32)
    out = csa.pcg_projected_scan(op, r, keep, Y, X, 0.45, 0.55, cal, snps=snps, rel_tol=1e-9, max_iter=500,
                                 exact_log10p=float("inf"))
    assert out["n_exact"] == 0 and np.isnan(out["dlog10p_bound_below"])
    ref, gamma, freq = oracle.pcg_projected_scan(body, N, M, keep, loco, Y, X, 0.45, 0.55, cal)
    print(f"gamma {out['gamma']:.6f} (oracle {gamma:.6f}), {out['pcg_iterations']} PCG iterations")
    assert abs(out["gamma"] / gamma - 1) <= 1e-4
    assert np.array_equal(out["freq"], freq[snps])
    for t in range(3):
        lp, lr = out["log10p"][:, t], ref[t]["log10p"][snps]
        assert np.array_equal(np.isnan(lp), np.isnan(lr))
        m = ~np.isnan(lr)
        assert np.max(np.abs(lp[m] - lr[m])) <= 1e-3
        b, br, sr = out["beta"][:, t][m], ref[t]["beta"][snps][m], ref[t]["se"][snps][m]
        assert np.max(np.abs(b - br) / sr) <= 1e-3
    # the default exact stage against textbook GLS, and against the oracle's restatement of the
    # combined estimator
    ex_out = csa.pcg_projected_scan(op, r, keep, Y, X, 0.45, 0.55, cal, snps=snps, rel_tol=1e-9, max_iter=500,
                                    exact_log10p=2.0)
    K = oracle.dense_kinship(body, N, M, keep, loco)
    Gm, okm = oracle.prepared_markers(body, N, M, keep, snps)
    check_exact_stage(ex_out, out, K, X, Y, Gm, okm, 0.45, 0.55, 2.0 + np.log10(3))
    refx, _, _ = oracle.pcg_projected_scan(body, N, M, keep, loco, Y, X, 0.45, 0.55, cal, exact_log10p=2.0)
    for t in range(3):
        lp, lr = ex_out["log10p"][:, t], refx[t]["log10p"][snps]
        m = ~np.isnan(lr)
        assert np.max(np.abs(lp[m] - lr[m])) <= 1e-3


def check_exact_stage(out, calib, K, X, Y, Gm, okm, sg, se, thr):
    """Rows whose calibrated (exact stage off: `calib`) largest -log10 P over traits reaches thr
    equal textbook GLS within 1e-3 in -log10 P and 1e-3 SE in beta; rows below keep their
    calibrated values and stay within the reported bound (thr >= 0)."""
    n = K.shape[0]
    V = sg * K + se * np.eye(n)
    T = Y.shape[1]
    L = out["log10p"]
    Lc = calib["log10p"]
    with np.errstate(invalid="ignore"):
        sel = calib["passes"] & (np.any(Lc >= thr, axis=1) if thr >= 0 else ~np.isnan(Lc[:, 0]))
    assert np.array_equal(L[~sel], Lc[~sel], equal_nan=True)
    assert out["n_exact"] == int(sel.sum())
    worst_above, worst_below = 0.0, 0.0
    for t in range(T):
        ex = oracle.gls_scan(V, X.astype(np.float64), Y[:, t], Gm)
        m = okm & ex["tested"] & ~np.isnan(L[:, t])
        d = np.abs(L[:, t] - ex["log10p"])
        db = np.abs(out["beta"][:, t] - ex["beta"]) / ex["se"]
        if np.any(m & sel):
            worst_above = max(worst_above, float(np.max(d[m & sel])))
            assert np.max(db[m & sel]) <= 1e-3
        if np.any(m & ~sel):
            worst_below = max(worst_below, float(np.max(d[m & ~sel])))
    print(f"exact stage: {int(sel.sum())} SNPs at or above {thr}, max |d log10 P| {worst_above:.3e}; "
          f"below: {worst_below:.3e} (stated bound {out['dlog10p_bound_below']:.3e}, "
          f"calibration spread {out['calibration_rel_sd']:.3e}), {out['exact_pcg_iterations']} PCG iterations",
          flush=True)
    assert worst_above <= 1e-3
    if thr >= 0:
        assert worst_below <= out["dlog10p_bound_below"]


def test_projected_scan_exact_gls_n2000(csa):
    """n = 2000 samples, 6000 LOCO kinship SNPs, 1500 scanned SNPs with 12 causal ones: the exact
    stage at the default threshold and at -1 (every SNP) against textbook GLS."""
    N, M = 2100, 7500
    body = oracle.gen_hwe_bed(2026, N, M, 0.05, 0.5, 0.01)
    rng = np.random.default_rng(21)
    keep = np.sort(rng.choice(N, 2000, replace=False)).astype(np.int32)
    n = keep.size
    loco = np.arange(0, 6000, dtype=np.int32)
    snps = np.arange(6000, 7500, dtype=np.int32)
    r = csa.PlinkReader.from_body(body, N, M)
    op = csa.PlinkKinshipOperator(r, keep, loco)
    Gm, okm = oracle.prepared_markers(body, N, M, keep, snps)
    X = np.column_stack([np.ones(n), rng.standard_normal((n, 2))]).astype(np.float32)
    causal = rng.choice(np.nonzero(okm)[0], 12, replace=False)
    Y = rng.standard_normal((n, 2)) + X[:, 1:2]
    Y += Gm[:, causal] @ rng.normal(0.0, 0.25, (12, 2))
    cal = rng.choice(snps, 24, replace=False).astype(np.int32)
    K = oracle.dense_kinship(body, N, M, keep, loco)
    calib = csa.pcg_projected_scan(op, r, keep, Y, X, 0.4, 0.6, cal, snps=snps, rel_tol=1e-6, max_iter=300,
                                   exact_log10p=float("inf"))
    # the default level (5), 2 and every SNP; the level applies to the smallest P over the two
    # traits times 2
    for level, thr in ((0.0, 5.0 + np.log10(2)), (2.0, 2.0 + np.log10(2)), (-1.0, -1.0)):
        out = csa.pcg_projected_scan(op, r, keep, Y, X, 0.4, 0.6, cal, snps=snps, rel_tol=1e-6, max_iter=300,
                                     exact_log10p=level)
        assert out["max_rel_residual"] <= 1e-6 and out["exact_log10p"] == thr
        check_exact_stage(out, calib, K, X, Y, Gm, okm, 0.4, 0.6, thr)
        assert out["n_exact"] >= 8
    op.close()
    r.close()


def test_pcg_refines_below_operator_floor(csa, data):
    """Tolerances under 1e-5 continue with 31-bit operands: the exact residual (dense FP64 K)
    meets rel_tol, and the operator is back at its own precision afterwards (bitwise apply)."""
    N, M, body, keep, loco = data["N"], data["M"], data["body"], data["keep"], data["loco"]
    r = csa.PlinkReader.from_body(body, N, M)
    op = csa.PlinkKinshipOperator(r, keep, loco)
    n = keep.size
    K = oracle.dense_kinship(body, N, M, keep, loco)
    rng = np.random.default_rng(9)
    B = rng.standard_normal((n, 5))
    probe = rng.standard_normal((n, 3))
    before = op.apply(probe)
    for tol in (1e-6, 1e-7):
        dX = csa.DeviceArray(n, 5)
        it, res = csa.pcg_solve(op, 0.7, 0.3, csa.DeviceArray.from_host(B), dX, rel_tol=tol, max_iter=400)
        X = dX.download()
        exact = np.linalg.norm(B - (0.7 * K @ X + 0.3 * X), axis=0) / np.linalg.norm(B, axis=0)
        print(f"tol {tol:.0e}: {it} iterations, reported {res.max():.2e}, exact residual {exact.max():.2e}",
              flush=True)
        assert res.max() <= tol and exact.max() <= 1.5 * tol
    assert np.array_equal(op.apply(probe), before)
    op.close()
