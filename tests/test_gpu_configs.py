"""GPU tests at the sizes of the benchmark workloads (the configurations the oracle cannot run in
full are checked through pieces it can run and through size-independent properties).

* Plain single-trait LMM, n = 2,000 x m = 20,000 (p ~ U(0.05, 0.5), 1 % missing, intercept + 3
  covariates, 100 causal SNPs, h2 = 0.5, seed 20261203, three LOCO chromosomes) through
  cs_gwas_dense_run against the oracle's full restatement of the dense path: GRM bit-exact,
  variance components 1e-6 relative, |d log10 P| <= 1e-4, |d beta| <= 1e-5 SE.
* GRM at n = 50,000 x m = 500,000 (p ~ U(0.01, 0.5), 0.5 % missing, seed 20261204): a 256 x 256
  tile of the exact GRM and its whole diagonal bit-exact against the oracle's int128 entries over
  all SNPs; the same tile of the standardised GRM within 8e-6 of the largest diagonal entry of
  the FP64 value.
* LOCO PCG at n = 200,000 x m = 600,000 (22 chromosomes by GRCh38 length, 16 replicates): every
  chromosome's solve has a true relative residual <= 1e-6 recomputed with an independent INT8_L4
  apply; on three chromosomes 64 sampled SNPs through the exact stage agree with a host FP64 GLS
  built from separately solved V^-1 [Y, X, G] (|d log10 P| <= 1e-4); the low-rank null model on
  two chromosomes gives h2 within [0.25, 0.55] (true 0.4).
* Multi-trait LMM at n = 50,000 x m = 300,000 (4 traits, intercept + 2 covariates, standardised
  GRM): the prepare takes the 64-bit eigen path (n > 32766).  Sampled eigenpairs of the
  projected GRM satisfy P_X G w = lambda w (P_X the projection orthogonal to X) within
  1e-10 ||G||, the sampled W columns are orthonormal and orthogonal to X within 1e-10, AI-REML
  converges, and 16 scanned SNPs agree with an FP64 host recomputation from the downloaded W
  (|d log10 P| <= 1e-3).
* Serving over n = 100,000 x m = 1,000,000 resident genotypes: two requests (all 1M SNPs) give
  bitwise equal results alone and mixed; a small-cohort request over a 4,000-SNP kinship and a
  1,000-SNP scan slice agrees with the oracle's dense restatement (|d log10 P| <= 1e-3).
"""
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
            f.write(f"F{i:05d} I{i:05d} 0 0 1 -9\n")
    with open(prefix + ".bim", "w") as f:
        for j in range(M):
            f.write(f"{chroms[j]}\trs{j:06d}\t0\t{1000 + 10 * j}\tA\tC\n")


def test_c1_dense_gwas_full_size(csa, tmp_path):
    N, M, seed = 2000, 20000, 20261203
    print("C1: n=2000 m=20000 dense GWAS", flush=True)
    body = oracle.gen_hwe_bed(seed, N, M, 0.05, 0.5, 0.01)
    chroms = ["1"] * 8000 + ["2"] * 7000 + ["3"] * 5000
    rng = np.random.default_rng(seed)
    D = oracle.decode_numpy(body, N, M)
    causal = rng.choice(M, 100, replace=False)
    Zc = D[:, causal]
    Zc = np.where(np.isnan(Zc), np.nanmean(Zc, axis=0), Zc)
    Zc = (Zc - Zc.mean(0)) / np.maximum(Zc.std(0), 1e-12)
    g = Zc @ rng.normal(0.0, np.sqrt(0.5 / 100), 100)
    cov = rng.standard_normal((N, 3))
    y = cov @ np.array([0.2, -0.1, 0.3]) + 1.0 + g + rng.normal(0.0, np.sqrt(0.5), N)
    prefix = str(tmp_path / "c1")
    write_plink(prefix, body, N, M, chroms)
    out = str(tmp_path / "c1run")
    sm = csa.gwas_dense_run(prefix, y, out, cov=cov)
    models, rows = oracle.gwas_dense(body, N, M, chroms, y, cov)
    assert [m[0] for m in sm["models"]] == ["1", "2", "3"]
    for (lg, vg, ve), (lo, vgo, veo) in zip(sm["models"], models):
        assert lg == lo and abs(vg - vgo) <= 1e-6 * max(vgo, 1e-3 * veo) and abs(ve - veo) <= 1e-6 * veo
    lines = open(out + ".assoc.txt").read().splitlines()[1:]
    rg = [ln.split("\t") for ln in lines]
    assert len(rg) == len(rows) == sm["n_tested"]
    worst = 0.0
    for r, (j, freq, beta, se, p, l10) in zip(rg, rows):
        assert r[1] == f"rs{j:06d}" and r[5] == f"{freq:.6g}"
        assert abs(float(r[7]) - beta) <= 1e-5 * se + 2e-5 * abs(beta)
        d = abs(float(r[10]) - l10)
        worst = max(worst, d)
        assert d <= 1e-4 + 1e-5 * abs(l10)
    print(f"C1: {len(rows)} SNPs, max |d log10 P| {worst:.2e}, VCs {sm['models']}", flush=True)
    # the exact GRM of the same genotypes, bit for bit
    reader = csa.PlinkReader.from_body(body, N, M)
    raw_o, nf_o, s_o, _ = oracle.grm_exact_centred(body, N, M)
    nt = N * (N + 1) // 2
    d_tri = csa.DeviceArray((nt + 1) // 2)
    s2pq, nv = csa.grm_build_device(reader, d_tri.ptr)
    t = d_tri.download().reshape(-1, order="F").view(np.float32)[:nt]
    assert s2pq == s_o
    il = np.tril_indices(N)
    assert np.array_equal(t, nf_o[il])
    d_tri.close()
    reader.close()


def test_c2_grm_tile_and_diagonal(csa):
    N, M = 50_000, 500_000
    print("C2: n=50000 m=500000 exact and standardised GRM", flush=True)
    reader = csa.PlinkReader.synthetic(20261204, N, M, 0.01, 0.5, 0.005)
    body = reader.download()
    nt = N * (N + 1) // 2
    d_tri = csa.DeviceArray((nt + 1) // 2)
    s2pq, nv = csa.grm_build_device(reader, d_tri.ptr, alg=1)
    tri = d_tri.download().reshape(-1, order="F").view(np.float32)[:nt]
    rows = np.arange(49_000, 49_256, dtype=np.int32)
    cols = np.arange(1_000, 1_256, dtype=np.int32)
    _, nf_o, s_o, nv_o = oracle.grm_exact_tile(body, N, M, rows, cols)
    assert s2pq == s_o and nv == nv_o
    idx = rows[:, None].astype(np.int64) * (rows[:, None] + 1) // 2 + cols[None, :]
    assert np.array_equal(tri[idx], nf_o)
    _, dg_o = oracle.grm_exact_diag(body, N, M)
    di = np.arange(N, dtype=np.int64)
    assert np.array_equal(tri[di * (di + 1) // 2 + di], dg_o)
    print("C2: exact tile and diagonal bit-exact", flush=True)
    _, nv0 = csa.grm_build_device(reader, d_tri.ptr, alg=0)
    tri0 = d_tri.download().reshape(-1, order="F").view(np.float32)[:nt]
    G0, nv0_o = oracle.grm_std_tile(body, N, M, rows, cols)
    dmax = float(np.max(tri0[di * (di + 1) // 2 + di]))
    err = float(np.max(np.abs(tri0[idx].astype(np.float64) - G0)))
    print(f"C2: standardised tile max |dG| {err:.2e} (diagonal max {dmax:.3f})", flush=True)
    assert nv0 == nv0_o and err <= 8e-6 * dmax
    d_tri.close()
    reader.close()


def test_c3_loco_solves_and_scan(csa):
    import bench
    n, m, T = 200_000, 600_000, 16
    print("C3: n=200000 m=600000 LOCO", flush=True)
    ctx = bench.c3_setup(csa, n, m, T, 20261205)
    reader, bounds, X, Y, grm_ok, rng = ctx["reader"], ctx["bounds"], ctx["X"], ctx["Y"], ctx["grm_ok"], ctx["rng"]
    B = np.asfortranarray(np.column_stack([Y, X.astype(np.float64)]))
    k = B.shape[1]
    dB = csa.DeviceArray.from_host(B)
    dZ = csa.DeviceArray(n, k)
    dK = csa.DeviceArray(n, k)
    bn = np.linalg.norm(B, axis=0)
    idx = np.arange(m)
    worst = 0.0
    for c in range(22):
        lo, hi = bounds[c], bounds[c + 1]
        loco = idx[grm_ok & ((idx < lo) | (idx >= hi))].astype(np.int32)
        op = csa.PlinkKinshipOperator(reader, None, loco)
        it, res = csa.pcg_solve(op, 0.4, 0.6, dB, dZ, rel_tol=1e-6, max_iter=200)
        op.set_precision(csa.CS_PREC_INT8_L4)
        op.apply_device(dZ.ptr, n, k, dK.ptr, n)
        op.set_precision(csa.CS_PREC_INT8_L3)
        Zs = dZ.download()
        r = np.linalg.norm(B - (0.4 * dK.download() + 0.6 * Zs), axis=0) / bn
        worst = max(worst, float(r.max()))
        assert res.max() <= 1e-6 and r.max() <= 1.05e-6, (c, it, res.max(), r.max())
        if c in (0, 11, 21):
            snps = np.sort(rng.choice(np.arange(lo, hi), 64, replace=False)).astype(np.int32)
            cal = np.sort(rng.choice(np.arange(lo, hi), 16, replace=False)).astype(np.int32)
            out = csa.pcg_projected_scan(op, reader, None, Y, X, 0.4, 0.6, cal, snps=snps, rel_tol=1e-6,
                                         exact_log10p=-1.0)
            assert out["n_exact"] == int(out["passes"].sum())
            dG = csa.DeviceArray(n, snps.size)
            ps, _ = csa.markers_device(reader, None, snps, dG)
            G = dG.download()
            # host FP64 GLS from a separate tight solve of V^-1 [Y, X, G]
            dB2 = csa.DeviceArray.from_host(np.column_stack([B, G]))
            dZ2 = csa.DeviceArray(n, k + snps.size)
            csa.pcg_solve(op, 0.4, 0.6, dB2, dZ2, rel_tol=1e-8, max_iter=300)
            Z2 = dZ2.download()
            Xd = X.astype(np.float64)
            ZY, ZX, ZG = Z2[:, :T], Z2[:, T:k], Z2[:, k:]
            C = np.linalg.inv(Xd.T @ ZX)
            PY = ZY - ZX @ (C @ (Xd.T @ ZY))
            q = Xd.T @ ZG
            var = np.einsum("ij,ij->j", G, ZG) - np.einsum("aj,ab,bj->j", q, C, q)
            for t in range(T):
                st = oracle._stats_from(G.T @ PY[:, t], var)
                mm = ps & st["tested"]
                assert np.max(np.abs(out["log10p"][mm, t] - st["log10p"][mm])) <= 1e-4
                assert np.max(np.abs(out["beta"][mm, t] - st["beta"][mm]) / st["se"][mm]) <= 1e-4
            for a in (dG, dB2, dZ2):
                a.close()
        if c in (0, 21):
            lr = csa.LowRankLmm(op, np.ascontiguousarray(Y[:, 0]), X, rank=256, oversample=10)
            d = lr.diag
            h2 = d["sigma2_g"] / (d["sigma2_g"] + d["sigma2_e"])
            print(f"C3: chromosome {c + 1} low-rank h2 {h2:.3f}", flush=True)
            assert 0.25 <= h2 <= 0.55
            lr.close()
        op.close()
    print(f"C3: largest true relative residual over 22 solves {worst:.3e}", flush=True)
    for a in (dB, dZ, dK):
        a.close()
    reader.close()


def test_c4_serving_full_size(csa):
    n, m = 100_000, 1_000_000
    print("C4: n=100000 m=1000000 serving", flush=True)
    reader = csa.PlinkReader.synthetic(20261206, n, m, 0.05, 0.5, 0.005)
    rng = np.random.default_rng(31)

    def request(nc, T, c, h2):
        keep = np.sort(rng.choice(n, nc, replace=False)).astype(np.int32)
        X = np.column_stack([np.ones(nc)] + [rng.standard_normal(nc) for _ in range(c - 1)]).astype(np.float32)
        return dict(Y=rng.standard_normal((nc, T)), X=X, sigma2_g=h2, sigma2_e=1.0 - h2,
                    cal_snps=np.sort(rng.choice(m, 16, replace=False)).astype(np.int32), keep=keep)

    qa, qb = request(20_000, 3, 2, 0.3), request(50_000, 5, 4, 0.6)

    def same(a, b):
        for key in ("beta", "se", "p", "log10p", "freq"):
            assert np.array_equal(a[key], b[key], equal_nan=True), key
        assert a["stats"]["iterations"] == b["stats"]["iterations"]

    alone = []
    for q in (qa, qb):
        srv = csa.ScanServer(reader, max_columns=128)
        rid = srv.submit(**q)
        srv.run_until_idle()
        alone.append(srv.result(rid))
        srv.close()
    srv = csa.ScanServer(reader, max_columns=128)
    ia = srv.submit(**qa)
    srv.step()
    srv.step()
    ib = srv.submit(**qb)
    srv.run_until_idle()
    same(srv.result(ia), alone[0])
    same(srv.result(ib), alone[1])
    srv.close()
    print("C4: two full-size requests bitwise equal alone and mixed", flush=True)
    # a small cohort over a 4,000-SNP kinship and a 1,000-SNP scan slice against the oracle
    kin = np.sort(rng.choice(m, 4000, replace=False)).astype(np.int32)
    rest = np.setdiff1d(np.arange(m), kin)
    scan = np.sort(rng.choice(rest, 1000, replace=False)).astype(np.int32)
    sl = np.concatenate([kin, scan])
    body = reader.download_snps(sl)
    q = request(2000, 2, 2, 0.5)
    q["cal_snps"] = scan[:16]
    srv = csa.ScanServer(reader, variants=kin, max_columns=64)
    rid = srv.submit(**dict(q, snps=scan, rel_tol=1e-9, max_iter=400, exact_log10p=2.0))
    srv.run_until_idle()
    got = srv.result(rid)
    srv.close()
    kin_l = np.arange(4000, dtype=np.int32)
    scan_l = np.arange(4000, 5000, dtype=np.int32)
    st, gamma, _ = oracle.pcg_projected_scan(body, n, sl.size, q["keep"], kin_l, q["Y"], q["X"], 0.5, 0.5,
                                             scan_l[:16], exact_log10p=2.0)
    assert abs(got["stats"]["gamma"] / gamma - 1) <= 1e-4
    for t in range(2):
        lo = st[t]["log10p"][scan_l]
        mm = ~np.isnan(lo)
        assert np.array_equal(np.isnan(got["log10p"][:, t]), ~mm)
        assert np.max(np.abs(got["log10p"][mm, t] - lo[mm])) <= 1e-3
    print(f"C4: small cohort vs oracle ok ({got['stats']['n_exact']} exact SNPs)", flush=True)
    reader.close()


def test_c5_multitrait_full_size(csa):
    n, m, d = 50_000, 300_000, 4
    print("C5: n=50000 m=300000 d=4 multi-trait, 64-bit eigen path", flush=True)
    rng = np.random.default_rng(20261207)
    reader = csa.PlinkReader.synthetic(20261207, n, m, 0.05, 0.5, 0.01)
    causal = np.sort(rng.choice(m, 200, replace=False)).astype(np.int32)
    dZ = csa.DeviceArray(n, causal.size)
    csa.markers_device(reader, None, causal, dZ, min_maf=0.0)
    Zc = dZ.download()
    dZ.close()
    Zc /= np.maximum(Zc.std(axis=0), 1e-12)
    X = np.column_stack([np.ones(n), rng.standard_normal((n, 2))]).astype(np.float32)
    Y = Zc @ rng.normal(0, 0.08, (causal.size, d)) + rng.standard_normal((n, d))
    raw = csa.DeviceArray(n, n)
    S, _, nv = csa.grm_build_raw(reader, raw, min_maf=0.01, alg=0)
    G = csa.DeviceArray(n, n)
    csa.grm_loco_subset_scaled(raw, None, n, None, nv, 1.0, S, 0.0, G)
    raw.close()
    mv = csa.MvLmm(Y, X)
    mv.prepare(G)
    print("C5: prepare done", flush=True)
    Gh = G.download()
    G.close()
    D, Uty, W = mv.eigen(with_w=True)
    ne = D.size
    assert ne == n - 3 and np.all(np.diff(D) >= 0)
    idx = np.unique(np.concatenate([[0, 1, ne // 2, ne - 2, ne - 1], rng.choice(ne, 5, replace=False)]))
    Wi = np.ascontiguousarray(W[:, idx])
    Xd = X.astype(np.float64)
    gnorm = np.max(D)
    GW = Gh @ Wi
    del Gh
    GW -= Xd @ np.linalg.lstsq(Xd, GW, rcond=None)[0]
    res = np.max(np.abs(GW - Wi * D[idx][None, :])) / gnorm
    orth = np.max(np.abs(Wi.T @ Wi - np.eye(idx.size)))
    xw = np.max(np.abs(Xd.T @ Wi)) / np.sqrt(n)
    print(f"C5: eigenpair residual {res:.2e}, orthonormality {orth:.2e}, X^T w {xw:.2e}", flush=True)
    assert res <= 1e-10 and orth <= 1e-10 and xw <= 1e-10
    conv, em_it, ai_it = mv.runNullModel("aireml")
    Vg, Ve, _, _, _ = mv.variance()
    print(f"C5: AI-REML converged={conv} (EM {em_it}, AI {ai_it})", flush=True)
    assert conv and np.all(np.isfinite(Vg)) and np.all(np.isfinite(Ve))
    snps = np.sort(rng.choice(m, 2000, replace=False)).astype(np.int32)
    res_scan = mv.scan(reader, None, snps)
    chk = np.sort(rng.choice(snps.size, 16, replace=False))
    dG = csa.DeviceArray(n, chk.size)
    csa.markers_device(reader, None, snps[chk], dG, min_maf=0.0)
    Gm = dG.download()
    dG.close()
    Sr = W.T @ Gm
    del W
    Sinv = np.linalg.inv(D[:, None, None] * Vg + Ve)
    H = np.einsum("is,iab->sab", Sr * Sr, Sinv)
    r = Sr.T @ np.einsum("iab,ib->ia", Sinv, Uty)
    beta = np.linalg.solve(H, r[:, :, None])[:, :, 0]
    wald = np.einsum("sa,sab,sb->s", beta, H, beta)
    from scipy.special import chdtrc
    lp = -np.log10(chdtrc(d, wald))
    lp_gpu = -np.log10(res_scan["p"][chk])
    dl = np.max(np.abs(lp_gpu - lp))
    print(f"C5: 16 SNPs vs FP64 host recomputation |d log10 P| {dl:.2e}", flush=True)
    assert dl <= 1e-3
    mv.close()
    reader.close()
