"""Pins the CPU oracle against known answers, closed forms, brute force and invariants.

Every check here is independent of the formula it checks: decodes come from a separate numpy
unpack, dense products from numpy, p-values from scipy's closed form, REML from a brute-force
grid.  These are CPU-only tests (no GPU).
"""
import numpy as np
import pytest
import scipy.special

import oracle
from kinship_batch_fixture import MT19937_64


def dense_z(bed, N, M, means, inv, keep, variants):
    """The reference test's dense oracle (test_kinship_batch.cpp:61-71) from a numpy decode."""
    D = oracle.decode_numpy(bed, N, M)
    keep = np.arange(N) if keep is None else keep
    variants = np.arange(M) if variants is None else variants
    Z = (D[np.ix_(keep, variants)] - means[variants]) * inv[variants]
    return np.where(np.isnan(Z), 0.0, Z)


def rel(a, b):
    return np.max(np.abs(a - b)) / max(1e-30, np.max(np.abs(b)))


def test_mt19937_64_standard_known_answer():
    # C++ standard [rand.predef]: the 10000th output of a default-constructed mt19937_64.
    r = MT19937_64(5489)
    for _ in range(9999):
        r()
    assert r() == 9981545732273789042


def test_bed_magic_rejected(tmp_path):
    p = tmp_path / "x"
    (p.with_suffix(".fam")).write_text("a a 0 0 0 1\n")
    (p.with_suffix(".bim")).write_text("1 s 0 1 A C\n")
    (p.with_suffix(".bed")).write_bytes(bytes([0x6C, 0x1B, 0x00, 0]))
    with pytest.raises(RuntimeError, match="SNP-major"):
        oracle.read_bed(str(p))
    (p.with_suffix(".bed")).write_bytes(bytes([0x6C, 0x1C, 0x01, 0]))
    with pytest.raises(RuntimeError, match="magic"):
        oracle.read_bed(str(p))


@pytest.mark.parametrize("name", ["synth", "gemma", "ebfmt"])
def test_stats_closed_form_reference_beds(ref_datasets, name):
    bed, N, M, bim, fam = ref_datasets[name]
    means, inv = oracle.stats(bed, N, M)
    D = oracle.decode_numpy(bed, N, M)
    obs = ~np.isnan(D)
    cnt = obs.sum(0)
    s = np.nansum(D, 0)
    with np.errstate(invalid="ignore", divide="ignore"):
        mu = np.where(cnt > 0, s / cnt, 0.0)
    p = mu / 2
    var = 2 * p * (1 - p)
    exp_inv = np.where((cnt > 0) & (var >= 1e-9), 1 / np.sqrt(np.maximum(var, 1e-300)), 0.0)
    assert np.array_equal(means, mu)
    assert np.array_equal(inv, exp_inv)
    cc = oracle.class_counts(bed, N, M)
    codes = np.array([[3, 2, 0][int(v)] if not np.isnan(v) else 1 for v in D[:, 0]])
    assert cc[0].tolist() == [int((codes == c).sum()) for c in range(4)]
    assert np.array_equal(cc.sum(1), np.full(M, N))
    assert np.array_equal(cc[:, 0] * 2 + cc[:, 2], s.astype(np.int64))


def test_stats_kinship_fixture_special_snps(kb_fixture):
    bed, N, M, _ = kb_fixture
    means, inv = oracle.stats(bed, N, M)
    for j in range(M):
        if j % 31 == 0:
            assert means[j] == 0.0 and inv[j] == 0.0     # all code 3: dosage 0, monomorphic
        elif j % 37 == 0:
            assert means[j] == 0.0 and inv[j] == 0.0     # all missing
    assert np.count_nonzero(inv) > 900
    # partial last byte: padding codes (random) must not change the stats
    cc = oracle.class_counts(bed, N, M)
    assert np.array_equal(cc.sum(1), np.full(M, N))


def test_kinship_apply_matches_dense_all_selections(kb_fixture):
    bed, N, M, sels = kb_fixture
    means, inv = oracle.stats(bed, N, M)
    rng = np.random.default_rng(1)
    for keep, variants in sels:
        Z = dense_z(bed, N, M, means, inv, keep, variants)
        Mv = Z.shape[1]
        K = Z @ Z.T / Mv
        tr = oracle.kinship_trace(bed, N, means, inv, keep, variants)
        assert abs(tr - np.trace(K)) < 1e-10
        assert np.max(np.abs(oracle.kinship_diag(bed, N, means, inv, keep, variants)
                             - np.diag(K))) < 1e-12
        for width in (1, 7, 41):
            X = rng.standard_normal((Z.shape[0], width))
            Y = oracle.kinship_apply(bed, N, means, inv, keep, variants, X)
            assert rel(Y, K @ X) < 1e-10
            # the reference's 512-SNP block form (the CPU baseline), one and several threads
            for th in (1, 3):
                assert rel(oracle.kinship_apply_blocked(bed, N, means, inv, keep, variants, X, nthreads=th),
                           K @ X) < 1e-10
        with pytest.raises(RuntimeError):
            oracle.kinship_apply(bed, N, means, inv, keep, variants, np.ones((Z.shape[0] + 1, 1)))


def test_kinship_invariants(kb_fixture):
    bed, N, M, sels = kb_fixture
    means, inv = oracle.stats(bed, N, M)
    rng = np.random.default_rng(2)
    v, w = rng.standard_normal((2, N))
    Kv = oracle.kinship_apply(bed, N, means, inv, None, None, v)
    Kw = oracle.kinship_apply(bed, N, means, inv, None, None, w)
    assert abs(w @ Kv - v @ Kw) < 1e-10 * np.abs(w @ Kv)          # symmetry
    Kvw = oracle.kinship_apply(bed, N, means, inv, None, None, 2 * v - 3 * w)
    assert rel(Kvw, 2 * Kv - 3 * Kw) < 1e-12                      # linearity
    # principal submatrix: K_S v_S == (K_full P_S v)_S
    keep = np.array([3, 9, 0, 30, 12], dtype=np.int32)
    vs = rng.standard_normal(keep.size)
    full = np.zeros(N)
    np.add.at(full, keep, vs)
    sub = oracle.kinship_apply(bed, N, means, inv, keep, None, vs)
    assert rel(sub, oracle.kinship_apply(bed, N, means, inv, None, None, full)[keep]) < 1e-12
    # LOCO additivity: M K_full = M_loco K_loco + M_c K_c
    chrom = np.arange(M) % 3
    loco = np.flatnonzero(chrom != 1).astype(np.int32)
    only = np.flatnonzero(chrom == 1).astype(np.int32)
    a = M * Kv
    b = loco.size * oracle.kinship_apply(bed, N, means, inv, None, loco, v) + \
        only.size * oracle.kinship_apply(bed, N, means, inv, None, only, v)
    assert rel(a, b) < 1e-12


def test_grm_alg1_alg0_dense(ref_datasets):
    bed, N, M, _, _ = ref_datasets["gemma"]
    D = oracle.decode_numpy(bed, N, M)
    for alg in (1, 0):
        G, s2pq, nv = oracle.grm(bed, N, M, alg=alg, min_maf=0.01)
        valid, mean_f, scale_f, pq2_f, s2, nv2 = oracle.grm_snp_params(bed, N, M, alg, 0.01)
        obs = ~np.isnan(D)
        f = np.nansum(D, 0) / (2 * obs.sum(0))
        maf = np.minimum(f, 1 - f)
        assert np.array_equal(valid, (maf >= np.float32(0.01)) & (np.abs(maf) >= 1e-8))
        assert np.array_equal(mean_f, np.where(valid, (2 * f).astype(np.float32), 0))
        Z = (D[:, valid] - mean_f[valid].astype(np.float64)) * scale_f[valid].astype(np.float64)
        Z = np.where(np.isnan(Z), 0.0, Z)
        ref = Z @ Z.T / (nv if alg == 0 else s2pq)
        assert nv == valid.sum() and s2 == s2pq
        assert rel(G, ref) < 1e-13
        if alg == 1:
            assert abs(s2pq - np.sum(pq2_f[valid].astype(np.float64))) < 1e-9 * s2pq


def test_grm_alg0_mean_diag_no_missing():
    N, M = 300, 4000
    bed = oracle.gen_hwe_bed(7, N, M, 0.05, 0.5, 0.0)
    G, _, _ = oracle.grm(bed, N, M, alg=0, min_maf=0.01)
    assert abs(np.mean(np.diag(G)) - 1.0) < 0.02


def test_grm_exact_matches_python_integers():
    N, M = 9, 57
    bed = oracle.gen_hwe_bed(11, N, M, 0.05, 0.5, 0.08)
    raw, nf, s2pq, nv = oracle.grm_exact_centred(bed, N, M, min_maf=0.01)
    valid, mean_f, _, _, _, _ = oracle.grm_snp_params(bed, N, M, 1, 0.01)
    D = oracle.decode_numpy(bed, N, M)
    A = [int(np.ldexp(np.float64(m), 29)) for m in mean_f]
    from fractions import Fraction
    for i in range(N):
        for k in range(N):
            tot = 0
            for j in range(M):
                if not valid[j] or np.isnan(D[i, j]) or np.isnan(D[k, j]):
                    continue
                tot += ((int(D[i, j]) << 29) - A[j]) * ((int(D[k, j]) << 29) - A[j])
            exact = Fraction(tot, 1 << 58)
            assert raw[i, k] == float(exact)                           # correctly rounded
            q = exact / Fraction(s2pq)
            f32 = np.float32(float(q))
            # correct rounding to float: the nearest float to q (ties even) -- check neighbours
            lo, hi = np.nextafter(f32, np.float32(-np.inf)), np.nextafter(f32, np.float32(np.inf))
            d = abs(Fraction(float(f32)) - q)
            assert d <= abs(Fraction(float(lo)) - q) and d <= abs(Fraction(float(hi)) - q)
            assert nf[i, k] == f32


def test_grm_exact_tile_and_diag_any_mean():
    # rare SNPs (float means below 2^-6 that are not multiples of 2^-29): the tile and diagonal
    # oracles against Python rationals with the float mean taken exactly; at the default MAF the
    # tile equals the full exact matrix
    N, M = 300, 40
    bed = oracle.gen_hwe_bed(12, N, M, 0.0005, 0.02, 0.03)
    valid, mean_f, _, _, _, _ = oracle.grm_snp_params(bed, N, M, 1, 0.0001)
    assert np.any(valid & (mean_f < 2 ** -6) & (np.ldexp(mean_f.astype(np.float64), 29) % 1 != 0))
    rows = np.array([3, 0, 299, 17, 3], dtype=np.int32)
    cols = np.array([5, 3, 120], dtype=np.int32)
    raw, nf, s2pq, nv = oracle.grm_exact_tile(bed, N, M, rows, cols, min_maf=0.0001)
    draw, dnf = oracle.grm_exact_diag(bed, N, M, min_maf=0.0001)
    D = oracle.decode_numpy(bed, N, M)
    from fractions import Fraction
    mu = [Fraction(float(m)) for m in mean_f]

    def exact(i, k):
        tot = Fraction(0)
        for j in range(M):
            if not valid[j] or np.isnan(D[i, j]) or np.isnan(D[k, j]):
                continue
            tot += (Fraction(int(D[i, j])) - mu[j]) * (Fraction(int(D[k, j])) - mu[j])
        return tot

    def check(r, f, i, k):
        e = exact(i, k)
        assert r == float(e)
        q = e / Fraction(s2pq)
        f32 = np.float32(float(q))
        lo, hi = np.nextafter(f32, np.float32(-np.inf)), np.nextafter(f32, np.float32(np.inf))
        d = abs(Fraction(float(f32)) - q)
        assert d <= abs(Fraction(float(lo)) - q) and d <= abs(Fraction(float(hi)) - q)
        assert f == f32

    for a, i in enumerate(rows):
        for b, k in enumerate(cols):
            check(raw[a, b], nf[a, b], i, k)
    for i in (0, 3, 150, 299):
        check(draw[i], dnf[i], i, i)
    for maf in (0.0001, 0.02):
        full_raw, full_nf, s_full, _ = oracle.grm_exact_centred(bed, N, M, min_maf=maf)
        t_raw, t_nf, s_t, _ = oracle.grm_exact_tile(bed, N, M, rows, cols, min_maf=maf)
        assert s_t == s_full and np.array_equal(t_raw, full_raw[np.ix_(rows, cols)])
        assert np.array_equal(t_nf, full_nf[np.ix_(rows, cols)])
        d_raw, d_nf = oracle.grm_exact_diag(bed, N, M, min_maf=maf)
        assert np.array_equal(d_raw, np.diag(full_raw)) and np.array_equal(d_nf, np.diag(full_nf))


def test_grm_std_tile_matches_full(ref_datasets):
    bed, N, M, _, _ = ref_datasets["gemma"]
    G, _, _ = oracle.grm(bed, N, M, alg=0, min_maf=0.01)
    rows = np.array([7, 0, N - 1], dtype=np.int32)
    cols = np.array([N - 1, 3, 7, 9], dtype=np.int32)
    T, nv = oracle.grm_std_tile(bed, N, M, rows, cols)
    assert rel(T, G[np.ix_(rows, cols)]) < 1e-12


def test_grm_exact_vs_fp64(ref_datasets):
    bed, N, M, _, _ = ref_datasets["gemma"]
    raw, nf, s2pq, nv = oracle.grm_exact_centred(bed, N, M, min_maf=0.01)
    G, s2, _ = oracle.grm(bed, N, M, alg=1, min_maf=0.01)
    assert s2 == s2pq
    assert rel(raw / s2pq, G) < 1e-12
    assert np.max(np.abs(nf.astype(np.float64) - G) / (np.abs(G) + 1e-30)) < 2 ** -23


def test_scan_sums_match_numpy(ref_datasets):
    bed, N, M, _, _ = ref_datasets["synth"]
    rng = np.random.default_rng(3)
    keep = np.sort(rng.choice(N, 45, replace=False)).astype(np.int32)
    R = rng.standard_normal((keep.size, 3))
    freq, nobs, ps, dot, gtg = oracle.scan_sums(bed, N, M, keep, R, min_maf=0.05)
    D = oracle.decode_numpy(bed, N, M)[keep]
    obs = ~np.isnan(D)
    mu = np.nansum(D, 0) / obs.sum(0)
    G = np.where(obs, D, mu) - mu
    f = np.nansum(D, 0) / (2 * obs.sum(0))
    assert np.allclose(freq, f, rtol=0, atol=1e-15)
    assert np.array_equal(nobs, obs.sum(0))
    assert np.array_equal(ps, (obs.sum(0) >= 3) & (np.minimum(f, 1 - f) >= 0.05))
    assert rel(dot, G.T @ R) < 1e-13
    assert rel(gtg, np.sum(G * G, 0)) < 1e-13


def test_chisq_prob_closed_form():
    stat = np.concatenate([np.linspace(0.01, 5, 50), np.linspace(5, 200, 50)])
    p = oracle.chisq_prob(stat, 1)
    exact = scipy.special.chdtrc(1, stat)
    assert np.all(np.abs(p - exact) <= 1e-6 * exact + 1e-300)
    assert oracle.chisq_prob(np.array([0.0, -1.0]), 1).tolist() == [1.0, 1.0]
    # deep tail: continued fraction keeps |d(-log10 p)| tiny
    assert np.max(np.abs(np.log10(p[p > 0]) - np.log10(exact[p > 0]))) < 1e-6
    p4 = oracle.chisq_prob(np.array([1.0, 9.0, 30.0]), 4)
    assert np.allclose(p4, scipy.special.chdtrc(4, [1.0, 9.0, 30.0]), rtol=1e-6)


def _small_lmm(seed=5, N=240, M=800):
    rng = np.random.default_rng(seed)
    bed = oracle.gen_hwe_bed(seed, N, M, 0.05, 0.5, 0.01)
    G, _, _ = oracle.grm(bed, N, M, alg=1, min_maf=0.01)
    X = np.column_stack([np.ones(N), rng.standard_normal((N, 2)).astype(np.float32)])
    L = np.linalg.cholesky(G + 1e-6 * np.eye(N))
    y = X @ np.array([1.0, 0.5, -0.2]) + L @ rng.standard_normal(N) + rng.standard_normal(N)
    return bed, N, M, G, X, y


def test_solve_diagonal_brute_force_grid():
    bed, N, M, G, X, y = _small_lmm()
    D, W, Uty = oracle.eigen_vce_prepare(G, y, X)
    vg, ve, delta = oracle.solve_diagonal(D, Uty)
    # explicit REML likelihood on V = G + delta I (X projected out), brute-force grid
    def reml(dl):
        V = G + dl * np.eye(N)
        Vi = np.linalg.inv(V)
        XtVi = X.T @ Vi
        P = Vi - XtVi.T @ np.linalg.solve(XtVi @ X, XtVi)
        q = y @ P @ y
        nef = N - X.shape[1]
        return -0.5 * (np.linalg.slogdet(V)[1] + np.linalg.slogdet(XtVi @ X)[1] + nef * np.log(q / nef))
    grid = 10.0 ** np.linspace(-2, 2, 161)
    best = grid[np.argmax([reml(d) for d in grid])]
    assert abs(np.log10(best) - np.log10(delta)) <= 0.026
    # invariance to the null-space basis: a rotated basis gives the same D and statistic
    assert np.all(D >= 0)


def test_rotated_scan_equals_explicit_gls():
    bed, N, M, G, X, y = _small_lmm(seed=9, N=200, M=600)
    D, W, Uty = oracle.eigen_vce_prepare(G, y, X)
    vg, ve, delta = oracle.solve_diagonal(D, Uty)
    keep = np.arange(N, dtype=np.int32)
    Dd = oracle.decode_numpy(bed, N, M)[:, :50]
    mu = np.nanmean(Dd, 0)
    Gm = np.where(np.isnan(Dd), mu, Dd) - mu
    a = oracle.rotated_scan(W, D, Uty, vg, ve, Gm)
    b = oracle.gls_scan(vg * G + ve * np.eye(N), X, y, Gm)
    assert rel(a["score"], b["score"]) < 1e-8
    assert rel(a["var"], b["var"]) < 1e-8
    assert np.max(np.abs(a["log10p"] - b["log10p"])) < 1e-6


def test_gen_hwe_moments():
    N, M = 4000, 300
    bed = oracle.gen_hwe_bed(20261204, N, M, 0.01, 0.5, 0.005)
    D = oracle.decode_numpy(bed, N, M)
    miss = np.isnan(D).mean()
    assert abs(miss - 0.005) < 0.001
    f = np.nanmean(D, 0) / 2
    assert f.min() > 0.0 and f.max() < 0.56
    # HWE: het fraction ~ 2pq per SNP
    het = np.nanmean(D == 1, 0)
    assert np.mean(np.abs(het - 2 * f * (1 - f))) < 0.02
    # padding bits are zero
    bed2 = oracle.gen_hwe_bed(1, 5, 3, 0.1, 0.5, 0.0)
    assert np.all(bed2.reshape(3, 2)[:, 1] >> 2 == 0)


def test_loco_formula_equals_complement_grm(ref_datasets):
    # gwas_scanner.cpp:481-495: G * (S / S_loco) - raw_c / S_loco must equal the GRM of the other
    # chromosomes' SNPs, (raw - raw_c) / S_loco -- computed here from scratch on the complement.
    body, N, M, bim, _ = ref_datasets["gemma"]
    chroms = [b[0] for b in bim]
    raw, _, S, _ = oracle.grm_exact_centred(body, N, M)
    for c in sorted(set(chroms)):
        on = [j for j in range(M) if chroms[j] == c]
        off = [j for j in range(M) if chroms[j] != c]
        raw_c = oracle.grm_exact_centred(oracle.select_snps(body, N, on), N, len(on))[0]
        raw_o, _, S_o, _ = oracle.grm_exact_centred(oracle.select_snps(body, N, off), N, len(off))
        S_c = oracle.sequential_sum_2pq(body, N, M, 0.01, on)
        G_loco = (raw / S) * (S / (S - S_c)) - raw_c / (S - S_c)
        assert abs((S - S_c) - S_o) <= 1e-12 * S
        assert rel(G_loco, raw_o / S_o) < 1e-12


def test_gwas_dense_reference_readme_row_count(ref_datasets):
    # README.md:63: the bundled synth fixture (60 x 180) gives 180 association rows.
    import os
    import cosmicscanner_amd as csa
    body, N, M, bim, fam = ref_datasets["synth"]
    prefix = os.path.join(os.path.dirname(__file__), "golden", "ref_synth", "synth")
    y = csa.read_pheno(prefix[:-5] + "quantitative.pheno", fam)
    models, rows = oracle.gwas_dense(body, N, M, [b[0] for b in bim], y)
    assert len(rows) == 180
    assert [m[0] for m in models] == ["1", "2"]


def test_gwas_dense_gemma_truth(ref_datasets):
    # tests/data/gemma_accuracy/gemma_accuracy.truth.txt: simulated causal SNPs and effects; the
    # two large effects are among the top hits with the right sign, the third has the right sign.
    import os
    import cosmicscanner_amd as csa
    body, N, M, bim, fam = ref_datasets["gemma"]
    prefix = os.path.join(os.path.dirname(__file__), "golden", "ref_gemma_accuracy", "gemma_accuracy")
    y = csa.read_pheno(prefix + ".pheno", fam)
    _, rows = oracle.gwas_dense(body, N, M, [b[0] for b in bim], y)
    by_id = {bim[r[0]][1]: r for r in rows}
    truth = {}
    for ln in open(prefix + ".truth.txt").read().splitlines()[1:]:
        k, v = ln.split()
        truth[k] = float(v)
    rank = [bim[r[0]][1] for r in sorted(rows, key=lambda r: r[4])]
    for snp, b in truth.items():
        assert np.sign(by_id[snp][2]) == np.sign(b)
    assert rank.index("rs000013") < 5 and rank.index("rs000212") < 5


def test_projected_scan_reduces_to_ols_without_kinship():
    # sigma_g = 0: V = sigma_e I, so g^T P g = g~^T g~ / sigma_e exactly -> gamma = 1 / sigma_e, and
    # the calibrated scan is exactly the textbook GLS (here OLS) scan of the same markers.
    N, M = 120, 300
    body = oracle.gen_hwe_bed(9, N, M, 0.05, 0.5, 0.02)
    rng = np.random.default_rng(0)
    X = np.column_stack([np.ones(N), rng.standard_normal(N)]).astype(np.float32)
    y = rng.standard_normal(N)
    st, gamma, _ = oracle.pcg_projected_scan(body, N, M, None, None, y, X, 0.0, 0.7, np.arange(0, 40))
    assert abs(gamma * 0.7 - 1.0) < 1e-12
    G, ok = oracle.prepared_markers(body, N, M, None, np.arange(M))
    ref = oracle.gls_scan(0.7 * np.eye(N), X.astype(np.float64), y, G[:, ok])
    lp = st[0]["log10p"][ok]
    m = ~np.isnan(lp)
    assert np.array_equal(m, ref["tested"])
    assert np.max(np.abs(lp[m] - ref["log10p"][m])) < 1e-9
    assert np.max(np.abs(st[0]["beta"][ok][m] - ref["beta"][m])) < 1e-9


def test_prepared_markers_match_scan_sums():
    N, M = 77, 200
    body = oracle.gen_hwe_bed(5, N, M, 0.0, 0.5, 0.05)
    keep = np.array([3, 1, 1, 40, 76, 20, 9, 9, 9, 50, 60, 61], dtype=np.int32)
    G, ok = oracle.prepared_markers(body, N, M, keep, np.arange(M), 0.05)
    _, _, ps, _, gtg = oracle.scan_sums(body, N, M, keep, np.ones((keep.size, 1)), 0.05)
    assert np.array_equal(ok, ps.astype(bool))
    assert np.max(np.abs(np.sum(G * G, 0)[ok] - gtg[ok])) < 1e-9


def test_std_normal_known_answers():
    # tests/golden/std_normal_mt64.txt: this image's libstdc++ std::normal_distribution<double>
    # over std::mt19937_64 (the reference's sketch generator), made by make_std_normal.cpp.
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "std_normal_mt64.txt")
    blocks, cur = {}, None
    for ln in open(path).read().splitlines():
        if ln.startswith("seed"):
            cur = int(ln.split()[1])
            blocks[cur] = []
        else:
            blocks[cur].append(float(ln))
    for seed, vals in blocks.items():
        g = oracle.StdNormal(oracle.MT19937_64(seed))
        assert [g() for _ in range(len(vals))] == vals
    om = oracle.sketch_omega(42, 3, 2)
    assert om[0, 0] == blocks[42][0] and om[2, 0] == blocks[42][2] and om[0, 1] == blocks[42][3]


def test_std_shuffle_known_answers():
    # tests/golden/std_shuffle_mt64.txt: this image's libstdc++ std::shuffle over
    # std::mt19937_64 (the matrix-free driver's calibration draw), made by make_std_shuffle.cpp;
    # odd and even lengths, both seeds.
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "std_shuffle_mt64.txt")
    lines = open(path).read().splitlines()
    assert len(lines) == 28
    for head, body in zip(lines[::2], lines[1::2]):
        t = head.split()
        seed, n = int(t[1]), int(t[3])
        assert oracle.std_shuffle(range(n), seed) == [int(x) for x in body.split()]
    # a permutation of the input whatever the values
    v = list(range(100, 137))
    assert sorted(oracle.std_shuffle(v, 3)) == v


def test_lowrank_exact_when_rank_covers_kinship():
    # With K of rank < sketch size the sketch is exact: the leading eigenpairs are K's own, r0 = 0
    # beyond them, and the low-rank REML equals the dense REML on the same model.
    rng = np.random.default_rng(2)
    n, q = 60, 8
    Zr = rng.standard_normal((n, q))
    K = Zr @ Zr.T / q
    X = np.column_stack([np.ones(n), rng.standard_normal(n)]).astype(np.float32)
    y = rng.standard_normal(n) + 0.5 * Zr[:, 0]
    m = oracle.lowrank_prepare(K, y, X, rank=q, oversample=4, trace_threshold=0.0)
    w = np.linalg.eigvalsh(K)[::-1][:q]
    assert np.max(np.abs(m["lam"] - w)) < 1e-10 * w[0]
    assert abs(m["r0"]) < 1e-10
    # REML with V = sg K + se I directly (brute-force grid around the optimum)
    Xd = X.astype(np.float64)

    def reml(sg, se):
        V = sg * K + se * np.eye(n)
        Vi = np.linalg.inv(V)
        XVX = Xd.T @ Vi @ Xd
        P = Vi - Vi @ Xd @ np.linalg.solve(XVX, Xd.T @ Vi)
        return -0.5 * (np.linalg.slogdet(V)[1] + np.linalg.slogdet(XVX)[1] + y @ P @ y)
    best = reml(m["sg"], m["se"])
    for fg in (0.9, 1.1):
        for fe in (0.9, 1.1):
            assert reml(m["sg"] * fg, m["se"] * fe) <= best + 1e-9


# ------------------------------------------------------------------ multivariate (C5) oracle pins
def _mv_case(N=300, M=1200, seed=3, d=1):
    bed = oracle.gen_hwe_bed(seed, N, M, 0.05, 0.5, 0.01)
    raw, _, S, _ = oracle.grm_exact_centred(bed, N, M)
    rng = np.random.default_rng(seed)
    X = np.column_stack([np.ones(N), rng.standard_normal(N)]).astype(np.float32).astype(np.float64)
    Dm = oracle.decode_numpy(bed, N, M)
    Dm = np.where(np.isnan(Dm), np.nanmean(Dm, 0), Dm)
    g = (Dm[:, :60] - Dm[:, :60].mean(0)) @ rng.normal(0, 0.15, (60, d))
    Y = X[:, 1:2] * 0.3 + g + rng.normal(0, 1, (N, d))
    return bed, N, M, raw / S, X, Y


def test_mv_oracle_reduces_to_univariate():
    """d = 1: AI-REML reaches the univariate REML optimum of solveDiagonal, and the multivariate
    Wald, beta and Var(beta) equal the rotated scan's chi^2, beta and SE^2."""
    bed, N, M, G, X, Y = _mv_case()
    D, W, Uty, Vg0, Ve0 = oracle.mv_prepare(G, Y, X)
    Vg, Ve, conv, _, ai_it = oracle.mv_null(D, Uty, Vg0, Ve0, "aireml")
    D1, W1, U1 = oracle.eigen_vce_prepare(G, Y[:, 0], X)
    vg, ve, _ = oracle.solve_diagonal(D1, U1)
    assert conv and ai_it > 0
    assert abs(Vg[0, 0] - vg) <= 2e-5 * vg and abs(Ve[0, 0] - ve) <= 2e-5 * ve
    o = oracle.mv_scan(bed, N, M, None, W, D, Uty, np.array([[0.3]]), np.array([[0.8]]), np.arange(80))
    Dg = oracle.decode_numpy(bed, N, M)[:, :80]
    obs = ~np.isnan(Dg)
    Gm = np.where(obs, Dg - np.nansum(Dg, 0) / obs.sum(0), 0.0)
    st = oracle.rotated_scan(W1, D1, U1, 0.3, 0.8, Gm)
    chi = (st["beta"] / st["se"]) ** 2
    assert rel(o["wald"], chi) < 1e-12
    assert rel(o["beta"][:, 0], st["beta"]) < 1e-12
    assert rel(np.sqrt(o["vbeta"][:, 0, 0]), st["se"]) < 1e-12
    assert rel(o["p"], st["p"]) < 1e-12


def test_mv_oracle_kronecker_gls():
    """The per-eigen-sample scan equals brute-force GLS with V = Vg (x) K + Ve (x) I in the
    projected space (explicit (n-c) d x (n-c) d inverse)."""
    N, M = 60, 40
    bed = oracle.gen_hwe_bed(5, N, M, 0.1, 0.5, 0.02)
    raw, _, S, _ = oracle.grm_exact_centred(bed, N, M)
    G = raw / S
    rng = np.random.default_rng(7)
    X = np.column_stack([np.ones(N), rng.standard_normal(N)]).astype(np.float32).astype(np.float64)
    Y = rng.normal(0, 1, (N, 2))
    D, W, Uty, _, _ = oracle.mv_prepare(G, Y, X)
    Vg = np.array([[0.5, 0.2], [0.2, 0.4]])
    Ve = np.array([[0.7, -0.1], [-0.1, 0.9]])
    o = oracle.mv_scan(bed, N, M, None, W, D, Uty, Vg, Ve, np.arange(12))
    Q, _ = np.linalg.qr(X, mode="complete")
    A = Q[:, 2:]
    ne = N - 2
    Vi = np.linalg.inv(np.kron(Vg, A.T @ G @ A) + np.kron(Ve, np.eye(ne)))
    yv = (A.T @ Y).T.reshape(-1)
    Dg = oracle.decode_numpy(bed, N, M)[:, :12]
    obs = ~np.isnan(Dg)
    gm = np.where(obs, Dg - np.nansum(Dg, 0) / obs.sum(0), 0.0)
    for j in range(12):
        Gd = np.kron(np.eye(2), (A.T @ gm[:, j]).reshape(-1, 1))
        H = Gd.T @ Vi @ Gd
        beta = np.linalg.solve(H, Gd.T @ Vi @ yv)
        assert abs(beta @ H @ beta - o["wald"][j]) <= 1e-11 * o["wald"][j]
        assert np.max(np.abs(beta - o["beta"][j])) <= 1e-12 * np.max(np.abs(beta)) + 1e-15
        assert np.max(np.abs(np.linalg.inv(H) - o["vbeta"][j])) <= 1e-11 * np.max(np.abs(o["vbeta"][j]))


def test_mv_oracle_score_is_gradient_and_em_ascends():
    """The AI-REML score is the gradient of the REML log-likelihood (central differences on the
    packed Vg / Ve entries), the log-likelihood equals the explicit Kronecker form, and EM never
    decreases it."""
    bed, N, M, G, X, Y = _mv_case(N=200, M=800, seed=9, d=2)
    D, W, Uty, Vg, Ve = oracle.mv_prepare(G, Y, X)
    Vg = Vg + 0.05 * np.eye(2)
    AI, score = oracle.mv_ai(D, Uty, Vg, Ve)
    pairs = [(0, 0), (0, 1), (1, 1)]
    h = 1e-6
    for a, (j, k) in enumerate(pairs):
        for which, off in ((0, 0), (1, 3)):
            P = [Vg.copy(), Ve.copy()]
            M1 = [Vg.copy(), Ve.copy()]
            P[which][j, k] += h
            P[which][k, j] = P[which][j, k]
            M1[which][j, k] -= h
            M1[which][k, j] = M1[which][j, k]
            fd = (oracle.mv_logl(D, Uty, *P) - oracle.mv_logl(D, Uty, *M1)) / (2 * h)
            assert abs(fd - score[a + off]) <= 1e-5 * max(1.0, abs(fd))
    ne = D.size
    V = np.kron(Vg, np.diag(D)) + np.kron(Ve + 1e-8 * np.eye(2), np.eye(ne))
    yv = Uty.T.reshape(-1)
    s, ld = np.linalg.slogdet(V)
    assert abs(-0.5 * (ld + yv @ np.linalg.solve(V, yv)) - oracle.mv_logl(D, Uty, Vg, Ve)) < 1e-8 * ne
    ll = oracle.mv_logl(D, Uty, Vg, Ve)
    for _ in range(20):
        Vg, Ve = oracle.mv_em_update(D, Uty, Vg, Ve)
        nl = oracle.mv_logl(D, Uty, Vg, Ve)
        assert nl >= ll - 1e-9 * abs(ll)
        ll = nl


def test_multiply_z_zt_restatement():
    # plink_reader.cpp:241-438 by a per-genotype loop over the packed bytes on a tiny case, and
    # the identity Z (Z^T V) / M = K V against the kinship operator restatement (all SNPs)
    N, M = 37, 23
    bed = oracle.gen_hwe_bed(5, N, M, 0.0, 0.5, 0.05)
    means, inv = oracle.stats(bed, N, M)
    B = (N + 3) // 4
    raw = np.asarray(bed, dtype=np.uint8)
    lut = [2.0, None, 1.0, 0.0]
    keep = np.array([4, 0, 36, 4, 17, 9], dtype=np.int32)
    v = np.random.default_rng(1).standard_normal(keep.size)
    u_loop = np.zeros(M)
    for j in range(M):
        if inv[j] == 0.0:
            continue
        acc = 0.0
        for r, smp in enumerate(keep):
            code = (raw[j * B + smp // 4] >> (2 * (smp % 4))) & 3
            if code != 1:
                acc += (lut[code] - means[j]) * v[r]
        u_loop[j] = acc * inv[j]
    assert np.max(np.abs(oracle.multiply_zt(bed, N, M, v[:, None], keep)[:, 0] - u_loop)) < 1e-12
    w = np.random.default_rng(2).standard_normal(M)
    y_loop = np.zeros(keep.size)
    for j in range(M):
        if inv[j] == 0.0:
            continue
        for r, smp in enumerate(keep):
            code = (raw[j * B + smp // 4] >> (2 * (smp % 4))) & 3
            if code != 1:
                y_loop[r] += (lut[code] - means[j]) * w[j] * inv[j]
    assert np.max(np.abs(oracle.multiply_z(bed, N, M, w[:, None], keep)[:, 0] - y_loop)) < 1e-12
    X = np.random.default_rng(3).standard_normal((N, 3))
    KX = oracle.kinship_apply(bed, N, means, inv, None, None, X)
    ZZX = oracle.multiply_z(bed, N, M, oracle.multiply_zt(bed, N, M, X)) / M
    assert rel(ZZX, KX) < 1e-12


def test_hwe_exact_known_answers_and_enumeration():
    # tests/test_qc_pipeline.cpp:166-167 holds calculateHWE(3, 0, 3) = 5/231; beyond it, the exact
    # conditional distribution of the heterozygote count given the allele counts enumerated with
    # rationals: P(h) ~ n! 2^h / (h! homr! homc!), p = sum of P(h') <= P(h_obs)
    from fractions import Fraction
    from math import factorial
    assert abs(oracle.hwe_p(3, 0, 3) - 5.0 / 231) < 1e-14
    assert oracle.hwe_p(0, 0, 0) == 1.0 and oracle.hwe_p(10, 0, 0) == 1.0

    def exact(hom1, het, hom2):
        n = hom1 + het + hom2
        rare = 2 * min(hom1, hom2) + het
        probs = {}
        for h in range(rare % 2, rare + 1, 2):
            r = (rare - h) // 2
            c = n - h - r
            if c < 0:
                continue
            probs[h] = Fraction(factorial(n) * 2 ** h, factorial(h) * factorial(r) * factorial(c))
        tot = sum(probs.values())
        return float(sum(v for v in probs.values() if v <= probs[het]) / tot)

    for hom1, het, hom2 in [(3, 0, 3), (10, 20, 10), (0, 30, 0), (25, 2, 3), (1, 1, 40), (57, 3, 0),
                            (100, 80, 20), (7, 0, 9), (0, 1, 0), (12, 13, 11)]:
        assert abs(oracle.hwe_p(hom1, het, hom2) - exact(hom1, het, hom2)) <= 1e-12, (hom1, het, hom2)
        assert oracle.hwe_p(hom1, het, hom2) == oracle.hwe_p(hom2, het, hom1)
