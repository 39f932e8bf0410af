"""GPU parity of kernel (b), the dense GRM build (GrmBuilder::compute, algorithm 1 VanRaden,
libCosmicCore/src/grm_builder.cpp:63-399), against the CPU oracle through the C-ABI.

The GPU computes Z Z^T exactly in integers and rounds once, so it is compared BIT FOR BIT with
the oracle's exact-integer restatement (oracle.grm_exact_centred: correctly rounded FP64 raw
Z Z^T and correctly rounded float32 of Z Z^T / sum_2pq).  Against the oracle's FP64 restatement
of the reference's own arithmetic (oracle.grm: FP64 Z fill from the float32 mean, block GEMM
accumulation) the bound is max|G_gpu - G_fp64| <= 1e-12 * max|G| (FP64 accumulation error only).
sum_2pq and the valid-SNP count are compared exactly.

Cases: the reference's committed BED fixtures, the kinship-batch fixture (partial last byte,
monomorphic and all-missing SNPs), device-generated HWE data with and without missing genotypes
at sample counts that are not multiples of the 128-row block, the multi-panel path, and the
GCTA writer (GrmBuilder::save, grm_builder.cpp:1977-2026).

The standardised GRM (algorithms 0 and 3, grm_builder.cpp:241-243, 386-387) runs as FP16 hi/lo
split products with FP32 accumulation (flushed to FP64 every 16384 SNPs).  Its bound against the
oracle's FP64 restatement is max|G_gpu - G_fp64| <= STD_TOL * max diag(G_fp64): the split alone
contributes ~5e-8, the FP32 accumulation the rest (about 1.4e-6 at n=400, m=6000 in an emulation
of the same rounding; it grows with the chunk length, about chunk / sqrt(m) x 1e-7 of the diagonal,
so 1-3e-6 for the 16384-SNP chunks at m <= 40000 and below 1e-6 at m = 500000).  The valid-SNP count and sum_2pq are compared exactly.
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


def lower(a):
    return a[np.tril_indices(a.shape[0])]


def check_against_oracle(csa, reader, body, N, M, min_maf=0.01, min_hwe=0.0):
    b = csa.GrmBuilder(csa.GrmBuilder.Options(min_maf=min_maf, min_hwe=min_hwe))
    b.compute(reader)
    raw, nf, s2pq, nv = oracle.grm_exact_centred(body, N, M, min_maf, min_hwe=min_hwe)
    assert b.getSum2pq() == s2pq
    assert b.getTotalSnps() == nv
    assert np.array_equal(b.getRaw(), raw)
    assert np.array_equal(b.getLowerTriangle(), lower(nf))
    G64, s64, nv64 = oracle.grm(body, N, M, 1, min_maf, min_hwe=min_hwe)
    assert (s64, nv64) == (s2pq, nv)
    G = b.getGrm()
    assert np.max(np.abs(G - G64)) <= 1e-12 * np.max(np.abs(G64))
    return b


def test_grm_reference_fixtures(csa, ref_datasets):
    for name, (body, N, M, _, _) in ref_datasets.items():
        print(f"GRM fixture {name}: N={N} M={M}", flush=True)
        check_against_oracle(csa, csa.PlinkReader.from_body(body, N, M), body, N, M)


def test_grm_kinship_fixture(csa, kb_fixture):
    bed, N, M, _ = kb_fixture
    check_against_oracle(csa, csa.PlinkReader.from_body(bed, N, M), bed, N, M)


@pytest.mark.parametrize("N,M,miss", [(301, 1000, 0.02), (700, 3000, 0.005), (517, 2200, 0.0)])
def test_grm_synthetic(csa, N, M, miss):
    print(f"GRM synthetic N={N} M={M} miss={miss}", flush=True)
    r = csa.PlinkReader.synthetic(91 + N, N, M, 0.01, 0.5, miss)
    body = r.download()
    check_against_oracle(csa, r, body, N, M)


def test_grm_multi_panel(csa, monkeypatch):
    # a panel bound of one 128-row block forces the row-panel loop
    N, M = 450, 1500
    monkeypatch.setenv("CSAMD_GRM_PANEL_BYTES", "1")
    r = csa.PlinkReader.synthetic(7, N, M, 0.01, 0.5, 0.01)
    check_against_oracle(csa, r, r.download(), N, M)


def test_grm_min_maf_filter(csa):
    N, M = 200, 800
    r = csa.PlinkReader.synthetic(11, N, M, 0.0, 0.5, 0.01)
    body = r.download()
    for maf in (0.05, 0.2):
        check_against_oracle(csa, r, body, N, M, min_maf=maf)


@pytest.mark.parametrize("maf", [0.001, 0.0])
def test_grm_low_maf(csa, maf):
    # rare SNPs: float means below 2^-6 that are not multiples of 2^-29 need a larger common
    # shift (more fractional bits, more digits); bit-exact all the same
    N, M = 700, 2600
    r = csa.PlinkReader.synthetic(404, N, M, 0.0002, 0.05, 0.01)
    body = r.download()
    valid, mean_f, _, _, _, _ = oracle.grm_snp_params(body, N, M, 1, maf)
    assert np.any(valid & (np.ldexp(mean_f.astype(np.float64), 29) % 1 != 0))
    check_against_oracle(csa, r, body, N, M, min_maf=maf)


def test_grm_many_missing_pairs(csa, monkeypatch):
    # 12 % missing: pairs of samples missing at the same SNP are common (the D sums), over several
    # 128-row panels
    N, M = 390, 1800
    monkeypatch.setenv("CSAMD_GRM_PANEL_BYTES", "1")
    r = csa.PlinkReader.synthetic(77, N, M, 0.01, 0.5, 0.12)
    check_against_oracle(csa, r, r.download(), N, M)


def test_grm_long_missing_lists_and_snp_ranges(csa, monkeypatch):
    # 50 % missing over 9000 SNPs: every sample misses more SNPs than one cursor batch of the pair
    # sums holds (4096), and E passes forced onto SNP ranges of 8 blocks (2048 SNPs)
    N, M = 190, 9000
    monkeypatch.setenv("CSAMD_GRM_EPASS_KB", "8")
    r = csa.PlinkReader.synthetic(5150, N, M, 0.01, 0.5, 0.5)
    check_against_oracle(csa, r, r.download(), N, M)


def test_grm_hwe_filter(csa):
    # GrmBuilder's optional exact-HWE SNP filter (grm_builder.cpp:220-223): SNPs with an excess of
    # heterozygotes or homozygotes are dropped from both GRMs at min_hwe = 1e-4
    N, M = 400, 1500
    r0 = csa.PlinkReader.synthetic(55, N, M, 0.05, 0.5, 0.01)
    body = r0.download().reshape(M, -1).copy()
    rng = np.random.default_rng(6)
    for j in rng.choice(M, 60, replace=False):
        codes = np.where(rng.random(N) < 0.8, 2, rng.choice([0, 3], N))    # mostly heterozygous
        codes[rng.random(N) < 0.01] = 1
        pad = np.zeros(body.shape[1] * 4, dtype=np.uint8)
        pad[:N] = codes
        q = pad.reshape(-1, 4)
        body[j] = q[:, 0] | (q[:, 1] << 2) | (q[:, 2] << 4) | (q[:, 3] << 6)
    body = body.reshape(-1)
    r = csa.PlinkReader.from_body(body, N, M)
    b = check_against_oracle(csa, r, body, N, M, min_hwe=1e-4)
    assert b.getTotalSnps() <= M - 55
    b0 = csa.GrmBuilder(csa.GrmBuilder.Options(algorithm=0, min_hwe=1e-4))
    b0.compute(r)
    _, _, nv0 = oracle.grm(body, N, M, 0, 0.01, min_hwe=1e-4)
    assert b0.getTotalSnps() == nv0 == b.getTotalSnps()
    with pytest.raises(ValueError, match="min_hwe"):
        csa.GrmBuilder(csa.GrmBuilder.Options(min_hwe=1.5)).compute(r)


def test_grm_device_output(csa):
    N, M = 333, 1200
    r = csa.PlinkReader.synthetic(5, N, M, 0.01, 0.5, 0.005)
    nt = N * (N + 1) // 2
    d = csa.DeviceArray((nt + 1) // 2)
    s2pq, nv = csa.grm_build_device(r, d.ptr)
    tri = d.download().reshape(-1).view(np.float32)[:nt]
    b = csa.GrmBuilder()
    b.compute(r, fp64=False)
    assert np.array_equal(tri, b.getLowerTriangle())
    assert (s2pq, nv) == (b.getSum2pq(), b.getTotalSnps())


def test_grm_errors(csa, kb_fixture):
    bed, N, M, _ = kb_fixture
    r = csa.PlinkReader.from_body(bed, N, M)
    with pytest.raises(ValueError, match="algorithms 0/3"):
        csa.GrmBuilder(csa.GrmBuilder.Options(algorithm=2)).compute(r)
    with pytest.raises(RuntimeError, match="subsetting"):
        csa.GrmBuilder().compute(r, keep_indices=[0, 1])
    with pytest.raises(RuntimeError, match="not computed"):
        csa.GrmBuilder().save("/nonexistent")


def test_grm_save_gcta(csa, ref_datasets, tmp_path):
    body, N, M, bim, fam = ref_datasets["synth"]
    prefix = os.path.join(os.path.dirname(__file__), "golden", "ref_synth", "synth")
    r = csa.PlinkReader.load(prefix)
    b = csa.GrmBuilder()
    b.compute(r)
    out = str(tmp_path / "g")
    b.save(out)
    tri = np.fromfile(out + ".grm.bin", dtype="<f4")
    nb = np.fromfile(out + ".grm.N.bin", dtype="<f4")
    _, nf, _, nv = oracle.grm_exact_centred(body, N, M)
    assert np.array_equal(tri, lower(nf))
    assert nb.size == N * (N + 1) // 2 and np.all(nb == np.float32(nv))
    ids = open(out + ".grm.id").read().splitlines()
    assert ids == [f"{row[0]}\t{row[1]}" for row in fam]


# ------------------------------------------------------------------ standardised GRM (alg 0, 3)
STD_TOL = 8e-6


def check_std(csa, reader, body, N, M, min_maf=0.01, alg=0):
    b = csa.GrmBuilder(csa.GrmBuilder.Options(min_maf=min_maf, algorithm=alg))
    b.compute(reader)
    G64, s64, nv64 = oracle.grm(body, N, M, 0, min_maf)
    assert b.getTotalSnps() == nv64
    assert b.getSum2pq() == s64
    G = b.getGrm()
    scale = np.max(np.abs(np.diag(G64)))
    err = np.max(np.abs(G - G64))
    print(f"  standardised GRM max|dG| = {err:.3e} ({err / scale:.3e} of max diag)", flush=True)
    assert err <= STD_TOL * scale
    assert np.array_equal(G, G.T)
    tri = b.getLowerTriangle()
    assert np.array_equal(tri, lower(G).astype(np.float32))
    return b, err / scale


def test_grm_std_reference_fixtures(csa, ref_datasets):
    for name, (body, N, M, _, _) in ref_datasets.items():
        print(f"standardised GRM fixture {name}: N={N} M={M}", flush=True)
        check_std(csa, csa.PlinkReader.from_body(body, N, M), body, N, M)


def test_grm_std_kinship_fixture(csa, kb_fixture):
    bed, N, M, _ = kb_fixture
    check_std(csa, csa.PlinkReader.from_body(bed, N, M), bed, N, M, alg=3)


@pytest.mark.parametrize("N,M,miss", [(301, 1000, 0.02), (700, 3000, 0.005), (517, 2200, 0.0)])
def test_grm_std_synthetic(csa, N, M, miss):
    print(f"standardised GRM synthetic N={N} M={M} miss={miss}", flush=True)
    r = csa.PlinkReader.synthetic(191 + N, N, M, 0.01, 0.5, miss)
    check_std(csa, r, r.download(), N, M)


def test_grm_std_two_chunks_and_panels(csa, monkeypatch):
    # 40000 SNPs span three FP32 accumulation chunks; a panel bound of one 256-row block forces
    # the row-panel loop
    N, M = 600, 40000
    monkeypatch.setenv("CSAMD_GRM_PANEL_BYTES", "1")
    r = csa.PlinkReader.synthetic(17, N, M, 0.01, 0.5, 0.01)
    check_std(csa, r, r.download(), N, M)


def test_grm_std_min_maf_and_device(csa):
    N, M = 280, 1500
    r = csa.PlinkReader.synthetic(23, N, M, 0.0, 0.5, 0.01)
    body = r.download()
    for maf in (0.05, 0.2):
        check_std(csa, r, body, N, M, min_maf=maf)
    nt = N * (N + 1) // 2
    d = csa.DeviceArray((nt + 1) // 2)
    s2pq, nv = csa.grm_build_device(r, d.ptr, alg=0)
    tri = d.download().reshape(-1).view(np.float32)[:nt]
    b = csa.GrmBuilder(csa.GrmBuilder.Options(algorithm=0))
    b.compute(r, fp64=False)
    assert np.array_equal(tri, b.getLowerTriangle())
    assert (s2pq, nv) == (b.getSum2pq(), b.getTotalSnps())
