"""GPU parity of kernels (a) decode/stats and (b) K.V against the CPU oracle, through the C-ABI.

Mirrors the reference's own kinship test (tests/test_kinship_batch.cpp): the same n = 37 x m =
1031 fixture (partial last byte, monomorphic and all-missing SNPs), the six sample / variant
selections, RHS widths 1 / 7 / 41, trace, repeat determinism, empty RHS and dimension errors.

Tolerances.  Integer work (class counts, generated bytes) and the FP64 statistics are compared
bit for bit.  K.V in the default INT8_L3 mode quantises each RHS column to 23-bit fixed point (and
the intermediate Z^T V likewise), so the bound is relative to the column maximum:
max|Y - K X| <= 2e-6 * max|K X| (per column).  INT8_L2 (15-bit): 1e-3.  INT8_L4 (31-bit): 3e-8.
FP32: 2e-5.
trace/diag are FP64: relative 1e-10 as in the reference test.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def csa():
    import cosmicscanner_amd as csa
    return csa


def colrel(Y, R):
    den = np.maximum(np.max(np.abs(R), axis=0), 1e-300)
    return float(np.max(np.max(np.abs(Y - R), axis=0) / den))


def test_stats_bitwise(csa, kb_fixture, ref_datasets):
    bed, N, M, _ = kb_fixture
    cases = [(bed, N, M)] + [(v[0], v[1], v[2]) for v in ref_datasets.values()]
    for body, n, m in cases:
        r = csa.PlinkReader.from_body(body, n, m)
        means, inv = oracle.stats(body, n, m)
        assert np.array_equal(r.getMeans(), means)
        assert np.array_equal(r.getInvStds(), inv)
        assert np.array_equal(r.class_counts(), oracle.class_counts(body, n, m))
        w = np.random.default_rng(n).integers(0, 3, n).astype(np.int32)
        assert np.array_equal(r.class_counts(w), oracle.class_counts(body, n, m, w))


def test_load_file_and_magic(csa, tmp_path):
    import os
    prefix = os.path.join(os.path.dirname(__file__), "golden", "ref_synth", "synth")
    r = csa.PlinkReader.load(prefix)
    body, N, M, _, _ = oracle.read_bed(prefix)
    assert (r.getNumSamples(), r.getNumSnps()) == (N, M)
    assert np.array_equal(r.download(), body)
    p = tmp_path / "bad"
    (p.with_suffix(".fam")).write_text("a a 0 0 0 1\n")
    (p.with_suffix(".bim")).write_text("1 s 0 1 A C\n")
    (p.with_suffix(".bed")).write_bytes(bytes([0x6C, 0x1B, 0x00, 0]))
    with pytest.raises(RuntimeError, match="SNP-major"):
        csa.PlinkReader.load(str(p))


@pytest.mark.parametrize("N,M", [(1001, 300), (256, 257), (37, 1031)])
def test_synthetic_generator_bytes(csa, N, M):
    r = csa.PlinkReader.synthetic(20261204, N, M, 0.01, 0.5, 0.005)
    assert np.array_equal(r.download(), oracle.gen_hwe_bed(20261204, N, M, 0.01, 0.5, 0.005))


def test_kinship_batch_selections(csa, kb_fixture):
    bed, N, M, sels = kb_fixture
    reader = csa.PlinkReader.from_body(bed, N, M)
    means, inv = oracle.stats(bed, N, M)
    rng = np.random.default_rng(8128)
    for keep, variants in sels:
        op = csa.createKinshipOperator(reader, keep, variants)
        n = N if keep is None else keep.size
        assert op.rows() == n and op.cols() == n
        assert op.numVariants() == (M if variants is None else variants.size)
        tr = oracle.kinship_trace(bed, N, means, inv, keep, variants)
        assert abs(op.trace() - tr) <= 1e-10 * abs(tr)
        d = oracle.kinship_diag(bed, N, means, inv, keep, variants)
        assert np.max(np.abs(op.getDiagonal() - d)) <= 1e-10 * np.max(np.abs(d))
        for width in (1, 7, 41):
            X = rng.standard_normal((n, width))
            ref = oracle.kinship_apply(bed, N, means, inv, keep, variants, X)
            Y = op.apply(X)
            assert colrel(Y, ref) < 2e-6, (keep is None, variants is None, width)
            assert np.array_equal(op.apply(X), Y)                   # bitwise repeat
            v = X[:, 0]
            assert colrel(op.applyTranspose(v)[:, None], ref[:, :1]) < 2e-6
        out = op.apply(np.zeros((n, 0)))
        assert out.shape == (n, 0)
        with pytest.raises(RuntimeError, match="dimension mismatch"):
            op.apply(np.ones((n + 1, 1)))
        op.close()


def test_kinship_errors(csa, kb_fixture):
    bed, N, M, _ = kb_fixture
    reader = csa.PlinkReader.from_body(bed, N, M)
    with pytest.raises(RuntimeError, match="sample index out of range"):
        csa.PlinkKinshipOperator(reader, [0, N])
    with pytest.raises(RuntimeError, match="variant index out of range"):
        csa.PlinkKinshipOperator(reader, None, [M])


@pytest.mark.parametrize("mode,tol", [(3, 2e-6), (2, 1e-3), (32, 2e-5), (4, 3e-8)])
def test_kinship_precision_modes_larger(csa, mode, tol):
    N, M = 3000, 4100
    body = oracle.gen_hwe_bed(77, N, M, 0.01, 0.5, 0.01)
    reader = csa.PlinkReader.from_body(body, N, M)
    means, inv = oracle.stats(body, N, M)
    rng = np.random.default_rng(5)
    keep = np.sort(rng.choice(N, 2200, replace=False)).astype(np.int32)
    chrom = np.arange(M) * 22 // M
    loco = np.flatnonzero(chrom != 3).astype(np.int32)
    for kp, vp in [(None, None), (keep, loco)]:
        op = csa.PlinkKinshipOperator(reader, kp, vp)
        op.set_precision(mode)
        X = rng.standard_normal((op.rows(), 41 if mode == 4 else 16))   # INT8_L4: three groups
        ref = oracle.kinship_apply(body, N, means, inv, kp, vp, X)
        err = colrel(op.apply(X), ref)
        print(f"mode {mode}: column-relative error {err:.2e}", flush=True)
        assert err < tol
        op.close()


def test_kinship_fp32_comparator_widths(csa):
    """The FP32 comparator's kernels for 1, 4 and 16 columns per pass (and several passes),
    with a keep list that repeats samples and a variant selection, against the oracle."""
    N, M = 1700, 3300
    body = oracle.gen_hwe_bed(81, N, M, 0.01, 0.5, 0.01)
    reader = csa.PlinkReader.from_body(body, N, M)
    means, inv = oracle.stats(body, N, M)
    rng = np.random.default_rng(9)
    keep = rng.choice(N, 1400, replace=True).astype(np.int32)
    snps = np.sort(rng.choice(M, 2900, replace=False)).astype(np.int32)
    op = csa.PlinkKinshipOperator(reader, keep, snps)
    op.set_precision(csa.CS_PREC_FP32)
    for k in (1, 3, 5, 20):
        X = rng.standard_normal((op.rows(), k))
        ref = oracle.kinship_apply(body, N, means, inv, keep, snps, X)
        err = colrel(op.apply(X), ref)
        print(f"FP32 k {k}: column-relative error {err:.2e}", flush=True)
        assert err < 2e-5
    op.close()


@pytest.mark.parametrize("mode,tol", [(2, 1e-3), (3, 2e-6), (4, 3e-8)])
def test_kinship_packed_limbs_small_k(csa, mode, tol):
    """Up to 16 / L right-hand sides share one MFMA tile (limb l of column c in tile column
    c L + l); widths on both sides of that limit, against the oracle."""
    N, M = 1500, 2300
    body = oracle.gen_hwe_bed(78, N, M, 0.01, 0.5, 0.01)
    reader = csa.PlinkReader.from_body(body, N, M)
    means, inv = oracle.stats(body, N, M)
    rng = np.random.default_rng(8)
    keep = rng.choice(N, 1300, replace=True).astype(np.int32)
    op = csa.PlinkKinshipOperator(reader, keep, None)
    op.set_precision(mode)
    for k in sorted({1, 16 // mode, 16 // mode + 1}):
        X = rng.standard_normal((op.rows(), k))
        ref = oracle.kinship_apply(body, N, means, inv, keep, None, X)
        err = colrel(op.apply(X), ref)
        print(f"mode {mode} k {k}: column-relative error {err:.2e}", flush=True)
        assert err < tol
    op.close()


@pytest.mark.parametrize("mode", [2, 3, 4])
def test_partial_last_group_bitwise(csa, mode):
    """k = 16 g + r with r L <= 16 runs the last r columns as their own packed-limb apply; every
    column is quantised and multiplied on its own, so the columns equal the same columns of a
    wider apply that needs no split (bitwise), and the oracle within the precision's bound."""
    N, M = 1400, 2700
    body = oracle.gen_hwe_bed(79, N, M, 0.01, 0.5, 0.01)
    reader = csa.PlinkReader.from_body(body, N, M)
    means, inv = oracle.stats(body, N, M)
    rng = np.random.default_rng(10)
    keep = np.sort(rng.choice(N, 1100, replace=False)).astype(np.int32)
    op = csa.PlinkKinshipOperator(reader, keep, None)
    op.set_precision(mode)
    r = 16 // mode
    k = 32 + r
    X = rng.standard_normal((op.rows(), 48))
    wide = op.apply(X)                      # 48 columns: three whole groups, no split
    split = op.apply(X[:, :k])              # two groups + r packed-limb columns
    assert np.array_equal(split, wide[:, :k])
    ref = oracle.kinship_apply(body, N, means, inv, keep, None, X[:, :k])
    tol = {2: 1e-3, 3: 2e-6, 4: 3e-8}[mode]
    assert colrel(split, ref) < tol
    op.close()


def test_kinship_sketch_width_266(csa):
    # The reference's randomized low-rank path applies K to n x (256 + 10) sketches
    # (lmm_engine.cpp:915-946); widths above 256 columns exercise the per-column scale grid.
    N, M = 700, 2300
    body = oracle.gen_hwe_bed(31, N, M, 0.01, 0.5, 0.01)
    reader = csa.PlinkReader.from_body(body, N, M)
    means, inv = oracle.stats(body, N, M)
    op = csa.PlinkKinshipOperator(reader)
    X = np.random.default_rng(2).standard_normal((N, 266))
    ref = oracle.kinship_apply(body, N, means, inv, None, None, X)
    assert colrel(op.apply(X), ref) < 2e-6


@pytest.mark.parametrize("miss", [0.005, 0.05, 0.001, 0.0])
def test_sparse_and_dense_missing_forms_bitwise(csa, miss):
    # Both forms compute the same exact integer sums: the results must agree bit for bit, with
    # keep lists (repeats, any order), a LOCO variant selection and a partial last byte.  Auto
    # (mode 0) takes the sparse form at 0.1 % and 0 % missing, the dense one above 0.15 %.
    N, M = 2999, 4100
    r = csa.PlinkReader.synthetic(8, N, M, 0.01, 0.5, miss)
    rng = np.random.default_rng(1)
    keep = rng.choice(N, 2500, replace=True).astype(np.int32)
    variants = np.sort(rng.choice(M, 3000, replace=False)).astype(np.int32)
    op = csa.PlinkKinshipOperator(r, keep, variants)
    X = rng.standard_normal((keep.size, 37))
    outs = []
    for mode in (1, 2, 0):
        op.set_missing_mode(mode)
        outs.append(op.apply(X))
    assert np.array_equal(outs[0], outs[1])
    assert np.array_equal(outs[0], outs[2])
    body = r.download()
    means, inv = oracle.stats(body, N, M)
    ref = oracle.kinship_apply(body, N, means, inv, keep, variants, X)
    assert colrel(outs[1], ref) < 2e-6


@pytest.mark.parametrize("prec,k,tol", [(3, 64, 2e-6), (3, 20, 2e-6), (2, 20, 1e-3), (2, 64, 1e-3)])
def test_sparse_missing_form_precisions_bitwise(csa, prec, k, tol):
    """The 16-quad list gathers (64 columns: 16 plain int32 quads per row) and the two-limb
    genotype-only products (INT8_L2): sparse and dense forms bitwise equal, auto (mode 0, sparse
    at 0.1 % missing) equal to both, and within the precision's bound of the oracle."""
    N, M = 1800, 3100
    r = csa.PlinkReader.synthetic(31, N, M, 0.01, 0.5, 0.001)
    rng = np.random.default_rng(12)
    keep = rng.choice(N, 1500, replace=True).astype(np.int32)
    op = csa.PlinkKinshipOperator(r, keep, None)
    op.set_precision(prec)
    X = rng.standard_normal((keep.size, k))
    outs = []
    for mode in (1, 2, 0):
        op.set_missing_mode(mode)
        outs.append(op.apply(X))
    assert np.array_equal(outs[0], outs[1])
    assert np.array_equal(outs[0], outs[2])
    body = r.download()
    means, inv = oracle.stats(body, N, M)
    ref = oracle.kinship_apply(body, N, means, inv, keep, None, X)
    err = colrel(outs[1], ref)
    print(f"prec {prec} k {k}: column-relative error {err:.2e}", flush=True)
    assert err < tol
    op.close()


@pytest.mark.parametrize("k", [120, 250])
def test_sparse_missing_form_wide_blocks_bitwise(csa, k):
    """Wide blocks take the full-wave list gathers: 120 columns (eight groups, 128 plain int32
    columns = 32 quads per row) and 250 (sixteen groups, 64 quads).  Bitwise equal to the dense
    form, and within the operator's bound of the oracle."""
    N, M = 1500, 2300
    r = csa.PlinkReader.synthetic(21, N, M, 0.01, 0.5, 0.01)
    op = csa.PlinkKinshipOperator(r, None, None)
    X = np.random.default_rng(5).standard_normal((N, k))
    outs = []
    for mode in (1, 2):
        op.set_missing_mode(mode)
        outs.append(op.apply(X))
    assert np.array_equal(outs[0], outs[1])
    body = r.download()
    means, inv = oracle.stats(body, N, M)
    ref = oracle.kinship_apply(body, N, means, inv, None, None, X)
    assert colrel(outs[1], ref) < 2e-6


def test_register_kernels_match_lds_kernels(csa, tmp_path):
    """CSAMD_GENO_LDSB=0 selects the register-only kernels (no LDS sharing, two groups per wave
    for several groups, no limb packing); CSAMD_GENO_DMA=0 stages the B operands of the LDS
    kernels through registers instead of direct global-to-LDS loads.  The integer products are exact, so every
    column must be bitwise identical to the default kernels' (each setting in a child process:
    the switches are read once per process)."""
    import os
    import subprocess
    import sys
    N, M = 900, 2100
    script = f"""
import sys, numpy as np
sys.path.insert(0, {repr(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))})
import oracle, cosmicscanner_amd as csa
body = oracle.gen_hwe_bed(79, {N}, {M}, 0.01, 0.5, 0.01)
r = csa.PlinkReader.from_body(body, {N}, {M})
keep = np.random.default_rng(3).choice({N}, 800, replace=True).astype(np.int32)
op = csa.PlinkKinshipOperator(r, keep, None)
X = np.random.default_rng(4).standard_normal((800, 41))
out = {{}}
for mode in (2, 3, 4):
    op.set_precision(mode)
    for k in (1, 16, 41):
        out[f"m{{mode}}k{{k}}"] = op.apply(X[:, :k])
np.savez(sys.argv[1], **out)
"""
    paths = {name: str(tmp_path / f"{name}.npz") for name in ("lds", "reg", "nodma")}
    envs = {"lds": {}, "reg": {"CSAMD_GENO_LDSB": "0"}, "nodma": {"CSAMD_GENO_DMA": "0"}}
    for name in ("lds", "reg", "nodma"):
        print(f"child process: {name} {envs[name]}", flush=True)
        subprocess.run([sys.executable, "-c", script, paths[name]], check=True, timeout=120,
                       env=dict(os.environ, **envs[name]))
    a = np.load(paths["lds"])
    for other in ("reg", "nodma"):
        b = np.load(paths[other])
        for key in a.files:
            assert np.array_equal(a[key], b[key]), (other, key)


def test_multiply_z_and_zt(csa):
    """PlinkReader::multiply_Zt_v / multiply_Z_v (plink_reader.cpp:241-438) over all SNPs, with and
    without a keep list, against the oracle's dense Z: the int8 products carry the 23-bit operand
    quantisation (<= 2e-6 of each output column's maximum); wrong leading dimensions raise the
    reference's dimension-mismatch errors."""
    N, M = 613, 2900
    body = oracle.gen_hwe_bed(404, N, M, 0.0, 0.5, 0.01)
    reader = csa.PlinkReader.from_body(body, N, M)
    rng = np.random.default_rng(8)
    for keep in (None, np.sort(rng.choice(N, 300, replace=False)).astype(np.int32)):
        op = csa.PlinkKinshipOperator(reader, keep, np.arange(0, M, 2, dtype=np.int32))   # selection ignored
        n = N if keep is None else keep.size
        for k in (1, 5, 37):
            V = rng.standard_normal((n, k))
            U = op.multiply_Zt_v(V)
            ref = oracle.multiply_zt(body, N, M, V, keep)
            assert np.max(np.abs(U - ref), axis=0).max() <= 2e-6 * np.max(np.abs(ref))
            W = rng.standard_normal((M, k))
            Y = op.multiply_Z_v(W)
            ref2 = oracle.multiply_z(body, N, M, W, keep)
            assert np.max(np.abs(Y - ref2)) <= 2e-6 * np.max(np.abs(ref2))
        u1 = op.multiply_Zt_v(V[:, 0])
        assert u1.shape == (M,)
        L = csa.lib()
        X = np.zeros((n, 1))
        Yb = np.zeros((M, 1))
        rc = L.cs_kinship_multiply_zt(op._h, X.ctypes.data_as(csa._dp), n - 1, 1, Yb.ctypes.data_as(csa._dp), M)
        assert rc == csa.CS_ERR_RUNTIME and b"multiply_Zt_v" in L.cs_last_error()
        op.close()
    reader.close()


@pytest.mark.parametrize("miss", [0.001, 0.02])
def test_compact_rows_bitwise(csa, miss):
    """cs_kinship_compact_rows: the genotype products on a compact copy of the cohort's rows give
    bitwise the full-sample products' results -- unsorted and sorted cohorts, a variant
    selection, INT8_L2/L3/L4, dense / sparse / auto missing forms (the copy has its own missing
    fraction), packed (k L <= 16) and grouped operands, Z^T V and Z U alone, and the
    no-allocation entry point.  FP32 (summation order changes) within its bound of the oracle;
    trace and diag unchanged; repeated samples refused."""
    N, M = 1800, 3100
    r = csa.PlinkReader.synthetic(41, N, M, 0.01, 0.5, miss)
    rng = np.random.default_rng(9)
    variants = np.sort(rng.choice(M, 2400, replace=False)).astype(np.int32)
    body = r.download()
    means, inv = oracle.stats(body, N, M)
    for keep in (rng.permutation(N)[:700].astype(np.int32), np.sort(rng.choice(N, 1200, replace=False)).astype(np.int32)):
        n = keep.size
        print(f"compact rows: miss={miss} n={n}", flush=True)
        full = csa.PlinkKinshipOperator(r, keep, variants)
        comp = csa.PlinkKinshipOperator(r, keep, variants)
        b0 = comp.device_bytes()
        comp.compact_rows()
        comp.compact_rows()                       # idempotent
        npad, mpad = -(-n // 256) * 256, -(-M // 256) * 256
        assert comp.device_bytes() - b0 >= 2 * npad * mpad // 4
        assert abs(comp.trace() / full.trace() - 1) <= 1e-12
        assert np.array_equal(comp.getDiagonal(), full.getDiagonal())
        for prec in (csa.CS_PREC_INT8_L3, csa.CS_PREC_INT8_L2, csa.CS_PREC_INT8_L4):
            full.set_precision(prec)
            comp.set_precision(prec)
            for k in (3, 20):
                X = rng.standard_normal((n, k))
                for mode in (1, 2, 0):
                    full.set_missing_mode(mode)
                    comp.set_missing_mode(mode)
                    assert np.array_equal(comp.apply(X), full.apply(X)), (prec, k, mode)
        full.set_precision(csa.CS_PREC_INT8_L3)
        comp.set_precision(csa.CS_PREC_INT8_L3)
        X = rng.standard_normal((n, 9))
        U = rng.standard_normal((M, 9))
        assert np.array_equal(comp.multiply_Zt_v(X), full.multiply_Zt_v(X))
        assert np.array_equal(comp.multiply_Z_v(U), full.multiply_Z_v(U))
        dX = csa.DeviceArray.from_host(X)
        dY = csa.DeviceArray(n, 9)
        comp.apply_device_noalloc(dX.ptr, n, 9, dY.ptr, n)
        assert np.array_equal(dY.download(), full.apply(X))
        dX.close()
        dY.close()
        comp.set_precision(csa.CS_PREC_FP32)
        ref = oracle.kinship_apply(body, N, means, inv, keep, variants, X)
        assert colrel(comp.apply(X), ref) < 2e-5
        full.close()
        comp.close()
    dup = csa.PlinkKinshipOperator(r, np.array([3, 5, 3, 9], dtype=np.int32), None)
    with pytest.raises(ValueError, match="twice"):
        dup.compact_rows()
    dup.close()
    allrows = csa.PlinkKinshipOperator(r, None, None)
    b0 = allrows.device_bytes()
    allrows.compact_rows()                        # the all-sample cohort: nothing to copy
    assert allrows.device_bytes() == b0
    allrows.close()


def test_device_cache_release_between_applies(csa):
    """A cohort operator's compact row copies (two layouts of 18.5 MB here) go back to the device
    cache when the operator is destroyed and are handed to the next one; cs_dev_release_cache
    frees them.  Each apply matches the uncompacted operator bitwise, whichever path its
    memory came from."""
    N, M, n = 7000, 12000, 6000
    reader = csa.PlinkReader.synthetic(515, N, M, 0.01, 0.5, 0.01)
    rng = np.random.default_rng(12)
    keep = np.sort(rng.choice(N, n, replace=False)).astype(np.int32)
    X = rng.standard_normal((n, 24))
    plain = csa.PlinkKinshipOperator(reader, keep, None)
    want = plain.apply(X)
    plain.close()
    for release in (False, False, True, False):
        op = csa.PlinkKinshipOperator(reader, keep, None)
        op.compact_rows()
        assert np.array_equal(op.apply(X), want), release
        op.close()
        if release:
            csa._check(csa.lib().cs_dev_release_cache())
    reader.close()
