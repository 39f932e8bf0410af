"""GPU test of cs_kinship_apply_device_noalloc, the K.V entry point that allocates nothing (fixed
per-GPU workspace): through the C-ABI, bitwise equal to cs_kinship_apply_device on the same
operator and inputs, on two calls in a row whose second input is smaller (fewer samples and
columns, another operator), both with a stale workspace from the call before; a request larger
than the workspace is refused before any device work.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def csa():
    import cosmicscanner_amd as csa
    return csa


def run_pair(csa, op, n, k, seed):
    X = np.asfortranarray(np.random.default_rng(seed).standard_normal((n, k)))
    dX = csa.DeviceArray.from_host(X)
    dY = csa.DeviceArray(n, k)
    dR = csa.DeviceArray(n, k)
    rc = csa.lib().cs_kinship_apply_device_noalloc(op._h, dX.ptr, n, k, dY.ptr, n)
    assert rc == csa.CS_OK, csa.lib().cs_last_error()
    csa._check(csa.lib().cs_kinship_apply_device(op._h, dX.ptr, n, k, dR.ptr, n))
    y, r = dY.download(), dR.download()
    for a in (dX, dY, dR):
        a.close()
    return X, y, r


def test_noalloc_matches_apply_device(csa):
    N, M = 1500, 5200
    body = oracle.gen_hwe_bed(99, N, M, 0.02, 0.5, 0.004)
    reader = csa.PlinkReader.from_body(body, N, M)
    means, inv = reader.getMeans(), reader.getInvStds()
    # first call: all samples, a LOCO-like variant list, k = 37 (three column groups)
    variants = np.arange(0, 4000, dtype=np.int32)
    op1 = csa.PlinkKinshipOperator(reader, None, variants)
    print("noalloc: n=1500 k=37", flush=True)
    X1, y1, r1 = run_pair(csa, op1, N, 37, 1)
    assert np.array_equal(y1, r1)
    ref = oracle.kinship_apply(body, N, means, inv, None, variants, X1)
    assert np.max(np.abs(y1 - ref)) <= 2e-6 * np.max(np.abs(ref))
    # second call, smaller: a keep list with repeats, k = 5 (packed limbs) and k = 16
    keep = np.sort(np.random.default_rng(2).choice(N, 700, replace=True)).astype(np.int32)
    op2 = csa.PlinkKinshipOperator(reader, keep, None)
    for k in (5, 16):
        print(f"noalloc: n=700 k={k}", flush=True)
        X2, y2, r2 = run_pair(csa, op2, keep.size, k, 3 + k)
        assert np.array_equal(y2, r2)
        ref2 = oracle.kinship_apply(body, N, means, inv, keep, None, X2)
        assert np.max(np.abs(y2 - ref2)) <= 2e-6 * np.max(np.abs(ref2))
    # the same operator through the sparse missing-genotype form of apply_device: still bitwise
    op2.set_missing_mode(2)
    X3, y3, r3 = run_pair(csa, op2, keep.size, 20, 7)
    assert np.array_equal(y3, r3)
    # too large for the 256 MiB workspace: refused with a message, before any device work
    dX = csa.DeviceArray(N, 1)
    rc = csa.lib().cs_kinship_apply_device_noalloc(op1._h, dX.ptr, N, 30000, dX.ptr, N)
    assert rc == csa.CS_ERR_INVALID_ARGUMENT and b"workspace" in csa.lib().cs_last_error()
    dX.close()
    op1.close()
    op2.close()
    reader.close()
