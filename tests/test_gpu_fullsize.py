"""K.V at the headline size (n = 50,000 samples x m = 500,000 SNPs, 0.5 % missing, genotypes
generated on the device), where the oracle would take minutes per apply: size-independent
properties instead of an oracle comparison.

* Column independence.  Every right-hand-side column is quantised with its own scale and the
  integer products are exact, so a column's result does not depend on the other columns nor on
  how the columns are tiled: the packed single-column kernel (k = 1), the one-group kernel
  (k = 16) and the three-group XCD-ordered kernel (k = 48) give bitwise identical columns.
* Symmetry.  w^T (K v) = v^T (K w) within the operator's 23-bit quantisation (1e-6 relative).
* Linearity.  K (a v + b w) = a K v + b K w within the same bound (per-column max).
* Diagonal.  (K e_i)_i = diag(K)_i (FP64 diag) within 2e-6 of the column maximum, and
  sum(diag) = trace (FP64, 1e-10 relative; both from the per-SNP statistics).
* Missing-genotype forms.  The dense observed-indicator products and the sparse form (missing
  lists) are exact integer sums, so they agree bit for bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, M = 50_000, 500_000


@pytest.fixture(scope="module")
def setup():
    import cosmicscanner_amd as csa
    print(f"full-size K.V n={N} m={M}", flush=True)
    reader = csa.PlinkReader.synthetic(20261204, N, M, 0.01, 0.5, 0.005)
    op = csa.PlinkKinshipOperator(reader, None, None)
    yield csa, op
    op.close()
    reader.close()


def apply_dev(csa, op, X):
    dX = csa.DeviceArray.from_host(np.asfortranarray(X))
    dY = csa.DeviceArray(X.shape[0], X.shape[1])
    op.apply_device(dX.ptr, dX.n, X.shape[1], dY.ptr, dY.n)
    Y = dY.download()
    dX.close()
    dY.close()
    return Y


def test_columns_independent_of_tiling(setup):
    csa, op = setup
    rng = np.random.default_rng(1)
    X = rng.standard_normal((N, 48))
    Y48 = apply_dev(csa, op, X)
    Y16 = apply_dev(csa, op, X[:, 16:32])
    Y1 = apply_dev(csa, op, X[:, 20:21])
    assert np.array_equal(Y48[:, 16:32], Y16)
    assert np.array_equal(Y16[:, 4:5], Y1)


def test_symmetry_and_linearity(setup):
    csa, op = setup
    rng = np.random.default_rng(2)
    v, w = rng.standard_normal(N), rng.standard_normal(N)
    a, b = 0.7, -1.3
    Y = apply_dev(csa, op, np.column_stack([v, w, a * v + b * w]))
    Kv, Kw, Kvw = Y[:, 0], Y[:, 1], Y[:, 2]
    s1, s2 = w @ Kv, v @ Kw
    print(f"  w.Kv {s1:.10e}  v.Kw {s2:.10e}", flush=True)
    assert abs(s1 - s2) <= 1e-6 * max(abs(s1), np.linalg.norm(v) * np.linalg.norm(Kw))
    lin = a * Kv + b * Kw
    assert np.max(np.abs(Kvw - lin)) <= 1e-6 * np.max(np.abs(lin))


def test_diagonal_and_trace(setup):
    csa, op = setup
    d = op.getDiagonal()
    t = op.trace()
    assert abs(d.sum() - t) <= 1e-10 * t
    idx = np.array([0, 1, 4097, 25_000, N - 1])
    E = np.zeros((N, idx.size))
    E[idx, np.arange(idx.size)] = 1.0
    Y = apply_dev(csa, op, E)
    cm = np.max(np.abs(Y), axis=0)
    err = np.abs(Y[idx, np.arange(idx.size)] - d[idx]) / cm
    print(f"  (K e_i)_i vs diag: max column-relative error {err.max():.2e}", flush=True)
    assert err.max() <= 2e-6


def test_sparse_missing_form_bitwise_at_full_size(setup):
    """The sparse observed-indicator form (genotype-only MFMA products plus exact corrections from
    the missing lists) against the dense form at the full size: 125 M missing genotypes, the
    list gathers striding over 500,000 SNP rows and 50,000 sample rows, k = 20 (two groups)."""
    csa, op = setup
    X = np.random.default_rng(3).standard_normal((N, 20))
    try:
        op.set_missing_mode(1)
        Yd = apply_dev(csa, op, X)
        op.set_missing_mode(2)
        Ys = apply_dev(csa, op, X)
    finally:
        op.set_missing_mode(0)
    assert np.array_equal(Yd, Ys)
