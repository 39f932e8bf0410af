"""GPU tests of the serving loop (continuous batching of scan requests).

A request's columns are numerically independent of the rest of the batch (per-column
quantisation scales, exact integer products, per-column PCG recurrences and fixed-order dots), so
its results must be BIT-IDENTICAL whether it runs alone or mixed with other requests that arrive
while it is on the GPU -- also when a small column capacity forces requests to queue and be
admitted as slots free up.  Against the standalone projected scan with a cohort-restricted
operator (same estimator, compact rows, different dot partition) the bound is
|d log10 P| <= 1e-5 (PCG tolerance 1e-9).
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
def world(csa):
    N, M = 640, 2600
    reader = csa.PlinkReader.synthetic(4242, N, M, 0.05, 0.5, 0.01)
    rng = np.random.default_rng(11)
    reqs = []
    for r in range(5):
        n = int(rng.integers(220, N))
        keep = np.sort(rng.choice(N, n, replace=False)).astype(np.int32)
        T = int(rng.integers(1, 5))
        c = int(rng.integers(1, 4))
        X = np.column_stack([np.ones(n)] + [rng.standard_normal(n) for _ in range(c - 1)]).astype(np.float32)
        Y = rng.standard_normal((n, T)) + 0.3 * X[:, -1:]
        h2 = float(rng.uniform(0.1, 0.8))
        cal = rng.choice(M, 16, replace=False).astype(np.int32)
        snps = np.sort(rng.choice(M, 700, replace=False)).astype(np.int32)
        reqs.append(dict(Y=Y, X=X, sigma2_g=h2, sigma2_e=1 - h2, cal_snps=cal, keep=keep, snps=snps, rel_tol=1e-9,
                         max_iter=400))
    return reader, reqs


def run_alone(csa, reader, req, cap=128):
    srv = csa.ScanServer(reader, max_columns=cap)
    rid = srv.submit(**req)
    srv.run_until_idle()
    return srv.result(rid)


def same(a, b):
    for k in ("beta", "se", "p", "log10p", "freq"):
        x, y = a[k], b[k]
        assert np.array_equal(np.isnan(x), np.isnan(y)), k
        assert np.array_equal(x[~np.isnan(x)], y[~np.isnan(y)]), k
    assert a["stats"]["iterations"] == b["stats"]["iterations"]


@pytest.mark.parametrize("cap", [128, 40])
def test_alone_equals_mix(csa, world, cap):
    reader, reqs = world
    alone = [run_alone(csa, reader, q) for q in reqs]
    srv = csa.ScanServer(reader, max_columns=cap)
    ids = [srv.submit(**reqs[0]), srv.submit(**reqs[1])]
    srv.step()
    srv.step()
    ids.append(srv.submit(**reqs[2]))
    srv.step()
    ids += [srv.submit(**reqs[3]), srv.submit(**reqs[4])]
    saw_pending = False
    for _ in range(100000):
        a, p, _ = srv.step()
        saw_pending = saw_pending or p > 0
        if a == 0 and p == 0:
            break
    if cap == 40:
        assert saw_pending            # the small capacity made requests queue
    for rid, ref in zip(ids, alone):
        same(srv.result(rid), ref)
        st = srv.status(rid)[1]
        assert st["t_submit"] <= st["t_admit"] <= st["t_done"]


def test_server_matches_projected_scan(csa, world):
    reader, reqs = world
    q = reqs[1]
    got = run_alone(csa, reader, q)
    op = csa.PlinkKinshipOperator(reader, q["keep"])
    ref = csa.pcg_projected_scan(op, reader, q["keep"], q["Y"], q["X"], q["sigma2_g"], q["sigma2_e"], q["cal_snps"],
                                 snps=q["snps"], rel_tol=1e-9, max_iter=400)
    assert abs(got["stats"]["gamma"] / ref["gamma"] - 1) <= 1e-7
    m = ~np.isnan(ref["log10p"])
    assert np.array_equal(np.isnan(got["log10p"]), ~m)
    assert np.max(np.abs(got["log10p"][m] - ref["log10p"][m])) <= 1e-5
    # and the oracle's dense-solve restatement of the estimator
    body = reader.download()
    st, gamma, _ = oracle.pcg_projected_scan(body, reader.num_samples, reader.num_snps, q["keep"], None, q["Y"], q["X"],
                                             q["sigma2_g"], q["sigma2_e"], q["cal_snps"])
    assert abs(got["stats"]["gamma"] / gamma - 1) <= 1e-4
    for t in range(q["Y"].shape[1]):
        lo = st[t]["log10p"][q["snps"]]
        mm = ~np.isnan(lo)
        assert np.max(np.abs(got["log10p"][:, t][mm] - lo[mm])) <= 1e-3


def test_server_rejects_bad_requests(csa, world):
    reader, reqs = world
    srv = csa.ScanServer(reader, max_columns=64)
    q = dict(reqs[0])
    with pytest.raises(ValueError):
        srv.submit(**dict(q, sigma2_e=0.0))
    with pytest.raises(ValueError):
        srv.submit(**dict(q, keep=np.array([1, 1, 2] + list(range(3, q["keep"].size)), dtype=np.int32)))
    # a request wider than the column capacity fails on admission, without stalling the loop
    srv2 = csa.ScanServer(reader, max_columns=4)
    rid = srv2.submit(**q)
    srv2.run_until_idle()
    assert srv2.status(rid)[0] == -1
    with pytest.raises(RuntimeError, match="more columns"):
        srv2.result(rid)
