"""GPU tests of the serving loop (continuous batching of scan requests).

A request's columns are numerically independent of the rest of the batch (per-column
quantisation scales, exact integer products, per-column PCG recurrences and fixed-order dots), so
its results must be BIT-IDENTICAL whether it runs alone or mixed with other requests that arrive
while it is on the GPU -- also when a small column capacity forces requests to queue and be
admitted as slots free up.  The exact stage of a
request (its selected markers solved as a second phase of batch columns, in chunks that depend
on the free slots) keeps that property: its per-column dots run one workgroup per column.
Against the standalone projected scan with a cohort-restricted operator (same estimator, compact
rows, different dot partition) the bound is |d log10 P| <= 1e-5 (PCG tolerance 1e-9).

Admission and failure: a request that fails after admission (no calibration marker passes the
filter) gives its slots and budget back, so later requests still run; result arrays stay charged
to the HBM budget until the client releases them.
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
                         max_iter=400, exact_log10p=2.0))
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
                                 snps=q["snps"], rel_tol=1e-9, max_iter=400, exact_log10p=2.0)
    assert abs(got["stats"]["gamma"] / ref["gamma"] - 1) <= 1e-7
    m = ~np.isnan(ref["log10p"])
    assert np.array_equal(np.isnan(got["log10p"]), ~m)
    assert np.max(np.abs(got["log10p"][m] - ref["log10p"][m])) <= 1e-5
    assert got["stats"]["n_exact"] == ref["n_exact"] and ref["n_exact"] > 0
    # and the oracle's dense-solve restatement of the estimator, exact stage included
    body = reader.download()
    st, gamma, _ = oracle.pcg_projected_scan(body, reader.num_samples, reader.num_snps, q["keep"], None, q["Y"], q["X"],
                                             q["sigma2_g"], q["sigma2_e"], q["cal_snps"], exact_log10p=2.0)
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


def test_failed_request_releases_slots(csa, world):
    reader, reqs = world
    q = dict(reqs[0])
    T, c = q["Y"].shape[1], q["X"].shape[1]
    k_good = T + c + 16
    # capacity for one good request only: leaked slots of the failed one would block it for good
    srv = csa.ScanServer(reader, max_columns=k_good)
    bad = srv.submit(**dict(q, min_maf=0.6))        # every SNP fails the filter: no calibration marker
    srv.step()
    good = srv.submit(**q)
    for _ in range(2000):
        a, p, _ = srv.step()
        if a == 0 and p == 0:
            break
    assert srv.status(bad)[0] == -1
    with pytest.raises(RuntimeError, match="calibration"):
        srv.result(bad)
    assert srv.status(good)[0] == 2
    same(srv.result(good), run_alone(csa, reader, q))
    srv.release(bad)
    srv.release(good)


def test_compact_and_full_paths_bitwise(csa):
    """A cohort's request on its compact row copy (products, scan and exact-stage markers read the
    cohort's rows) and the same request on the full matrix through its keep list (a budget one
    byte short of the copy) give the same results bit for bit, exact stage included.  (Many SNPs
    per sample, so the copy outweighs the full-matrix workspace the second server needs.)"""
    N, M = 1000, 12000
    reader = csa.PlinkReader.synthetic(99, N, M, 0.05, 0.5, 0.01)
    rng = np.random.default_rng(23)
    n = int(0.3 * N)
    keep = np.sort(rng.choice(N, n, replace=False)).astype(np.int32)
    T, c = 3, 2
    q = dict(Y=rng.standard_normal((n, T)), X=np.column_stack([np.ones(n), rng.standard_normal(n)]).astype(np.float32),
             sigma2_g=0.6, sigma2_e=0.4, cal_snps=np.sort(rng.choice(M, 16, replace=False)).astype(np.int32),
             keep=keep, exact_log10p=1.0)
    big = csa.ScanServer(reader, max_columns=128)
    rid = big.submit(**q)
    big.run_until_idle()
    got_c = big.result(rid)
    st = got_c["stats"]
    assert st["compact_rows"] == 1 and st["n_exact"] > 0
    res = M * (4 * T + 3) * 8
    tight = csa.ScanServer(reader, max_columns=128, hbm_budget_bytes=st["admit_bytes"] + res - 1)
    rid2 = tight.submit(**q)
    tight.run_until_idle()
    got_f = tight.result(rid2)
    assert got_f["stats"]["compact_rows"] == 0
    same(got_c, got_f)
    for k in ("n_obs", "passes"):
        assert np.array_equal(got_c[k], got_f[k]), k
    assert got_c["stats"]["n_exact"] == got_f["stats"]["n_exact"]
    # (exact_iterations sums over the exact-stage chunks, whose size follows the free budget: the
    # tight server runs more, smaller chunks; each column's arithmetic is its own either way)
    big.close()
    tight.close()
    reader.close()


def test_budget_holds_unreleased_results(csa, world):
    reader, reqs = world
    N = reader.num_samples
    rng = np.random.default_rng(5)
    T, c = 2, 2
    X = np.column_stack([np.ones(N), rng.standard_normal(N)]).astype(np.float32)
    snps = np.arange(0, 700, dtype=np.int32)
    base = dict(X=X, sigma2_g=0.5, sigma2_e=0.5, cal_snps=np.arange(1000, 1016, dtype=np.int32), snps=snps,
                rel_tol=1e-8, max_iter=400)
    qa = dict(base, Y=rng.standard_normal((N, T)))
    qb = dict(base, Y=rng.standard_normal((N, T)))
    # what a request holds from admission (operator, preconditioner, projection, solve buffers and
    # workspace; reported in its stats) and its result arrays until release
    alone = csa.ScanServer(reader, max_columns=64)
    rid = alone.submit(**qa)
    alone.run_until_idle()
    st = alone.status(rid)[1]
    need = st["admit_bytes"]
    assert st["compact_rows"] == 0 and need > 10 * N * (T + c + 16) * 8
    res = snps.size * (4 * T + 3) * 8
    srv = csa.ScanServer(reader, max_columns=64, hbm_budget_bytes=need + res + res // 2)
    ia, ib = srv.submit(**qa), srv.submit(**qb)
    for _ in range(2000):
        srv.step()
        if srv.status(ia)[0] == 2:
            break
    assert srv.status(ia)[0] == 2
    for _ in range(30):
        srv.step()
    assert srv.status(ib)[0] == 0          # A's results still hold the budget
    got_a = srv.result(ia)
    srv.release(ia)
    srv.run_until_idle()
    assert srv.status(ib)[0] == 2
    same(got_a, run_alone(csa, reader, qa))
    same(srv.result(ib), run_alone(csa, reader, qb))


def test_fetch_on_another_thread_while_stepping(csa, world):
    # a client thread copies finished results out while the main thread keeps running steps; the
    # arrays are those of a request run alone, and release / destroy wait for no lock of their own
    from concurrent.futures import ThreadPoolExecutor
    reader, reqs = world
    alone = [run_alone(csa, reader, q) for q in reqs[:4]]
    srv = csa.ScanServer(reader, max_columns=128)
    ids = [srv.submit(**q) for q in reqs[:4]]
    futures = {}
    with ThreadPoolExecutor(max_workers=1) as pool:
        for _ in range(100000):
            a, p, _ = srv.step()
            for rid in ids:
                if rid not in futures and srv.status(rid)[0] == 2:
                    futures[rid] = pool.submit(srv.result, rid)
            if a == 0 and p == 0 and len(futures) == len(ids):
                break
        got = {rid: f.result() for rid, f in futures.items()}
    for rid, ref in zip(ids, alone):
        same(got[rid], ref)
        srv.release(rid)
    srv.close()
