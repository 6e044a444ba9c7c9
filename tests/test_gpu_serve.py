"""Serving loop (continuous batching): every request's answer is the same
alone and in the mix, and equals the oracle's linear scan on the same rows;
admission by batch size and HBM budget; requests from several threads."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mixture(n, d, comps, seed):
    from fuel_amd import workloads
    rng = np.random.default_rng(seed)
    return workloads.gaussian_mixture(rng, n, d, comps, -10.0, 10.0)


@pytest.fixture(scope="module")
def served(gpu):
    # clustered rows: the serving configuration's shape (50-component mixture,
    # d = 96) at 120k rows, where the bf16 bound sends queries down the chain
    X = _mixture(120_000, 96, 50, 61)
    Q = _mixture(3000, 96, 50, 62)
    return X, Q


def _requests(Q, rng, ks=(10,)):
    reqs, off = [], 0
    while off < Q.shape[0]:
        m = int(min(Q.shape[0] - off, rng.geometric(1.0 / 40)))
        reqs.append((off, m, int(rng.choice(ks))))
        off += m
    return reqs


@pytest.mark.parametrize("mode", ["bf16", "bf16r", "fp8r", "fp32", "exact"])
def test_mix_equals_alone_and_oracle(gpu, oracle, served, mode):
    from fuel_amd import Server
    X, Q = served
    ix = gpu.Index(X, "euclidean", mode=mode)
    rng = np.random.default_rng(7)
    reqs = _requests(Q, rng, ks=(10, 10, 10, 3))
    results = {}
    with Server(ix, workers=2, max_batch=512) as srv:
        lock = threading.Lock()

        def client(part):
            tickets = [(r, srv.submit(Q[r[0]:r[0] + r[1]], r[2])) for r in part]
            for r, t in tickets:
                idx, dist, lat, bq = srv.wait(t)
                assert lat >= 0.0 and r[1] <= bq <= 512
                with lock:
                    results[r] = (idx, dist)

        threads = [threading.Thread(target=client, args=(reqs[i::3],)) for i in range(3)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        st = srv.stats()
    assert len(results) == len(reqs)
    assert st["requests"] == len(reqs) and st["failed"] == 0
    assert st["queries"] == Q.shape[0] and st["max_batch_queries"] <= 512
    assert st["batches"] < len(reqs)  # requests were coalesced
    assert 0.0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    # alone: each sampled request by itself through the synchronous entry point
    for r in reqs[::7]:
        ai, ad = ix.knn(Q[r[0]:r[0] + r[1]], r[2])
        gi, gd = results[r]
        assert np.array_equal(ai, gi) and ai.dtype == gi.dtype
        assert ad.tobytes() == gd.tobytes()
    # and against the oracle on a sample of the rows
    for r in reqs[:: max(1, len(reqs) // 6)]:
        q = Q[r[0]:r[0] + r[1]]
        oi, od = oracle.knn(X, q, r[2], "euclidean", "none")
        gi, gd = results[r]
        assert np.array_equal(oi, gi) and od.tobytes() == gd.tobytes()


def test_budget_and_edge_requests(gpu):
    from fuel_amd import Server
    from fuel_amd._lib import FuelAmdError
    rng = np.random.default_rng(3)
    X = rng.random((5000, 16), dtype=np.float32)
    ix = gpu.Index(X, "sqeuclidean")
    # a budget below the smallest batch's estimate rejects the request
    with Server(ix, workers=1, max_batch=64, hbm_budget=1 << 20) as srv:
        t = srv.submit(X[:4], 5)
        with pytest.raises(FuelAmdError):
            srv.wait(t)
        with pytest.raises(ValueError):
            srv.submit(X[:65], 5)  # above max_batch
    with Server(ix, workers=2, max_batch=64) as srv:
        t0 = srv.submit(X[:0], 5)
        t1 = srv.submit(X[:3], 0)
        i0, d0, _, _ = srv.wait(t0)
        i1, d1, _, _ = srv.wait(t1)
        assert i0.shape == (0, 5) and i1.shape == (3, 0)
        with pytest.raises(FuelAmdError):
            srv.wait(123456789)  # unknown ticket (checked before the C call)
    # k beyond n pads with -1 / +inf like the synchronous call
    small = gpu.Index(X[:8], "sqeuclidean")
    with Server(small, workers=1) as srv:
        gi, gd, _, _ = srv.wait(srv.submit(X[:2], 10))
    ai, ad = small.knn(X[:2], 10)
    assert np.array_equal(ai, gi) and ad.tobytes() == gd.tobytes() and gi[0, 9] == -1


def test_unwaited_tickets_and_f64(gpu, oracle):
    from fuel_amd import Server
    rng = np.random.default_rng(4)
    X = rng.standard_normal((3000, 12))
    ix = gpu.Index(X, "euclidean")
    srv = Server(ix, workers=2, max_batch=256)
    tickets = [srv.submit(X[i:i + 17], 6) for i in range(0, 300, 17)]
    gi, gd, _, _ = srv.wait(tickets[3])
    oi, od = oracle.knn(X, X[51:68], 6, "euclidean", "none")
    assert np.array_equal(gi, oi) and gd.tobytes() == od.tobytes()
    srv.close()  # answers and releases the tickets never waited for


def test_two_servers_and_direct_calls_on_one_index(gpu, oracle):
    # per-caller workspaces: two servers and the synchronous entry point share
    # one resident index from several threads at once
    from fuel_amd import Server
    rng = np.random.default_rng(8)
    X = _mixture(30_000, 32, 20, 81)
    Q = _mixture(600, 32, 20, 82)
    ix = gpu.Index(X, "euclidean", mode="bf16r")
    ref_i, ref_d = oracle.knn(X, Q, 10, "euclidean", "none")
    out = {}
    errs = []

    def serve(srv, lo, hi, key):
        try:
            ts = [(i, srv.submit(Q[i:i + 25], 10)) for i in range(lo, hi, 25)]
            out[key] = [(i, srv.wait(t)[:2]) for i, t in ts]
        except Exception as e:  # surfaced below
            errs.append(e)

    def direct(lo, hi):
        try:
            out["direct"] = [(lo, ix.knn(Q[lo:hi], 10))]
        except Exception as e:
            errs.append(e)

    with Server(ix, workers=2, max_batch=256) as s1, Server(ix, workers=1, max_batch=128) as s2:
        ths = [threading.Thread(target=serve, args=(s1, 0, 300, "s1")),
               threading.Thread(target=serve, args=(s2, 300, 600, "s2")),
               threading.Thread(target=direct, args=(100, 400))]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
    assert not errs, errs
    for key in ("s1", "s2", "direct"):
        for lo, (gi, gd) in out[key]:
            n = gi.shape[0]
            assert np.array_equal(gi, ref_i[lo:lo + n]) and gd.tobytes() == ref_d[lo:lo + n].tobytes(), key


@pytest.mark.parametrize("k", [127, 200])
def test_large_k_through_the_server(gpu, oracle, k):
    # k + 1 = 128 is the widest MFMA kept list; beyond it the exact scan answers
    from fuel_amd import Server
    X = _mixture(8000, 16, 10, 90)
    Q = X[:40] + np.float32(0.01)
    ix = gpu.Index(X, "sqeuclidean", mode="bf16r")
    with Server(ix) as srv:
        gi, gd, _, _ = srv.wait(srv.submit(Q, k))
    oi, od = oracle.knn(X, Q, k, "sqeuclidean", "none")
    assert np.array_equal(gi, oi) and gd.tobytes() == od.tobytes()


def test_one_wait_per_ticket(gpu):
    # a second wait on a ticket fails cleanly (never a double release), whether it
    # comes while the first waits or after the ticket was released
    from fuel_amd import Server, _lib
    X = _mixture(200_000, 64, 20, 95)
    Q = _mixture(2000, 64, 20, 96)
    ix = gpu.Index(X, "euclidean", mode="bf16")
    lib = _lib.load()
    with Server(ix, workers=1, max_batch=4096) as srv:
        for _ in range(3):
            t = srv.submit(Q, 10)
            rcs = []
            outs = [(np.empty((2000, 10), np.int64), np.empty((2000, 10), np.float32)) for _ in range(2)]

            def waiter(o):
                rcs.append(lib.fuel_amd_server_wait(srv.handle, t, o[0].ctypes.data, o[1].ctypes.data, None, None))

            ths = [threading.Thread(target=waiter, args=(o,)) for o in outs]
            for th in ths:
                th.start()
            for th in ths:
                th.join()
            assert sorted(rcs) == sorted([_lib.OK, _lib.ERR_INVALID]), rcs
            srv._k.pop(t, None)
            # the waiter that succeeded holds the answer of the synchronous entry point
            ai, ad = ix.knn(Q[:50], 10)
            good = [o for o in outs if np.array_equal(o[0][:50], ai)]
            assert good and good[0][1][:50].tobytes() == ad.tobytes()
