"""Concurrent kNN calls on one index (fuel's SearchIndex is shared by many
threads): every thread's answers equal the same calls made one after another
and the oracle's, in every candidate mode, with range searches on the same
index running beside them."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["exact", "bf16", "bf16r", "fp8r", "fp32"])
def test_threads_share_one_index(gpu, oracle, mode):
    from fuel_amd import workloads
    rng = np.random.default_rng(81)
    X = workloads.gaussian_mixture(rng, 30000, 48, 10, -6.0, 6.0)
    Qs = [workloads.gaussian_mixture(rng, int(rng.integers(1, 400)), 48, 10, -6.0, 6.0) for _ in range(24)]
    ks = [int(rng.choice([1, 7, 10, 33])) for _ in Qs]
    ix = gpu.Index(X, "euclidean", mode=mode)
    alone = [ix.knn(q, k) for q, k in zip(Qs, ks)]
    got = [None] * len(Qs)
    errors = []

    def worker(t):
        try:
            for i in range(t, len(Qs), 6):
                got[i] = ix.knn(Qs[i], ks[i])
        except Exception as e:  # surfaced below
            errors.append(e)

    def ranger():
        try:
            for r in (1.0, 2.0, 3.0):
                ix.range(X[:50], r)
        except Exception as e:
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(t,)) for t in range(6)] + [threading.Thread(target=ranger)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for (ai, ad), (gi, gd) in zip(alone, got):
        assert np.array_equal(ai, gi) and ad.tobytes() == gd.tobytes()
    for i in (0, 5, 11):
        oi, od = oracle.knn(X, Qs[i], ks[i], "euclidean", "none")
        assert np.array_equal(got[i][0], oi) and got[i][1].tobytes() == od.tobytes()
    st = ix.stats()
    assert st["queries"] in {q.shape[0] for q in Qs}
