"""fuel's LinearScanPrioritySearcher on the GPU (fuel_amd_priority_*): every
call -- next, next with a filter, all_lower_bound, decrease_cutoff, reset,
reset_with_limits -- returns what the oracle's restatement of
src/search/linear_scan.rs:70-235 returns, bit for bit, over seeded random
scripts of calls on many slots (so that groups of slots are materialized
together while others are reset or advanced), on grids with heavy ties and on
random f32 / f64 data."""
import itertools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same(a, b):
    if a is None or b is None:
        return a is None and b is None
    return a[0] == b[0] and np.asarray(a[1]).tobytes() == np.asarray(b[1]).tobytes()


def _script(gpu, oracle, X, Q, metric, seed, steps=400, arith="avx"):
    rng = np.random.default_rng(seed)
    ix = gpu.Index(X, metric, mode="exact", arith=arith)
    ps = ix.priority(Q)
    refs = [oracle.PrioritySearcher(X, Q[q], metric, arith) for q in range(Q.shape[0])]
    cut = [float("inf")] * Q.shape[0]
    for step in range(steps):
        q = int(rng.integers(0, Q.shape[0]))
        op = rng.random()
        if op < 0.55:
            assert _same(ps.next(q), refs[q].next()), (step, q)
        elif op < 0.65:
            f = lambda i: i % 3 == 0  # noqa: E731
            assert _same(ps.next(q, f), refs[q].next(f)), (step, q)
        elif op < 0.75:
            a, b = ps.all_lower_bound(q), refs[q].all_lower_bound()
            assert np.asarray(a).tobytes() == np.asarray(b).tobytes(), (step, q)
        elif op < 0.85:
            lb = float(refs[q].all_lower_bound())
            t = lb + rng.random() * 3.0 if np.isfinite(lb) else float(rng.random() * 5.0)
            t = min(t, cut[q])
            cut[q] = t
            ps.decrease_cutoff(q, t)
            refs[q].decrease_cutoff(t)
        elif op < 0.93:
            c = float(rng.random() * 6.0)
            s = float(rng.random() * c)
            cut[q] = c
            ps.reset_with_limits(q, c, s)
            refs[q].reset_with_limits(c, s)
        else:
            cut[q] = float("inf")
            ps.reset(q)
            refs[q].reset()
    # drain a few slots completely
    for q in range(0, Q.shape[0], 7):
        while True:
            a, b = ps.next(q), refs[q].next()
            assert _same(a, b), q
            if a is None:
                break
    ps.close()


def test_grid_with_ties(gpu, oracle):
    g = np.array(list(itertools.product(range(7), range(7), range(4))), np.float32)
    X = np.concatenate([g, g[:40]])
    _script(gpu, oracle, X, X[::3], "euclidean", 1)
    _script(gpu, oracle, X, X[::5], "sqeuclidean", 2)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_random_data(gpu, oracle, dtype):
    rng = np.random.default_rng(5)
    X = rng.standard_normal((3000, 6)).astype(dtype)
    Q = rng.standard_normal((150, 6)).astype(dtype)
    _script(gpu, oracle, X, Q, "euclidean", 3)


def test_fallback_arithmetic(gpu, oracle):
    rng = np.random.default_rng(6)
    X = (rng.standard_normal((1500, 21)) * rng.choice([0.01, 1.0, 30.0], (1, 21))).astype(np.float32)
    _script(gpu, oracle, X, X[::19], "sqeuclidean", 4, arith="fallback")


def test_fresh_slot_semantics(gpu, oracle):
    # before its first next() a slot reports a lower bound of 0 even when its
    # list was built together with another slot's
    X = np.array([[0, 0], [1, 0], [0, 1], [1, 1], [2, 2]], np.float64)
    ix = gpu.Index(X, "euclidean", mode="exact")
    ps = ix.priority(X)
    assert ps.next(0)[0] == 0
    assert float(ps.all_lower_bound(1)) == 0.0
    assert [ps.next(1)[0] for _ in range(5)] == [1, 0, 3, 2, 4]
    assert ps.next(1) is None
    ps.decrease_cutoff(0, 1.0)
    with pytest.raises(ValueError):
        ps.decrease_cutoff(0, 2.0)
    with pytest.raises(ValueError):
        ps.reset_with_limits(0, 1.0, 2.0)
    ps.close()
