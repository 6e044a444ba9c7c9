"""The residual operand (FUEL_AMD_MODE_BF16R): rows partitioned around coarse
centroids, the bf16 MFMA on (q - c_j).(x - c_j) with per-list bounds folded
into the C input.  Parity with the oracle on clustered data (where the plain
bf16 bound is too wide to certify), and the certification rate that makes the
mode worth having: on such data nearly every query must certify at the first
level, which also fails if the residual keys were not lower bounds (parity)
or far too loose (rate)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mixture(rng, n, d, comps, lo, hi):
    from fuel_amd import workloads
    return workloads.gaussian_mixture(rng, n, d, comps, lo, hi)


def _check(gpu, oracle, X, Q, k, metric="euclidean", exclude=0, mode="bf16r"):
    ix = gpu.Index(X, metric, mode=mode)
    gi, gd = ix.knn(Q, k, exclude)
    st = ix.stats()
    oi, od = oracle.knn(X, Q, k, metric, exclude)
    bad = np.argwhere(np.any(gi != oi, axis=1)).ravel()
    assert bad.size == 0, (bad[:5], st)
    assert gd.tobytes() == od.tobytes()
    return st


@pytest.mark.parametrize("d", [8, 24, 96, 128, 200])
@pytest.mark.parametrize("metric", ["euclidean", "sqeuclidean"])
def test_clustered_parity_and_rate(gpu, oracle, d, metric):
    rng = np.random.default_rng(100 + d)
    means = rng.uniform(-10, 10, size=(50, d))
    X = (means[rng.integers(0, 50, 40000)] + rng.standard_normal((40000, d))).astype(np.float32)
    Q = (means[rng.integers(0, 50, 400)] + rng.standard_normal((400, d))).astype(np.float32)
    st = _check(gpu, oracle, X, Q, 10, metric)
    assert st["mode_used"] == 4
    assert st["fallback_queries"] <= 0.10 * Q.shape[0], st
    plain = gpu.Index(X, metric, mode="bf16")
    plain.knn(Q, 10)
    assert st["fallback_queries"] <= plain.stats()["fallback_queries"], (st, plain.stats())


@pytest.mark.parametrize("exclude", [1, 2])
@pytest.mark.parametrize("k", [1, 15, 40])
def test_self_knn_exclusion(gpu, oracle, exclude, k):
    rng = np.random.default_rng(7 + k)
    X = _mixture(rng, 12000, 64, 12, -5.0, 5.0)
    _check(gpu, oracle, X, X[::37], k, "euclidean", exclude)


@pytest.mark.parametrize("n", [1, 63, 64, 65, 511, 512, 1000, 5001])
def test_list_padding_and_small_n(gpu, oracle, n):
    # lists padded to 64-row blocks; n < 512 is one list (centring only)
    rng = np.random.default_rng(n)
    X = _mixture(rng, n, 32, 5, -8.0, 8.0)
    Q = _mixture(rng, 50, 32, 5, -8.0, 8.0)
    _check(gpu, oracle, X, Q, 7)


def test_duplicates_and_ties(gpu, oracle):
    import itertools
    g = np.array(list(itertools.product(range(6), range(6), range(5), range(4))), np.float32) * 3.0
    X = np.concatenate([g, g[:200], g + 0.5]).astype(np.float32)
    _check(gpu, oracle, X, X[::11], 12)
    _check(gpu, oracle, X, X[::11], 12, "sqeuclidean", 1)


def test_uniform_and_integer_data(gpu, oracle):
    rng = np.random.default_rng(5)
    X = rng.random((30000, 128), dtype=np.float32)
    _check(gpu, oracle, X, rng.random((300, 128), dtype=np.float32), 10)
    from fuel_amd import workloads
    C = workloads.c3(n=6000)
    _check(gpu, oracle, C, C[:120], 20, "euclidean", 1)


@pytest.mark.parametrize("mode", ["bf16r", "fp8r"])
@pytest.mark.parametrize("d", [1, 2, 3])
def test_low_dimension_uniform(gpu, oracle, mode, d):
    # One to three coordinates: the rounding of a single query coordinate is
    # not averaged out over many, so a residual key whose bound undercounted
    # the bf16 rounding (2^-8 of a coordinate, not 2^-9) lay above ||q - x||^2 / 2
    # and the pass from the certified threshold missed the nearest row
    # (2,000 rows in 1-D, 700 queries, k = 5: one query in 700).
    rng = np.random.default_rng(190153 + d)
    X = rng.random((2000, d), dtype=np.float32)
    Q = rng.random((700, d), dtype=np.float32)
    for k, metric, exclude in [(5, "euclidean", 0), (17, "sqeuclidean", 0), (10, "euclidean", 2)]:
        _check(gpu, oracle, X, Q if exclude == 0 else X[::3], k, metric, exclude, mode)
    if d == 1:
        # the sweep case that showed it, drawn as the sweep drew it
        r = np.random.default_rng(190153)
        r.choice([2000, 8500, 12000, 20000, 30000])
        r.choice([1, 5, 8, 16, 24, 31, 40, 48, 64, 65, 80, 96, 100, 112, 128, 200, 224, 240, 256])
        r.choice([1, 5, 64, 65, 200, 257, 700])
        r.choice([1, 2, 5, 10, 17, 40, 100])
        r.choice(["euclidean", "sqeuclidean"])
        r.choice([0, 1, 2])
        X1 = r.random((2000, 1), dtype=np.float32)
        Q1 = r.random((700, 1), dtype=np.float32)
        assert X1[1882, 0] == np.float32(0.97617567) and Q1[488, 0] == np.float32(0.97618175)
        _check(gpu, oracle, X1, Q1, 5, "euclidean", 0, mode)


def test_result_independent_of_batch_composition(gpu):
    rng = np.random.default_rng(9)
    X = _mixture(rng, 20000, 96, 50, -10.0, 10.0)
    Q = _mixture(rng, 700, 96, 50, -10.0, 10.0)
    ix = gpu.Index(X, "euclidean", mode="bf16r")
    ai, ad = ix.knn(Q, 10)
    for lo, hi in [(0, 1), (5, 70), (300, 700)]:
        bi, bd = ix.knn(Q[lo:hi], 10)
        assert np.array_equal(ai[lo:hi], bi) and ad[lo:hi].tobytes() == bd.tobytes()


@pytest.mark.parametrize("kind", ["clustered", "uniform"])
def test_auto_mode_chooses_by_probe(gpu, oracle, kind):
    # AUTO probes the plain bf16 certificate on 256 index rows: on the serving
    # mixture (means U[-10, 10]^96) it fails, and the index takes the residual
    # form; on uniform data it certifies and the index stays on plain bf16.
    # Either way the answer is the oracle's.
    rng = np.random.default_rng(31)
    if kind == "clustered":
        X = _mixture(rng, 60000, 96, 50, -10.0, 10.0)
        Q = X[::301] + np.float32(0.25)
    else:
        X = rng.random((60000, 96), dtype=np.float32)
        Q = rng.random((200, 96), dtype=np.float32)
    st = _check(gpu, oracle, X, Q, 10, mode="auto")
    assert st["mode_used"] == (4 if kind == "clustered" else 3), st
    # decided once: a second call keeps the choice
    ix = gpu.Index(X, "euclidean", mode="auto")
    ix.knn(Q[:5], 10)
    first = ix.stats()["mode_used"]
    ix.knn(Q, 3)
    assert ix.stats()["mode_used"] == first
