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
