"""k beyond one candidate slot.  fuel's KNNHeap takes any k
(src/api/search.rs:80-151, linear_scan_knn :169-184, SearchIndex.knn
python/search.rs:560-569, precompute=max_k); the engine finds the top k + 1
composites in rounds of 256 (each an exact scan that keeps only candidates
ordering after the previous round's last), so every mode answers any k.  The
answers must equal the oracle's bit for bit: packed rows with -1 / +inf
padding and both exclusion semantics, the tie-keeping CSR, the precomputed
table of SearchIndex(precompute=1000), and the batched problems."""
import itertools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _grid_with_ties():
    g = np.array(list(itertools.product(range(12), range(12), range(8))), np.float32)
    return np.concatenate([g, g[:300]])  # exact duplicates on top of grid ties


@pytest.mark.parametrize("k", [255, 256, 257, 300, 511, 513, 1000])
@pytest.mark.parametrize("exclude", [0, 1, 2])
def test_packed_rows_large_k(gpu, oracle, k, exclude):
    X = _grid_with_ties()
    Q = X[::53]
    name = {0: "none", 1: "for_each_knn", 2: "python_linear"}[exclude]
    for mode in ("exact", "bf16"):
        ix = gpu.Index(X, "euclidean", mode=mode)
        gi, gd = ix.knn(Q, k, exclude)
        oi, od = oracle.knn(X, Q, k, "euclidean", name)
        assert np.array_equal(gi, oi), (k, exclude, mode)
        assert gd.tobytes() == od.tobytes()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("metric", ["euclidean", "sqeuclidean"])
def test_random_data_large_k(gpu, oracle, dtype, metric):
    rng = np.random.default_rng(77)
    X = rng.standard_normal((5000, 19)).astype(dtype)
    Q = rng.standard_normal((70, 19)).astype(dtype)
    ix = gpu.Index(X, metric)
    for k in (300, 1000, 1777):
        gi, gd = ix.knn(Q, k)
        oi, od = oracle.knn(X, Q, k, metric)
        assert np.array_equal(gi, oi), k
        assert gd.tobytes() == od.tobytes()


def test_k_at_and_beyond_n(gpu, oracle):
    rng = np.random.default_rng(3)
    X = rng.random((600, 5)).astype(np.float32)
    ix = gpu.Index(X, "euclidean")
    for k in (599, 600, 601, 1000):
        for exclude, name in [(0, "none"), (1, "for_each_knn")]:
            gi, gd = ix.knn(X[:40], k, exclude)
            oi, od = oracle.knn(X, X[:40], k, "euclidean", name)
            assert np.array_equal(gi, oi) and gd.tobytes() == od.tobytes(), (k, exclude)
            assert (gi[:, max(0, 600 - exclude):] == -1).all()


def test_ties_csr_large_k(gpu, oracle):
    X = _grid_with_ties()
    ix = gpu.Index(X, "sqeuclidean")
    for k in (300, 1000):
        go, gi, gd = ix.knn_ties(X[::97], k)
        oo, oi, od = oracle.knn_ties(X, X[::97], k, "sqeuclidean")
        assert np.array_equal(go, oo) and np.array_equal(gi, oi) and gd.tobytes() == od.tobytes(), k
        assert (np.diff(go) >= min(k, X.shape[0])).all()


@pytest.mark.parametrize("exclude_self", [False, True])
def test_precompute_1000(gpu, oracle, exclude_self):
    from fuel_amd.search import SearchIndex
    X = _grid_with_ties()[:1400]
    si = SearchIndex(X, tree="linear", precompute=1000)
    for k in (10, 400, 1000):
        gi, gd = si.knn(X, k, exclude_self=exclude_self)
        oi, od = oracle.knn(X, X, k, "euclidean", "for_each_knn" if exclude_self else "none")
        assert np.array_equal(gi, oi) and gd.tobytes() == od.tobytes(), k


def test_batch_with_large_k(gpu, oracle):
    from fuel_amd import BatchKnn, workloads
    rng = np.random.default_rng(8)
    probs = [workloads.gaussian_mixture(rng, n, d, 5, -5.0, 5.0) for n, d in [(900, 8), (2500, 32), (400, 64)]]
    b = BatchKnn(probs, 300, mode="bf16")
    b.run()
    for X, (gi, gd) in zip(probs, b.results()):
        oi, od = oracle.knn(X, X, 300, "euclidean", "for_each_knn")
        assert np.array_equal(gi, oi) and gd.tobytes() == od.tobytes(), X.shape
    b.close()
