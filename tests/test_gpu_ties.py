"""kNN with fuel's tie semantics (KNNHeap keeps every point at the k-th
distance) and the GPU-filled precomputed kNN table of
PrecomputedKnnSearcher (src/search/precomputed.rs:36-90), against the oracle:
identical CSR offsets, indices and distance bits."""
import itertools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same_csr(a, b):
    return all(x.shape == y.shape and x.dtype == y.dtype and x.tobytes() == y.tobytes() for x, y in zip(a, b))


@pytest.mark.parametrize("metric", ["euclidean", "sqeuclidean"])
@pytest.mark.parametrize("k", [0, 1, 3, 7, 20])
def test_knn_ties_integer_grid(gpu, oracle, metric, k):
    # integer grid: many exact ties at the k-th distance
    X = np.array(list(itertools.product(range(8), range(8), range(4))), np.float32)
    ix = gpu.Index(X, metric)
    got = ix.knn_ties(X[::5], k)
    exp = oracle.knn_ties(X, X[::5], k, metric)
    assert _same_csr(got, exp)
    if k > 0:  # at least min(k, n) per query, more where the k-th distance is tied
        counts = np.diff(got[0])
        assert np.all(counts >= min(k, X.shape[0]))


@pytest.mark.parametrize("mode", ["exact", "fp32", "bf16"])
def test_knn_ties_continuous(gpu, oracle, mode):
    rng = np.random.default_rng(21)
    X = rng.random((7000, 24), dtype=np.float32)
    X[100:110] = X[5]  # duplicates: ties at distance 0 and beyond
    Q = X[:300]
    ix = gpu.Index(X, "euclidean", mode=mode)
    assert _same_csr(ix.knn_ties(Q, 10), oracle.knn_ties(X, Q, 10))


def test_knn_ties_k_beyond_n_and_f64(gpu, oracle):
    rng = np.random.default_rng(22)
    X = rng.standard_normal((40, 5))
    ix = gpu.Index(X, "euclidean")
    for k in (39, 40, 41, 100):
        assert _same_csr(ix.knn_ties(X, k), oracle.knn_ties(X, X, k))


@pytest.mark.parametrize("exclude_self", [False, True])
def test_precomputed_search_index(gpu, oracle, exclude_self):
    from fuel_amd.search import SearchIndex
    X = np.array(list(itertools.product(range(6), range(6), range(6))), np.float32)
    X = np.concatenate([X, X[:20]])  # duplicates
    si = SearchIndex(X, tree="linear", precompute=12)
    plain = SearchIndex(X, tree="linear")
    for k in (1, 5, 11, 12, 20):
        gi, gd = si.knn(X, k, exclude_self=exclude_self)
        # precomputed arm: k (+1) with ties, drop self, truncate to k
        oi, od = oracle.knn(X, X, k, "euclidean", "for_each_knn" if exclude_self else "none")
        assert np.array_equal(gi, oi) and gd.tobytes() == od.tobytes(), k
        # queries from another array take the plain scan
        Q = X.copy()
        qi, qd = si.knn(Q, k)
        pi, pd = plain.knn(Q, k)
        assert np.array_equal(qi, pi) and qd.tobytes() == pd.tobytes()


def test_subset_with_ties_matches_table(gpu, oracle):
    # the table row cut by subset_with_ties equals a kNN-with-ties at the smaller k
    rng = np.random.default_rng(23)
    X = np.round(rng.random((500, 3)) * 6).astype(np.float32)
    ix = gpu.Index(X, "sqeuclidean")
    offs, idx, dist = ix.knn_ties(X, 15)
    for k in (1, 4, 9, 15):
        so, si, sd = oracle.knn_ties(X, X, k, "sqeuclidean")
        for i in range(0, 500, 37):
            row = dist[offs[i]:offs[i + 1]]
            cut = oracle.subset_with_ties(row, k)
            assert np.array_equal(idx[offs[i]:offs[i] + cut], si[so[i]:so[i + 1]])
