"""Parity of the GPU exact path (restated distance + fused selection) with the
oracle, through the C ABI: bit-exact indices and distances, fuel's tie order,
padding and exclusion semantics, range search, pairwise matrices, LOF."""
import itertools

import numpy as np
import pytest

from conftest import load_gaussian4d, load_reference_scores

pytestmark = pytest.mark.gpu


def _same(a, b):
    return a.shape == b.shape and a.dtype == b.dtype and a.tobytes() == b.tobytes()


def _knn_case(gpu, oracle, X, Q, k, metric, exclude, mode="exact"):
    ix = gpu.Index(X, metric, mode=mode)
    gi, gd = ix.knn(Q, k, exclude)
    oi, od = oracle.knn(X, Q, k, metric, exclude)
    assert np.array_equal(gi, oi), (np.argwhere(gi != oi)[:5], gi[gi != oi][:5], oi[gi != oi][:5])
    assert _same(gd, od)
    return ix


@pytest.mark.parametrize("metric", ["euclidean", "sqeuclidean"])
@pytest.mark.parametrize("exclude", [0, 1, 2])
def test_c1_self_knn_exact(gpu, oracle, metric, exclude):
    from fuel_amd import workloads
    X = workloads.c1(n=3000)
    _knn_case(gpu, oracle, X, X, 10, metric, exclude)


@pytest.mark.parametrize("d", [1, 5, 8, 13, 16, 31, 64, 100, 128, 600])
def test_dims(gpu, oracle, d):
    rng = np.random.default_rng(d)
    X = rng.standard_normal((1500, d)).astype(np.float32)
    Q = rng.standard_normal((70, d)).astype(np.float32)
    _knn_case(gpu, oracle, X, Q, 7, "euclidean", 0)


def test_large_d_chunked_staging(gpu, oracle):
    rng = np.random.default_rng(5)
    X = rng.standard_normal((700, 1100)).astype(np.float32)
    Q = rng.standard_normal((40, 1100)).astype(np.float32)
    _knn_case(gpu, oracle, X, Q, 5, "sqeuclidean", 0)


@pytest.mark.parametrize("k", [1, 2, 3, 10, 33, 64, 100])
def test_ties_on_integer_grid(gpu, oracle, k):
    X = np.array(list(itertools.product(range(12), range(12), range(3))), np.float32)
    _knn_case(gpu, oracle, X, X, k, "euclidean", 2)
    _knn_case(gpu, oracle, X, X, k, "sqeuclidean", 1)


def test_duplicates_and_small_n(gpu, oracle):
    X = np.array([[0, 0], [0, 0], [1, 0], [1, 0], [2, 0]], np.float32)
    for k in range(0, 9):
        for ex in (0, 1, 2):
            if k == 0:
                ix = gpu.Index(X)
                gi, gd = ix.knn(X, 0, ex)
                assert gi.shape == (5, 0)
                continue
            _knn_case(gpu, oracle, X, X, k, "euclidean", ex)


def test_reference_known_answer(gpu):
    # src/search/linear_scan.rs:247-265
    X = np.array([[0, 0], [1, 0], [2, 0]], np.float32)
    gi, gd = gpu.Index(X).knn(np.array([[1, 0]], np.float32), 3, 0)
    assert gi.tolist() == [[1, 0, 2]]
    assert gd.tolist() == [[0.0, 1.0, 1.0]]


def test_f64(gpu, oracle):
    rng = np.random.default_rng(9)
    X = rng.standard_normal((2000, 19))
    Q = rng.standard_normal((65, 19))
    _knn_case(gpu, oracle, X, Q, 12, "euclidean", 0)
    _knn_case(gpu, oracle, X, X, 12, "sqeuclidean", 2)


def test_many_queries_multiple_splits(gpu, oracle):
    rng = np.random.default_rng(11)
    X = (rng.random((20000, 24)) * 4).round(1).astype(np.float32)  # many ties
    Q = (rng.random((300, 24)) * 4).round(1).astype(np.float32)
    _knn_case(gpu, oracle, X, Q, 15, "euclidean", 0)


def test_device_entry_point_with_torch_buffers(gpu, oracle):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(12)
    X = rng.standard_normal((4000, 40)).astype(np.float32)
    Q = rng.standard_normal((257, 40)).astype(np.float32)
    ix = gpu.Index(X, "euclidean", mode="exact")
    q = torch.from_numpy(Q).cuda()
    oi = torch.empty((257, 9), dtype=torch.int64, device="cuda")
    od = torch.empty((257, 9), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    ix.knn_device(q.data_ptr(), 257, 40, 9, 0, oi.data_ptr(), od.data_ptr(), 0)
    torch.cuda.synchronize()
    ri, rd = oracle.knn(X, Q, 9)
    assert np.array_equal(oi.cpu().numpy(), ri)
    assert _same(od.cpu().numpy(), rd)


# -- range --------------------------------------------------------------------

@pytest.mark.parametrize("metric", ["euclidean", "sqeuclidean"])
def test_range_parity(gpu, oracle, metric):
    from fuel_amd import workloads
    X = workloads.c1(n=2500)
    Q = X[:300]
    r = 4.0 if metric == "euclidean" else 16.0
    ix = gpu.Index(X, metric)
    for ex in (False, True):
        go, gi, gd = ix.range(Q, r, ex)
        oo, oi, od = oracle.range_search(X, Q, r, metric, ex)
        assert np.array_equal(go, oo)
        assert np.array_equal(gi, oi)
        assert _same(gd, od)


def test_range_ties_and_boundary(gpu, oracle):
    X = np.array(list(itertools.product(range(9), range(9))), np.float32)
    ix = gpu.Index(X)
    for r in (0.0, 1.0, 1.4142135, np.sqrt(2.0), 2.0, 3.5):
        go, gi, gd = ix.range(X, r, False)
        oo, oi, od = oracle.range_search(X, X, r)
        assert np.array_equal(go, oo) and np.array_equal(gi, oi) and _same(gd, od)


def test_range_known_answer_f64(gpu):
    X = np.array([[0, 0], [1, 0], [2, 0]], np.float64)
    o, i, d = gpu.Index(X).range(np.array([[0.5, 0]]), 1.0)
    assert i.tolist() == [0, 1]


# -- pairwise -------------------------------------------------------------------

@pytest.mark.parametrize("form", ["square", "condensed"])
@pytest.mark.parametrize("metric", ["euclidean", "sqeuclidean"])
def test_pairwise_c1(gpu, oracle, form, metric):
    from fuel_amd import workloads
    X = workloads.c1()[:2000]
    g = gpu.pairwise(X, metric, form)
    o = oracle.pairwise(X, metric, form)
    assert _same(g, o)


def test_pairwise_known_answers(gpu):
    p = np.array([[0, 1], [2, 3], [1, 4]], np.float32)
    assert gpu.pairwise(p, "sqeuclidean").ravel().tolist() == [0, 8, 10, 8, 0, 2, 10, 2, 0]
    p = np.array([[1, 2, 3, 4], [4, 3, 2, 1], [0, 0, 0, 0]], np.float64)
    assert gpu.pairwise(p, "sqeuclidean").ravel().tolist() == [0, 20, 30, 20, 0, 30, 30, 30, 0]


# -- LOF / kNN outlier ----------------------------------------------------------

def test_lof_knn_outlier_elki(gpu, oracle):
    X = load_gaussian4d()
    ref = load_reference_scores()
    ix = gpu.Index(X)
    np.testing.assert_allclose(ix.lof(10), ref["LOF-10"], rtol=0, atol=1e-6)
    np.testing.assert_allclose(ix.knn_outlier(10), ref["KNN-10"], rtol=0, atol=1e-6)
    assert _same(ix.lof(10), oracle.lof(X, 10))
    X32 = X.astype(np.float32)
    assert _same(gpu.Index(X32).lof(10), oracle.lof(X32, 10))


def test_lof_closed_form_and_degenerate(gpu):
    P = np.array([[1, 1], [1, 2], [2, 1]], np.float64)
    s = gpu.Index(P).lof(2)
    r2 = np.sqrt(2.0)
    assert abs(s[0] - 2 * r2 / (1 + r2)) < 1e-12
    assert gpu.Index(np.zeros((1, 3))).lof(5).tolist() == [1.0]
    assert gpu.Index(np.zeros((6, 2))).lof(3).tolist() == [1.0] * 6


def test_lof_c3_like_subset(gpu, oracle):
    from fuel_amd import workloads
    X = workloads.c3(n=3000)
    g = gpu.Index(X).lof(20)
    o = oracle.lof(X, 20)
    assert _same(g, o)


# -- Python mirror ------------------------------------------------------------------

def test_search_index_python_api(gpu, oracle):
    import fuel_amd
    from fuel_amd import workloads
    X = workloads.c1(n=2000)
    si = fuel_amd.SearchIndex(X, tree="linear")
    gi, gd = si.knn(X, 10)  # same source -> exclude_self (python linear semantics)
    oi, od = oracle.knn(X, X, 10, "euclidean", "python_linear")
    assert np.array_equal(gi, oi) and _same(gd, od)
    gi2, _ = si.knn(X.copy(), 10)  # different buffer -> no exclusion
    oi2, _ = oracle.knn(X, X, 10)
    assert np.array_equal(gi2, oi2)
    with pytest.raises(ValueError):
        si.knn(X.copy(), 3, exclude_self=True)
    idx_list, dist_list = si.radius_search(X[:50], 3.0, exclude_self=False)
    oo, oi3, od3 = oracle.range_search(X, X[:50], 3.0)
    for q in range(50):
        assert idx_list[q].tolist() == oi3[oo[q]:oo[q + 1]].tolist()
    res = fuel_amd.local_outlier_factor(X, 10)
    assert _same(res.scores, oracle.lof(X, 10))
    assert res.metadata["label"] == "LOF"
