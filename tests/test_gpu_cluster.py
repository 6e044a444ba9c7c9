"""DBSCAN, parallel DBSCAN and OPTICS over the GPU range search against the
oracle's restatements (src/cluster/dbscan.rs, parallel_dbscan.rs, optics.rs)
and the reference's own known answers: identical labels, orderings,
predecessors and distance bits."""
import itertools
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _blobs(seed, n_per=400, d=16, dtype=np.float32):
    rng = np.random.default_rng(seed)
    means = rng.uniform(-10, 10, (6, d))
    X = np.concatenate([m + rng.standard_normal((n_per, d)) for m in means])
    X = np.concatenate([X, rng.uniform(-12, 12, (60, d))])
    return X.astype(dtype)


@pytest.mark.parametrize("variant", ["dbscan", "parallel"])
@pytest.mark.parametrize("eps,min_points", [(4.0, 5), (5.5, 10), (3.0, 1), (7.0, 40)])
def test_dbscan_blobs(gpu, oracle, variant, eps, min_points):
    from fuel_amd.cluster import dbscan
    X = _blobs(1)
    got = dbscan(X, eps, min_points, variant=variant)
    exp = oracle.dbscan(X, eps, min_points, variant=variant)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("variant", ["dbscan", "parallel"])
def test_dbscan_reference_known_answers(gpu, variant):
    from fuel_amd.cluster import dbscan
    X = np.array([[0.0], [2.0], [3.0], [4.0], [6.0], [8.0], [10.0]])
    cases = [(1, [0, 1, 1, 1, 2, 3, 4]), (2, [-1, 0, 0, 0, -1, -1, -1]), (3, [-1, 0, 0, 0, -1, -1, -1]),
             (4, [-1] * 7)]
    for min_points, expected in cases:
        assert list(dbscan(X, 1.0, min_points, variant=variant)) == expected


def test_dbscan_grid_boundary_inclusive(gpu, oracle):
    # eps equal to grid spacings: the inclusive predicate decides every border
    from fuel_amd.cluster import dbscan
    X = np.array(list(itertools.product(range(12), range(12))), np.float32)
    X = X[np.random.default_rng(3).permutation(X.shape[0])]
    for eps in (1.0, np.sqrt(2.0), 2.0):
        for mp in (3, 5, 9):
            for variant in ("dbscan", "parallel"):
                assert np.array_equal(dbscan(X, eps, mp, variant=variant), oracle.dbscan(X, eps, mp, variant=variant))


def _same(a, b):
    return a.dtype == b.dtype and a.tobytes() == b.tobytes()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("max_eps,min_points", [(5.0, 5), (np.inf, 10), (3.5, 3)])
def test_optics_blobs(gpu, oracle, dtype, max_eps, min_points):
    from fuel_amd.cluster import optics
    X = _blobs(2, n_per=150, d=8, dtype=dtype)
    res = optics(X, max_eps, min_points)
    o, r, c, p, l = oracle.optics(X, max_eps, min_points)
    assert np.array_equal(res.ordering(), o)
    assert _same(res.reachability(), r)
    assert _same(res.core_distance(), c)  # NaN bits of non-core points included
    assert np.array_equal(res.predecessor(), p)
    assert np.array_equal(res.labels(), l)
    for xi, mp in ((0.05, 5), (0.2, 10)):
        assert np.array_equal(res.extract_xi(xi, mp), oracle.extract_xi(o, r, p, xi, mp))


def test_optics_elki_reference(gpu):
    from fuel_amd.cluster import optics
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "optics_elki_60.json")))
    res = optics(np.array(g["points"]), float("inf"), g["min_points"])
    assert list(res.ordering()) == g["expected_ordering"]
    exp = np.array([np.inf if v is None else v for v in g["expected_reachability_by_order"]])
    got = res.reachability()[res.ordering()]
    assert np.array_equal(np.isinf(got), np.isinf(exp))
    assert np.max(np.abs(got[~np.isinf(exp)] - exp[~np.isinf(exp)])) < 1e-9


def test_optics_sklearn_processing_order(gpu):
    from fuel_amd.cluster import optics
    res = optics(np.array([[0.0], [10.0], [-10.0], [25.0]]), 15.0, 3)
    assert list(res.ordering()) == [0, 1, 2, 3]
    r = res.reachability()
    assert np.isinf(r[0]) and list(r[1:]) == [10.0, 10.0, 15.0]
    c = res.core_distance()
    assert list(c[:2]) == [10.0, 15.0] and np.isnan(c[2:]).all()


def test_cluster_argument_errors(gpu):
    from fuel_amd.cluster import dbscan, optics
    X = np.zeros((4, 2), np.float32)
    with pytest.raises(ValueError):
        dbscan(X, -1.0, 3)
    with pytest.raises(ValueError):
        dbscan(X, 1.0, 3, variant="hdbscan")
    with pytest.raises(ValueError):
        optics(X, 1.0, 0)
