"""Seeded first-pass thresholds: of the plain MFMA scans (the KP-th smallest of
per-row-group key minima over a strided sample, computed on the matrix cores
before the scan) and of the residual scans (the KP-th smallest upper bound of
the exact distances to a sample of each query's nearest list).  The seed
changes only how many candidates the scan appends:
answers with and without it are identical and equal the oracle's, for every
operand type, every exclusion mode, widening up to the 128-candidate cap and
data with ties, duplicates and clusters."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cases():
    rng = np.random.default_rng(71)
    uni = rng.random((60000, 128), dtype=np.float32)
    g = np.round(rng.random((40000, 16)) * 6).astype(np.float32)  # many exact ties
    from fuel_amd import workloads
    mix = workloads.gaussian_mixture(rng, 50000, 64, 12, -5.0, 5.0)
    dup = np.concatenate([uni[:20000, :40], uni[:20000, :40]])  # every row twice
    return {"uniform128": uni, "grid16": g, "mixture64": mix, "dup40": dup}


@pytest.fixture(scope="module")
def cases(gpu):
    return _cases()


@pytest.mark.parametrize("name", ["uniform128", "grid16", "mixture64", "dup40"])
@pytest.mark.parametrize("mode", ["bf16", "fp32", "bf16r", "fp8r"])
def test_seed_does_not_change_answers(gpu, oracle, cases, name, mode):
    X = cases[name]
    rng = np.random.default_rng(len(name))
    Q = X[rng.choice(X.shape[0], 300, replace=False)] + np.float32(0.01)
    seeded = gpu.Index(X, "euclidean", mode=mode)
    cold = gpu.Index(X, "euclidean", mode=mode, seed=False)
    for k, exclude in [(10, 0), (15, 1), (7, 2)]:
        qq = Q if exclude == 0 else X[:300]
        a = seeded.knn(qq, k, exclude)
        b = cold.knn(qq, k, exclude)
        assert np.array_equal(a[0], b[0]) and a[1].tobytes() == b[1].tobytes(), (name, mode, k, exclude)
        oi, od = oracle.knn(X, qq, k, "euclidean", {0: "none", 1: "for_each_knn", 2: "python_linear"}[exclude])
        assert np.array_equal(a[0], oi) and a[1].tobytes() == od.tobytes(), (name, mode, k, exclude)


@pytest.mark.parametrize("widen", [2, 8, 13])
def test_seed_with_wide_candidate_lists(gpu, oracle, cases, widen):
    X = cases["uniform128"]
    Q = X[::211]
    ix = gpu.Index(X, "sqeuclidean", mode="bf16", widen=widen)
    for k in (1, 10, 64, 120):
        gi, gd = ix.knn(Q, k)
        oi, od = oracle.knn(X, Q, k, "sqeuclidean")
        assert np.array_equal(gi, oi) and gd.tobytes() == od.tobytes(), (widen, k, ix.stats())


@pytest.mark.parametrize("mode", ["bf16r", "fp8r"])
def test_sampled_seed_on_the_serving_shape(gpu, oracle, mode):
    # the serving mixture: the sampled threshold lies near each query's KP-th
    # key, and the answers equal the cold scan's and the oracle's
    from fuel_amd import workloads
    rng = np.random.default_rng(72)
    X = workloads.gaussian_mixture(rng, 150000, 96, 50, -10.0, 10.0)
    Q = workloads.gaussian_mixture(rng, 400, 96, 50, -10.0, 10.0)
    a = gpu.Index(X, "euclidean", mode=mode).knn(Q, 10)
    b = gpu.Index(X, "euclidean", mode=mode, seed=False).knn(Q, 10)
    assert np.array_equal(a[0], b[0]) and a[1].tobytes() == b[1].tobytes()
    oi, od = oracle.knn(X, Q, 10, "euclidean")
    assert np.array_equal(a[0], oi) and a[1].tobytes() == od.tobytes()
