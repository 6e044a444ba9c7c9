"""Indexes built with FUEL_AMD_FLAG_FALLBACK_ARITH reproduce fuel's portable
distance kernel (src/distance/squared_euclidean.rs:99-133: 8 lanes of
fl(t*t) + acc, a left fold from +0, an unfused tail), the one fuel runs on
hosts without AVX+FMA.  Every entry point of such an index -- kNN in every
mode, kNN with ties, range, pairwise, LOF -- must equal the oracle run with the
same arithmetic bit for bit, on scale-mixed data where the two roundings
differ (which the last test shows on the GPU itself)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _data(seed, n, d, dtype):
    rng = np.random.default_rng(seed)
    scale = rng.choice([1e-2, 1.0, 30.0], size=(1, d))
    X = (rng.standard_normal((n, d)) * scale).astype(dtype)
    Q = (rng.standard_normal((97, d)) * scale).astype(dtype)
    return X, Q


@pytest.mark.parametrize("mode", ["exact", "fp32", "bf16", "bf16r"])
@pytest.mark.parametrize("d", [7, 13, 64, 100])
def test_knn_fallback_f32(gpu, oracle, mode, d):
    X, Q = _data(10 + d, 6000, d, np.float32)
    ix = gpu.Index(X, "euclidean", mode=mode, arith="fallback")
    for exclude, name in [(0, "none"), (1, "for_each_knn")]:
        qq = Q if exclude == 0 else X[::61]
        gi, gd = ix.knn(qq, 10, exclude)
        oi, od = oracle.knn(X, qq, 10, "euclidean", name, arith="fallback")
        assert np.array_equal(gi, oi), (mode, d, exclude, ix.stats())
        assert gd.tobytes() == od.tobytes()


@pytest.mark.parametrize("metric", ["euclidean", "sqeuclidean"])
@pytest.mark.parametrize("d", [5, 9, 40])
def test_knn_fallback_f64(gpu, oracle, metric, d):
    X, Q = _data(20 + d, 3000, d, np.float64)
    ix = gpu.Index(X, metric, mode="exact", arith="fallback")
    gi, gd = ix.knn(Q, 12)
    oi, od = oracle.knn(X, Q, 12, metric, arith="fallback")
    assert np.array_equal(gi, oi)
    assert gd.tobytes() == od.tobytes()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_ties_range_pairwise_lof_fallback(gpu, oracle, dtype):
    X, Q = _data(33, 1500, 11, dtype)
    ix = gpu.Index(X, "euclidean", mode="exact", arith="fallback")
    go, gi, gd = ix.knn_ties(Q, 7)
    oo, oi, od = oracle.knn_ties(X, Q, 7, arith="fallback")
    assert np.array_equal(go, oo) and np.array_equal(gi, oi) and gd.tobytes() == od.tobytes()
    r = float(np.median(od))
    go, gi, gd = ix.range(Q, r)
    oo, oi, od = oracle.range_search(X, Q, r, arith="fallback")
    assert np.array_equal(go, oo) and np.array_equal(gi, oi) and gd.tobytes() == od.tobytes()
    for form in ("square", "condensed"):
        got = ix.pairwise(form)
        exp = oracle.pairwise(X, "euclidean", form, arith="fallback")
        assert got.tobytes() == exp.tobytes(), form
    got = ix.lof(10)
    exp = oracle.lof(X, 10, arith="fallback")
    assert got.tobytes() == exp.tobytes()


def test_the_two_roundings_differ_on_this_data(gpu, oracle):
    X, Q = _data(44, 4000, 100, np.float32)
    fb = gpu.Index(X, "sqeuclidean", mode="exact", arith="fallback")
    av = gpu.Index(X, "sqeuclidean", mode="exact")
    _, dfb = fb.knn(Q, 10)
    _, dav = av.knn(Q, 10)
    assert dfb.tobytes() != dav.tobytes()
    _, oav = oracle.knn(X, Q, 10, "sqeuclidean")
    assert dav.tobytes() == oav.tobytes()
    assert np.allclose(dfb, dav, rtol=1e-5)
