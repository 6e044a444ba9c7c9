"""The FP8 E4M3 resident copy (FUEL_AMD_MODE_FP8R): residual rows as E4M3 codes
with a power-of-two scale per row, the query side split into hi + lo E4M3
parts, both MFMAs of v_mfma_scale_f32_32x32x64_f8f6f4 against the same row
fragment, and the per-list bound folded into the C input.  Parity with the
oracle (identical indices and f32 bits) on clustered, uniform, integer and
wildly scaled data; the first-level certification rate on the serving shape
(which also fails if the E4M3 keys were not lower bounds, or far too loose);
and answers independent of the batch they ride in."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mixture(rng, n, d, comps, lo, hi):
    from fuel_amd import workloads
    return workloads.gaussian_mixture(rng, n, d, comps, lo, hi)


def _check(gpu, oracle, X, Q, k, metric="euclidean", exclude=0, widen=0):
    ix = gpu.Index(X, metric, mode="fp8r", widen=widen)
    gi, gd = ix.knn(Q, k, exclude)
    st = ix.stats()
    oi, od = oracle.knn(X, Q, k, metric, exclude)
    bad = np.argwhere(np.any(gi != oi, axis=1)).ravel()
    assert bad.size == 0, (bad[:5], st)
    assert gd.tobytes() == od.tobytes()
    return st


@pytest.mark.parametrize("d", [8, 24, 64, 96, 128, 200, 256])
@pytest.mark.parametrize("metric", ["euclidean", "sqeuclidean"])
def test_clustered_parity_and_rate(gpu, oracle, d, metric):
    rng = np.random.default_rng(200 + d)
    means = rng.uniform(-10, 10, size=(50, d))
    X = (means[rng.integers(0, 50, 40000)] + rng.standard_normal((40000, d))).astype(np.float32)
    Q = (means[rng.integers(0, 50, 400)] + rng.standard_normal((400, d))).astype(np.float32)
    st = _check(gpu, oracle, X, Q, 10, metric)
    assert st["mode_used"] == 5, st
    assert st["kp"] == 160, st  # widen 16 by default
    assert st["fallback_queries"] <= 0.10 * Q.shape[0], st


def test_serving_shape_rate(gpu, oracle):
    # the serving mixture (means U[-10, 10]^96, 50 components) at 200k rows:
    # nearly every query certifies at the first level with KP = 160
    rng = np.random.default_rng(66)
    X = _mixture(rng, 200_000, 96, 50, -10.0, 10.0)
    Q = _mixture(rng, 600, 96, 50, -10.0, 10.0)
    st = _check(gpu, oracle, X, Q, 10)
    assert st["fallback_queries"] <= 0.03 * Q.shape[0], st
    assert st["exact_queries"] == 0, st


@pytest.mark.parametrize("exclude", [1, 2])
@pytest.mark.parametrize("k", [1, 15, 40])
def test_self_knn_exclusion(gpu, oracle, exclude, k):
    rng = np.random.default_rng(17 + k)
    X = _mixture(rng, 12000, 64, 12, -5.0, 5.0)
    _check(gpu, oracle, X, X[::37], k, "euclidean", exclude)


@pytest.mark.parametrize("n", [1, 63, 64, 65, 511, 512, 1000, 5001])
def test_list_padding_and_small_n(gpu, oracle, n):
    rng = np.random.default_rng(50 + n)
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
    rng = np.random.default_rng(15)
    X = rng.random((30000, 128), dtype=np.float32)
    _check(gpu, oracle, X, rng.random((300, 128), dtype=np.float32), 10)
    from fuel_amd import workloads
    C = workloads.c3(n=6000)[:, :256]  # the E4M3 copy takes d <= 256
    _check(gpu, oracle, C, C[:120], 20, "euclidean", 1)


def test_rows_of_very_different_scales(gpu, oracle):
    # clusters whose spreads differ by 2^-20 .. 2^20: every row gets its own
    # power-of-two scale, so the tiny clusters keep their E4M3 precision
    rng = np.random.default_rng(21)
    d, parts = 48, []
    for j, sc in enumerate(np.exp2(np.linspace(-20, 20, 9))):
        c = rng.uniform(-5, 5, d) * sc
        parts.append(c + sc * rng.standard_normal((1500, d)))
    X = np.concatenate(parts).astype(np.float32)
    Q = X[::97] + np.float32(1e-3) * X[::97]
    _check(gpu, oracle, X, Q, 10)
    _check(gpu, oracle, X, X[::53], 6, "sqeuclidean", 1)


def test_widening_and_chain(gpu, oracle):
    # a narrow first level (widen 2) sends many queries to the certified second
    # pass; answers stay exact
    rng = np.random.default_rng(23)
    X = _mixture(rng, 30000, 96, 50, -10.0, 10.0)
    Q = _mixture(rng, 300, 96, 50, -10.0, 10.0)
    st = _check(gpu, oracle, X, Q, 10, widen=2)
    assert st["kp"] == 20 and st["fallback_queries"] > 0, st


def test_result_independent_of_batch_composition(gpu):
    rng = np.random.default_rng(19)
    X = _mixture(rng, 20000, 96, 50, -10.0, 10.0)
    Q = _mixture(rng, 700, 96, 50, -10.0, 10.0)
    ix = gpu.Index(X, "euclidean", mode="fp8r")
    ai, ad = ix.knn(Q, 10)
    for lo, hi in [(0, 1), (5, 70), (300, 700)]:
        bi, bd = ix.knn(Q[lo:hi], 10)
        assert np.array_equal(ai[lo:hi], bi) and ad[lo:hi].tobytes() == bd.tobytes()


def test_wider_than_the_e4m3_variants_falls_back_to_bf16r(gpu, oracle):
    rng = np.random.default_rng(25)
    X = _mixture(rng, 5000, 300, 8, -5.0, 5.0)
    st = _check(gpu, oracle, X, X[:40], 5)
    assert st["mode_used"] == 4, st
