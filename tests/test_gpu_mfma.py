"""Parity of the MFMA candidate modes (bf16 and f32 matrix cores + exact rerank +
certified fallback) with the oracle: identical indices and f32 bit patterns.
The fallback counters are checked too: on continuous data nearly every query
must certify, which fails if the MFMA tiles computed the wrong inner products
(a wrong lane layout would either break parity or send every query to the
fallback)."""
import itertools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MODES = ["bf16", "fp32", "bf16r", "fp8r"]


def _same(a, b):
    return a.shape == b.shape and a.dtype == b.dtype and a.tobytes() == b.tobytes()


def _run(gpu, oracle, X, Q, k, metric="euclidean", exclude=0, mode="bf16", max_fallback_frac=None):
    ix = gpu.Index(X, metric, mode=mode)
    gi, gd = ix.knn(Q, k, exclude)
    st = ix.stats()
    oi, od = oracle.knn(X, Q, k, metric, exclude)
    bad = np.argwhere(np.any(gi != oi, axis=1)).ravel()
    assert bad.size == 0, (mode, bad[:5], gi[bad[:2]], oi[bad[:2]], st)
    assert _same(gd, od)
    if max_fallback_frac is not None:
        assert st["fallback_queries"] <= max_fallback_frac * Q.shape[0], st
    return st


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("exclude", [0, 1, 2])
def test_c1_like(gpu, oracle, mode, exclude):
    from fuel_amd import workloads
    X = workloads.c1(n=5000)
    # clusters far from the origin: plain bf16 rounding (relative to |x|) is too
    # coarse to certify most queries; the f32 form and the residual operand
    # (rounding relative to the distance to a centroid) are held to a rate
    frac = {"fp32": 0.02, "bf16r": 0.10, "fp8r": 0.10}.get(mode)
    st = _run(gpu, oracle, X, X, 10, "euclidean", exclude, mode, max_fallback_frac=frac)
    assert st["mode_used"] == {"bf16": 3, "fp32": 2, "bf16r": 4, "fp8r": 5}[mode]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("d", [1, 3, 8, 16, 31, 64, 96, 128, 200, 256])
def test_dims_uniform(gpu, oracle, mode, d):
    rng = np.random.default_rng(100 + d)
    X = rng.random((6000, d), dtype=np.float32)
    Q = rng.random((300, d), dtype=np.float32)
    frac = 0.05 if d >= 8 else None
    _run(gpu, oracle, X, Q, 10, "euclidean", 0, mode, max_fallback_frac=frac)


@pytest.mark.parametrize("mode", MODES)
def test_sqeuclidean_gaussian(gpu, oracle, mode):
    rng = np.random.default_rng(7)
    X = (rng.standard_normal((8000, 48)) * 3 + 1).astype(np.float32)
    Q = (rng.standard_normal((500, 48)) * 3 + 1).astype(np.float32)
    _run(gpu, oracle, X, Q, 15, "sqeuclidean", 0, mode, max_fallback_frac=0.02)


@pytest.mark.parametrize("mode", MODES)
def test_ties_grid_exact_through_fallback(gpu, oracle, mode):
    # heavy ties at the boundary: certification must refuse them and the exact
    # fallback must restore fuel's index order
    X = np.array(list(itertools.product(range(10), range(10), range(10))), np.float32)
    _run(gpu, oracle, X, X, 6, "euclidean", 2, mode)
    _run(gpu, oracle, X, X[::7], 9, "sqeuclidean", 0, mode)


@pytest.mark.parametrize("mode", MODES)
def test_integer_c3_like(gpu, oracle, mode):
    from fuel_amd import workloads
    X = workloads.c3(n=2500, d=784 if mode == "bf16" else 256)
    st = _run(gpu, oracle, X, X, 20, "euclidean", 1, mode, max_fallback_frac=0.05)


def test_c2_like_subset(gpu, oracle):
    from fuel_amd import workloads
    X, Q = workloads.c2(n=200_000, m=256)
    st_b = _run(gpu, oracle, X, Q, 10, "euclidean", 0, "bf16")
    st_f = _run(gpu, oracle, X, Q, 10, "euclidean", 0, "fp32", max_fallback_frac=0.0)
    print("c2 subset fallback: bf16", st_b["fallback_queries"], "fp32", st_f["fallback_queries"], flush=True)
    assert st_b["fallback_queries"] <= 0.25 * 256


@pytest.mark.parametrize("mode", MODES)
def test_small_n_and_large_k(gpu, oracle, mode):
    rng = np.random.default_rng(3)
    X = rng.standard_normal((37, 12)).astype(np.float32)
    for k in (1, 5, 36, 37, 50, 100):
        _run(gpu, oracle, X, X, k, "euclidean", 0, mode)
        _run(gpu, oracle, X, X, k, "euclidean", 2, mode)


def test_batches_larger_than_one_chunk(gpu, oracle):
    rng = np.random.default_rng(4)
    X = rng.random((3000, 20), dtype=np.float32)
    Q = rng.random((5000, 20), dtype=np.float32)
    _run(gpu, oracle, X, Q, 4, "euclidean", 0, "bf16", max_fallback_frac=0.02)


def test_result_independent_of_batch_composition(gpu):
    rng = np.random.default_rng(5)
    X = rng.random((20000, 64), dtype=np.float32)
    Q = rng.random((700, 64), dtype=np.float32)
    ix = gpu.Index(X, "euclidean", mode="bf16")
    gi, gd = ix.knn(Q, 10, 0)
    for sl in (slice(0, 1), slice(13, 14), slice(100, 357), slice(600, 700)):
        si, sd = ix.knn(Q[sl], 10, 0)
        assert np.array_equal(si, gi[sl]) and _same(sd, gd[sl])


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("n", [5000, 6000, 6001, 6047])
def test_neighbours_in_partial_last_tile(gpu, oracle, mode, n):
    # the last tile mixes real rows with padding rows; queries sitting on the
    # last rows must find them (a padding row's sentinel must not hide its
    # neighbours in the same accumulator block)
    rng = np.random.default_rng(n)
    X = rng.random((n, 16), dtype=np.float32)
    Q = X[-60:] + np.float32(1e-3) * rng.standard_normal((60, 16)).astype(np.float32)
    _run(gpu, oracle, X, Q, 10, "euclidean", 0, mode)
    _run(gpu, oracle, X, X[-60:], 5, "euclidean", 1, mode)


@pytest.mark.parametrize("mode", MODES)
def test_huge_norms_take_the_exact_scan(gpu, oracle, mode):
    # squared norms beyond the scan's sentinel range: every query is answered
    # by the exact scan, still bit-identical
    rng = np.random.default_rng(11)
    X = (rng.random((3000, 8)) * 2e18).astype(np.float32)
    Q = (rng.random((50, 8)) * 2e18).astype(np.float32)
    st = _run(gpu, oracle, X, Q, 7, "euclidean", 0, mode)
    assert st["exact_queries"] == 50, st


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("m", [1, 63, 65, 191, 193, 255, 257])
def test_partial_query_blocks(gpu, oracle, mode, m):
    # a launch fills whole query blocks (64 slots for one-wave workgroups, 256
    # otherwise); the padding slots start below every key and keep nothing, and
    # the real queries of the last block are answered as if alone, at every level
    # of the chain (clustered rows send some of them to the second pass)
    from fuel_amd import workloads
    rng = np.random.default_rng(1000 + m)
    X = workloads.gaussian_mixture(rng, 30_000, 48, 12, -6.0, 6.0)
    Q = workloads.gaussian_mixture(rng, m, 48, 12, -6.0, 6.0)
    _run(gpu, oracle, X, Q, 10, "euclidean", 0, mode)
    ix = gpu.Index(X, "euclidean", mode=mode)
    li, ld = ix.knn(Q[-1:], 10, 0)  # the last query alone
    gi, gd = ix.knn(Q, 10, 0)
    assert np.array_equal(li[0], gi[-1]) and ld[0].tobytes() == gd[-1].tobytes()
