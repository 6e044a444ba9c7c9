"""The GPU's correctly rounded square root (fuel's f32::sqrt,
src/distance/euclidean.rs:7-9) on 10^8 distinct inputs: the 'euclidean'
pairwise matrix of 10,000 rows must equal numpy's IEEE f32 sqrt of the
'sqeuclidean' matrix of the same rows bit for bit (both come from the same
restated sum, so any difference is the sqrt).  Rows mix magnitudes from 2^-30
to 2^30 so that the inputs cover many exponents and every mantissa pattern is
likely."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_sqrt_on_1e8_squared_distances(gpu):
    rng = np.random.default_rng(91)
    n, d = 10_000, 16
    X = (rng.standard_normal((n, d)) * np.exp2(rng.uniform(-30, 30, (n, 1)))).astype(np.float32)
    s = gpu.pairwise(X, "sqeuclidean", "square")
    e = gpu.pairwise(X, "euclidean", "square")
    assert s.shape == (n, n) and e.shape == (n, n)
    ref = np.sqrt(s)  # IEEE f32 square root, round to nearest even
    bad = np.argwhere(ref.view(np.uint32) != e.view(np.uint32))
    assert bad.size == 0, (bad[:5], s[tuple(bad[0])] if bad.size else None)
    # the inputs really are varied: many distinct exponents and mantissas
    bits = s.view(np.uint32).ravel()[::97]
    assert len(np.unique(bits >> 23)) > 100 and len(np.unique(bits & 0x7FFFFF)) > 500_000
