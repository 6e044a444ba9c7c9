"""fuel's math::pairwise_sqdist (the k-means distance matrix) on the GPU
against the oracle: identical bits in the reference's rounding order (one
sequential fused chain per output for d >= 4, rounded squares below)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("m,n,d", [(1, 1, 1), (3, 7, 2), (10, 1000, 3), (64, 64, 4), (65, 130, 5),
                                   (100, 3000, 16), (7, 5000, 31), (257, 300, 64), (33, 70, 100), (16, 2000, 256)])
def test_pairwise_sqdist_matches_oracle(gpu, oracle, dtype, m, n, d):
    import fuel_amd
    rng = np.random.default_rng(m * 1000 + n + d)
    A = (rng.standard_normal((m, d)) * rng.choice([1e-2, 1.0, 30.0], (m, d))).astype(dtype)
    B = (rng.standard_normal((n, d)) * rng.choice([1e-2, 1.0, 30.0], (n, d))).astype(dtype)
    got = fuel_amd.pairwise_sqdist(A, B)
    exp = oracle.kmeans_pairwise(A, B)
    assert got.shape == (m, n) and got.tobytes() == exp.tobytes()


def test_pairwise_sqdist_reference_matrices(gpu):
    import fuel_amd
    p = np.array([[0.0, 1.0], [2.0, 3.0], [1.0, 4.0]], np.float32)
    assert fuel_amd.pairwise_sqdist(p, p).ravel().tolist() == [0, 8, 10, 8, 0, 2, 10, 2, 0]
    q = np.array([[1, 2, 3, 4], [4, 3, 2, 1], [0, 0, 0, 0]], np.float64)
    assert fuel_amd.pairwise_sqdist(q, q).ravel().tolist() == [0, 20, 30, 20, 0, 30, 30, 30, 0]


def test_pairwise_sqdist_device_entry(gpu, oracle):
    # k-means loop shape: centres x points on resident buffers with row pitches
    import ctypes

    import torch
    from fuel_amd import _lib
    rng = np.random.default_rng(9)
    pts = rng.standard_normal((5000, 24)).astype(np.float32)
    cen = pts[rng.choice(5000, 12, replace=False)].copy()
    dev = torch.device("cuda:0")
    P = torch.zeros((5000, 32), dtype=torch.float32, device=dev)
    P[:, :24] = torch.from_numpy(pts).to(dev)
    Cn = torch.from_numpy(cen).to(dev)
    out = torch.empty((12, 5008), dtype=torch.float32, device=dev)
    st = torch.cuda.current_stream().cuda_stream
    _lib.check(_lib.load().fuel_amd_pairwise_sqdist_device(Cn.data_ptr(), 12, 24, P.data_ptr(), 5000, 32, 24,
                                                          _lib.F32, out.data_ptr(), 5008, ctypes.c_void_p(st)))
    torch.cuda.synchronize()
    got = out.cpu().numpy()[:, :5000]
    assert got.tobytes() == oracle.kmeans_pairwise(cen, pts).tobytes()
