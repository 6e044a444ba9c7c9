"""fuel_amd_knn_device_noalloc: the exact kNN of a device query batch whose
scratch (candidate slots, list counts) is the library's fixed device array, so
the call allocates nothing.  Its answers must equal fuel_amd_knn_device's
(and the oracle's) bit for bit on two calls in a row, the second with fewer
queries (the array keeps the first call's contents), in both dtypes and every
exclusion mode; a batch beyond the array's capacity is refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(lib):
    f = lib.fuel_amd_knn_device_noalloc
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    f.restype = ctypes.c_int32
    return f


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("exclude", [0, 1, 2])
def test_noalloc_matches_knn_device(gpu, oracle, dtype, exclude):
    import torch
    lib = gpu.load()
    f = _noalloc(lib)
    rng = np.random.default_rng(40 + exclude)
    X = np.round(rng.random((20000, 24)) * 8).astype(dtype)  # ties
    dev = torch.device("cuda:0")
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    ix = gpu.Index(X, "euclidean", mode="bf16" if dtype == np.float32 else "exact")
    k = 9
    st = torch.cuda.current_stream().cuda_stream
    for m in (3000, 517):  # the second call is smaller
        Q = X[:m] if exclude else rng.random((m, 24)).astype(dtype) * 8
        Qd = torch.from_numpy(np.ascontiguousarray(Q)).to(dev)
        ai = torch.empty((m, k), dtype=torch.int64, device=dev)
        ad = torch.empty((m, k), dtype=tdt, device=dev)
        bi = torch.full((m, k), -7, dtype=torch.int64, device=dev)
        bd = torch.full((m, k), 7.0, dtype=tdt, device=dev)
        gpu.check(f(ix.handle, Qd.data_ptr(), m, 24, k, exclude, bi.data_ptr(), bd.data_ptr(), st))
        ix.knn_device(Qd.data_ptr(), m, 24, k, exclude, ai.data_ptr(), ad.data_ptr(), st)
        torch.cuda.synchronize()
        assert torch.equal(ai, bi), m
        assert ad.cpu().numpy().tobytes() == bd.cpu().numpy().tobytes(), m
        name = {0: "none", 1: "for_each_knn", 2: "python_linear"}[exclude]
        oi, od = oracle.knn(X, Q, k, "euclidean", name)
        assert np.array_equal(bi.cpu().numpy(), oi) and bd.cpu().numpy().tobytes() == od.tobytes()
    ix.close()


def test_noalloc_refuses_an_oversized_batch(gpu):
    import torch
    lib = gpu.load()
    f = _noalloc(lib)
    X = np.random.default_rng(1).random((1000, 8)).astype(np.float32)
    ix = gpu.Index(X, "euclidean")
    m = 40000  # 40000 x 4 KiB of slots per split > the 64 MiB array
    dev = torch.device("cuda:0")
    Qd = torch.zeros((m, 8), dtype=torch.float32, device=dev)
    oi = torch.empty((m, 3), dtype=torch.int64, device=dev)
    od = torch.empty((m, 3), dtype=torch.float32, device=dev)
    rc = f(ix.handle, Qd.data_ptr(), m, 8, 3, 0, oi.data_ptr(), od.data_ptr(),
           torch.cuda.current_stream().cuda_stream)
    assert rc == gpu.ERR_INVALID
    assert b"at most" in lib.fuel_amd_last_error()
    ix.close()
