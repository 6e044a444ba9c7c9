"""The C-ABI library loads, exports every symbol include/fuel_amd.h declares and,
without a GPU, refuses compute calls with FUEL_AMD_ERR_DEVICE (there is no CPU
fallback).  Also checks the Python layer's argument handling.  CPU only."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "fuel_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fuel_amd_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_are_exported():
    from fuel_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    declared = _declared_symbols()
    assert len(declared) >= 16
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(_lib.EXPORTED)


def test_library_has_gfx950_code_object():
    from fuel_amd import _lib
    blob = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_version_and_errors_without_gpu():
    from fuel_amd import _lib
    lib = _lib.load()
    assert lib.fuel_amd_version().startswith(b"fuel_amd")
    if _lib.device_count() > 0:
        pytest.skip("a GPU is visible")
    x = np.zeros((4, 3), np.float32)
    h = ctypes.c_void_p()
    rc = lib.fuel_amd_index_create(x.ctypes.data, 4, 3, _lib.F32, _lib.EUCLIDEAN, None, ctypes.byref(h))
    assert rc == _lib.ERR_DEVICE
    assert b"no HIP device" in lib.fuel_amd_last_error()
    with pytest.raises(_lib.FuelAmdError):
        _lib.Index(x)


def test_argument_errors_are_reported_before_the_device():
    from fuel_amd import _lib
    lib = _lib.load()
    h = ctypes.c_void_p()
    x = np.zeros((4, 3), np.float32)
    assert lib.fuel_amd_index_create(x.ctypes.data, 4, 3, 7, 0, None, ctypes.byref(h)) == _lib.ERR_INVALID
    assert lib.fuel_amd_index_create(x.ctypes.data, 4, 3, 0, 5, None, ctypes.byref(h)) == _lib.ERR_UNSUPPORTED
    assert lib.fuel_amd_index_create(x.ctypes.data, -1, 3, 0, 0, None, ctypes.byref(h)) == _lib.ERR_INVALID
    assert lib.fuel_amd_knn(None, None, 0, 1, 0, None, None) == _lib.ERR_INVALID


def test_distance_names():
    from fuel_amd import _lib
    assert _lib.metric_code("Euclidean") == 0 and _lib.metric_code("l2") == 0
    assert _lib.metric_code("sqeuclidean") == 1 and _lib.metric_code("squared_euclidean") == 1
    with pytest.raises(NotImplementedError):
        _lib.metric_code("cosine")
    with pytest.raises(ValueError):
        _lib.metric_code("nonsense")


def test_tree_names():
    from fuel_amd.search import _normalize_search_tree_name, _resolve_tree
    assert _resolve_tree(None) == "linear" and _resolve_tree("brute_force") == "linear"
    assert _normalize_search_tree_name("KD-tree") == "kd"
    with pytest.raises(NotImplementedError):
        _resolve_tree("vp")
    with pytest.raises(ValueError):
        _normalize_search_tree_name("oak")


def test_condensed_helpers():
    from fuel_amd.search import _condensed_to_square, _scipy_to_internal_condensed, pairwise_distances
    n = 6
    sq = np.arange(n * n, dtype=np.float64).reshape(n, n)
    sq = sq + sq.T
    np.fill_diagonal(sq, 0)
    iu = np.triu_indices(n, 1)
    scipy_cond = sq[iu]
    internal = _scipy_to_internal_condensed(scipy_cond)
    assert (_condensed_to_square(internal) == sq).all()
    assert (pairwise_distances(sq, "precomputed", "square") == sq).all()
    assert (pairwise_distances(scipy_cond, "precomputed", "square") == sq).all()
