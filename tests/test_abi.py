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


def test_struct_layouts_match_header(tmp_path):
    # the ctypes mirrors of every struct in include/fuel_amd.h, against the C
    # compiler's own sizeof / offsetof of the header's declarations
    import subprocess
    from fuel_amd import _lib
    mirrors = {"fuel_amd_opts": _lib.Opts, "fuel_amd_csr": _lib.Csr, "fuel_amd_stats": _lib.Stats,
               "fuel_amd_server_opts": _lib.ServerOpts, "fuel_amd_serve_stats": _lib.ServeStats,
               "fuel_amd_problem": _lib.Problem, "fuel_amd_batch_opts": _lib.BatchOpts}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "fuel_amd.h"', "int main(void) {"]
    for cname, ct in mirrors.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in ct._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cname, ct in mirrors.items():
        assert got[(cname, "size")] == ctypes.sizeof(ct), cname
        for f, _ in ct._fields_:
            assert got[(cname, f)] == getattr(ct, f).offset, (cname, f)


def test_server_and_batch_argument_errors():
    from fuel_amd import _lib
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.fuel_amd_server_create(None, None, ctypes.byref(h)) == _lib.ERR_INVALID
    assert lib.fuel_amd_server_submit(None, None, 1, 1, None) == _lib.ERR_INVALID
    assert lib.fuel_amd_server_destroy(None) == _lib.OK
    t = ctypes.c_uint64()
    assert lib.fuel_amd_server_wait(None, t, None, None, None, None) == _lib.ERR_INVALID
    x = np.zeros((16, 12), np.float32)
    oi = np.zeros((16, 3), np.int64)
    od = np.zeros((16, 3), np.float32)
    bad_ld = _lib.Problem(x.ctypes.data, 16, 12, 12, 3, oi.ctypes.data, od.ctypes.data)  # ld not a multiple of 8
    arr = (_lib.Problem * 1)(bad_ld)
    assert lib.fuel_amd_batch_create(arr, 1, _lib.EUCLIDEAN, 1, None, ctypes.byref(h)) == _lib.ERR_INVALID
    opts = _lib.BatchOpts(9, 4, 0, 0)
    assert lib.fuel_amd_batch_create(arr, 0, _lib.EUCLIDEAN, 1, ctypes.byref(opts), ctypes.byref(h)) == _lib.ERR_INVALID
    assert lib.fuel_amd_batch_create(arr, 0, 5, 1, None, ctypes.byref(h)) == _lib.ERR_UNSUPPORTED
    assert lib.fuel_amd_batch_destroy(None) == _lib.OK
