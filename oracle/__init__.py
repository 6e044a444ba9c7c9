"""ORACLE — test infrastructure only.

ctypes wrapper around ``oracle/liboracle.so``, the C++ restatement of fuel's CPU
linear-scan path (see ``oracle/fuel_oracle.cpp`` for the per-function citations
into the fuel sources and for how each function is pinned).  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this module, and only as the checker; the product package ``fuel_amd``
never does.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

METRICS = {"euclidean": 0, "l2": 0, "sqeuclidean": 1, "squared_euclidean": 1}
EXCLUDE = {"none": 0, "for_each_knn": 1, "python_linear": 2}
# distance arithmetic: "avx" = src/distance/squared_euclidean.rs:13-67 (x86 hosts with
# AVX+FMA), "fallback" = :99-133 (every other host); bit 4 of the metric code
ARITH = {"avx": 0, "fallback": 16}


def _metric(metric, arith="avx"):
    return METRICS[metric] | ARITH[arith]


def build(force=False):
    """Compile the oracle with its Makefile (g++, no GPU needed)."""
    src = os.path.join(_HERE, "fuel_oracle.cpp")
    if force or not os.path.exists(_LIB_PATH) or os.path.getmtime(_LIB_PATH) < os.path.getmtime(src):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "liboracle.so"], check=True)


class _Csr(ctypes.Structure):
    _fields_ = [("offsets", ctypes.POINTER(ctypes.c_int64)),
                ("index", ctypes.POINTER(ctypes.c_int64)),
                ("dist", ctypes.c_void_p),
                ("nnz", ctypes.c_int64)]


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    P = ctypes.c_void_p
    I64 = ctypes.c_int64
    I = ctypes.c_int
    lib.oracle_sqdist_f32.restype = ctypes.c_float
    lib.oracle_sqdist_f32.argtypes = [P, P, I64]
    lib.oracle_sqdist_f32_lanes.restype = ctypes.c_float
    lib.oracle_sqdist_f32_lanes.argtypes = [P, P, I64]
    lib.oracle_sqdist_f32_fallback.restype = ctypes.c_float
    lib.oracle_sqdist_f32_fallback.argtypes = [P, P, I64]
    lib.oracle_sqdist_f64.restype = ctypes.c_double
    lib.oracle_sqdist_f64.argtypes = [P, P, I64]
    lib.oracle_sqdist_f64_fallback.restype = ctypes.c_double
    lib.oracle_sqdist_f64_fallback.argtypes = [P, P, I64]
    lib.oracle_sqdist_rows_f32.argtypes = [P, I64, I64, P, P]
    lib.oracle_sqdist_fp64_pairs.argtypes = [P, I64, P, P, P, I64, P]
    for t in ("f32", "f64"):
        getattr(lib, f"oracle_knn_{t}").argtypes = [P, I64, I64, P, I64, I64, I, I, I, P, P]
        getattr(lib, f"oracle_knn_ties_{t}").argtypes = [P, I64, I64, P, I64, I64, I, I, ctypes.POINTER(_Csr)]
        getattr(lib, f"oracle_range_{t}").argtypes = [P, I64, I64, P, I64, ctypes.c_double, I, I, I,
                                                      ctypes.POINTER(_Csr)]
        getattr(lib, f"oracle_pairwise_{t}").argtypes = [P, I64, I64, I, I, I, P]
        getattr(lib, f"oracle_lof_{t}").argtypes = [P, I64, I64, I64, I, I, P]
        getattr(lib, f"oracle_knn_outlier_{t}").argtypes = [P, I64, I64, I64, I, I, P]
        getattr(lib, f"oracle_subset_with_ties_{t}").argtypes = [P, I64, I64]
        getattr(lib, f"oracle_subset_with_ties_{t}").restype = I64
        getattr(lib, f"oracle_kmeans_pairwise_{t}").argtypes = [P, I64, P, I64, I64, I, P]
        getattr(lib, f"oracle_dbscan_{t}").argtypes = [P, I64, I64, ctypes.c_double, I64, I, I, I, P]
        getattr(lib, f"oracle_optics_{t}").argtypes = [P, I64, I64, ctypes.c_double, I64, I, P, P, P, P, P]
        getattr(lib, f"oracle_extract_xi_{t}").argtypes = [P, P, P, I64, ctypes.c_double, I64, P]
    lib.oracle_csr_free.argtypes = [ctypes.POINTER(_Csr)]
    _lib = lib
    return lib


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _tag(dtype):
    return "f32" if np.dtype(dtype) == np.float32 else "f64"


def default_threads():
    return max(1, min(16, os.cpu_count() or 1))


def sqdist(a, b, kind="avx"):
    """Restated squared distance of two vectors (kind: avx | lanes | fallback)."""
    lib = _load()
    if np.asarray(a).dtype == np.float64:
        a, b = _c(a, np.float64), _c(b, np.float64)
        fn = lib.oracle_sqdist_f64_fallback if kind == "fallback" else lib.oracle_sqdist_f64
        return np.float64(fn(a.ctypes.data, b.ctypes.data, a.shape[0]))
    a, b = _c(a, np.float32), _c(b, np.float32)
    fn = {"avx": lib.oracle_sqdist_f32, "lanes": lib.oracle_sqdist_f32_lanes,
          "fallback": lib.oracle_sqdist_f32_fallback}[kind]
    return np.float32(fn(a.ctypes.data, b.ctypes.data, a.shape[0]))


def sqdist_rows(X, q):
    lib = _load()
    X, q = _c(X, np.float32), _c(q, np.float32)
    out = np.empty(X.shape[0], np.float32)
    lib.oracle_sqdist_rows_f32(X.ctypes.data, X.shape[0], X.shape[1], q.ctypes.data, out.ctypes.data)
    return out


def sqdist_fp64(X, Q, qi, xi):
    """FP64 squared distances of pairs (Q[qi[t]], X[xi[t]])."""
    lib = _load()
    X, Q = _c(X, np.float32), _c(Q, np.float32)
    qi, xi = _c(qi, np.int64), _c(xi, np.int64)
    out = np.empty(qi.shape[0], np.float64)
    lib.oracle_sqdist_fp64_pairs(X.ctypes.data, X.shape[1], Q.ctypes.data, qi.ctypes.data,
                                 xi.ctypes.data, qi.shape[0], out.ctypes.data)
    return out


def knn(X, Q, k, metric="euclidean", exclude="none", threads=None, arith="avx"):
    """Python-packed kNN: (int64[m,k] padded -1, F[m,k] padded +inf)."""
    lib = _load()
    dt = np.float64 if np.asarray(X).dtype == np.float64 else np.float32
    X, Q = _c(X, dt), _c(Q, dt)
    m = Q.shape[0]
    idx = np.empty((m, k), np.int64)
    dist = np.empty((m, k), dt)
    getattr(lib, f"oracle_knn_{_tag(dt)}")(
        X.ctypes.data, X.shape[0], X.shape[1], Q.ctypes.data, m, k, _metric(metric, arith),
        EXCLUDE[exclude] if isinstance(exclude, str) else int(exclude),
        threads or default_threads(), idx.ctypes.data, dist.ctypes.data)
    return idx, dist


def _csr_to_lists(c, dt, m):
    lib = _load()
    offs = np.ctypeslib.as_array(c.offsets, shape=(m + 1,)).copy()
    nnz = int(c.nnz)
    idx = np.ctypeslib.as_array(c.index, shape=(max(1, nnz),))[:nnz].copy()
    ptr = ctypes.cast(c.dist, ctypes.POINTER(ctypes.c_float if dt == np.float32 else ctypes.c_double))
    dist = np.ctypeslib.as_array(ptr, shape=(max(1, nnz),))[:nnz].copy()
    lib.oracle_csr_free(ctypes.byref(c))
    return offs, idx, dist


def knn_ties(X, Q, k, metric="euclidean", threads=None, arith="avx"):
    """KNNHeap rows with all ties, as CSR (offsets, index, dist)."""
    lib = _load()
    dt = np.float64 if np.asarray(X).dtype == np.float64 else np.float32
    X, Q = _c(X, dt), _c(Q, dt)
    c = _Csr()
    getattr(lib, f"oracle_knn_ties_{_tag(dt)}")(
        X.ctypes.data, X.shape[0], X.shape[1], Q.ctypes.data, Q.shape[0], k, _metric(metric, arith),
        threads or default_threads(), ctypes.byref(c))
    return _csr_to_lists(c, dt, Q.shape[0])


def range_search(X, Q, radius, metric="euclidean", exclude_self=False, threads=None, arith="avx"):
    """Range rows as CSR (offsets, index, dist)."""
    lib = _load()
    dt = np.float64 if np.asarray(X).dtype == np.float64 else np.float32
    X, Q = _c(X, dt), _c(Q, dt)
    c = _Csr()
    getattr(lib, f"oracle_range_{_tag(dt)}")(
        X.ctypes.data, X.shape[0], X.shape[1], Q.ctypes.data, Q.shape[0], float(radius),
        _metric(metric, arith), int(bool(exclude_self)), threads or default_threads(), ctypes.byref(c))
    return _csr_to_lists(c, dt, Q.shape[0])


def pairwise(X, metric="euclidean", form="square", threads=None, arith="avx"):
    lib = _load()
    dt = np.float64 if np.asarray(X).dtype == np.float64 else np.float32
    X = _c(X, dt)
    n = X.shape[0]
    out = np.empty((n, n) if form == "square" else (n * (n - 1) // 2,), dt)
    getattr(lib, f"oracle_pairwise_{_tag(dt)}")(
        X.ctypes.data, n, X.shape[1], _metric(metric, arith), 0 if form == "square" else 1,
        threads or default_threads(), out.ctypes.data)
    return out


def lof(X, k, metric="euclidean", threads=None, arith="avx"):
    lib = _load()
    dt = np.float64 if np.asarray(X).dtype == np.float64 else np.float32
    X = _c(X, dt)
    out = np.empty(X.shape[0], dt)
    getattr(lib, f"oracle_lof_{_tag(dt)}")(X.ctypes.data, X.shape[0], X.shape[1], k, _metric(metric, arith),
                                           threads or default_threads(), out.ctypes.data)
    return out


def knn_outlier(X, k, metric="euclidean", threads=None):
    lib = _load()
    dt = np.float64 if np.asarray(X).dtype == np.float64 else np.float32
    X = _c(X, dt)
    out = np.empty(X.shape[0], dt)
    getattr(lib, f"oracle_knn_outlier_{_tag(dt)}")(X.ctypes.data, X.shape[0], X.shape[1], k,
                                                   METRICS[metric], threads or default_threads(),
                                                   out.ctypes.data)
    return out


def subset_with_ties(dist_row, k):
    lib = _load()
    dt = np.float64 if np.asarray(dist_row).dtype == np.float64 else np.float32
    d = _c(dist_row, dt)
    return int(getattr(lib, f"oracle_subset_with_ties_{_tag(dt)}")(d.ctypes.data, d.shape[0], k))


def dbscan(X, eps, min_points, metric="euclidean", variant="dbscan", threads=None):
    """Labels of src/cluster/dbscan.rs:14-83 ('dbscan') or
    src/cluster/parallel_dbscan.rs:16-113 ('parallel') over the linear scan."""
    lib = _load()
    dt = np.float64 if np.asarray(X).dtype == np.float64 else np.float32
    X = _c(X, dt)
    labels = np.empty(X.shape[0], np.int64)
    getattr(lib, f"oracle_dbscan_{_tag(dt)}")(
        X.ctypes.data, X.shape[0], X.shape[1], float(eps), int(min_points), METRICS[metric],
        0 if variant == "dbscan" else 1, threads or default_threads(), labels.ctypes.data)
    return labels


def optics(X, eps, min_points, metric="euclidean"):
    """(ordering, reachability, core_distance (NaN = none), predecessor (-1 = none),
    labels) of src/cluster/optics.rs:56-133 over the linear scan."""
    lib = _load()
    dt = np.float64 if np.asarray(X).dtype == np.float64 else np.float32
    X = _c(X, dt)
    n = X.shape[0]
    ordering = np.empty(n, np.int64)
    reach = np.empty(n, dt)
    core = np.empty(n, dt)
    pred = np.empty(n, np.int64)
    labels = np.empty(n, np.int64)
    getattr(lib, f"oracle_optics_{_tag(dt)}")(
        X.ctypes.data, n, X.shape[1], float(eps), int(min_points), METRICS[metric], ordering.ctypes.data,
        reach.ctypes.data, core.ctypes.data, pred.ctypes.data, labels.ctypes.data)
    return ordering, reach, core, pred, labels


def extract_xi(ordering, reachability, predecessor, xi, min_points):
    """extract_xi_labels, src/cluster/optics.rs:148-304."""
    lib = _load()
    reach = np.asarray(reachability)
    dt = np.float64 if reach.dtype == np.float64 else np.float32
    ordering = np.ascontiguousarray(ordering, np.int64)
    reach = np.ascontiguousarray(reach, dt)
    pred = np.ascontiguousarray(predecessor, np.int64)
    labels = np.empty(ordering.shape[0], np.int64)
    getattr(lib, f"oracle_extract_xi_{_tag(dt)}")(
        ordering.ctypes.data, reach.ctypes.data, pred.ctypes.data, ordering.shape[0], float(xi), int(min_points),
        labels.ctypes.data)
    return labels


def kmeans_pairwise(A, B, threads=None):
    """math::pairwise_sqdist (src/math/mod.rs:562-579): (m, n) in the k-means
    rounding order."""
    lib = _load()
    dt = np.float64 if np.asarray(A).dtype == np.float64 else np.float32
    A, B = _c(A, dt), _c(B, dt)
    out = np.empty((A.shape[0], B.shape[0]), dt)
    getattr(lib, f"oracle_kmeans_pairwise_{_tag(dt)}")(A.ctypes.data, A.shape[0], B.ctypes.data, B.shape[0],
                                                       A.shape[1], threads or default_threads(), out.ctypes.data)
    return out


_SKIP_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int64)


class PrioritySearcher:
    """LinearScanPrioritySearcher (src/search/linear_scan.rs:70-235) of one
    query over X, restated in oracle/fuel_oracle.cpp."""

    def __init__(self, X, q, metric="euclidean", arith="avx"):
        lib = _load()
        dt = np.float64 if np.asarray(X).dtype == np.float64 else np.float32
        self._t = _tag(dt)
        self._dt = dt
        self._X = _c(X, dt)  # kept alive: the searcher reads it
        self._q = _c(q, dt)
        cty = ctypes.c_float if dt == np.float32 else ctypes.c_double
        P = ctypes.c_void_p
        f = getattr(lib, f"oracle_priority_create_{self._t}")
        f.restype = P
        f.argtypes = [P, ctypes.c_int64, ctypes.c_int64, P, ctypes.c_int]
        for name, res, args in [("destroy", None, [P]), ("reset", None, [P, ctypes.c_double, ctypes.c_double]),
                                ("decrease_cutoff", None, [P, ctypes.c_double]),
                                ("next", ctypes.c_int, [P, _SKIP_FN, P, ctypes.POINTER(ctypes.c_int64),
                                                        ctypes.POINTER(cty)]),
                                ("all_lower_bound", cty, [P])]:
            g = getattr(lib, f"oracle_priority_{name}_{self._t}")
            g.restype = res
            g.argtypes = args
        self._cty = cty
        self._h = f(self._X.ctypes.data, self._X.shape[0], self._X.shape[1], self._q.ctypes.data,
                    _metric(metric, arith))

    def __del__(self):
        try:
            getattr(_load(), f"oracle_priority_destroy_{self._t}")(self._h)
        except Exception:
            pass

    def reset(self):
        getattr(_load(), f"oracle_priority_reset_{self._t}")(self._h, float("inf"), 0.0)

    def reset_with_limits(self, cutoff, skip):
        getattr(_load(), f"oracle_priority_reset_{self._t}")(self._h, float(cutoff), float(skip))

    def decrease_cutoff(self, t):
        getattr(_load(), f"oracle_priority_decrease_cutoff_{self._t}")(self._h, float(t))

    def next(self, skip_point=None):
        idx = ctypes.c_int64()
        dist = self._cty()
        cb = ctypes.cast(None, _SKIP_FN) if skip_point is None else _SKIP_FN(lambda _c, i: 1 if skip_point(int(i)) else 0)
        found = getattr(_load(), f"oracle_priority_next_{self._t}")(self._h, cb, None, ctypes.byref(idx),
                                                                    ctypes.byref(dist))
        if not found:
            return None
        return int(idx.value), np.array(dist.value, dtype=self._dt)[()]

    def all_lower_bound(self):
        return np.array(getattr(_load(), f"oracle_priority_all_lower_bound_{self._t}")(self._h), dtype=self._dt)[()]
