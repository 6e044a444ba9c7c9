"""ctypes binding of libfuel_amd.so (the C ABI declared in include/fuel_amd.h).

The library is loaded from this package directory only; there is no CPU
fallback: if the library is missing or no GPU is visible, calls raise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfuel_amd.so")

OK = 0
ERR_INVALID = 1
ERR_UNSUPPORTED = 2
ERR_DEVICE = 3
ERR_INTERRUPTED = 4
ERR_NOMEM = 5

F32 = 0
F64 = 1
EUCLIDEAN = 0
SQEUCLIDEAN = 1
EXCLUDE_NONE = 0
EXCLUDE_FOR_EACH_KNN = 1
EXCLUDE_PY_LINEAR = 2
MODE_AUTO = 0
MODE_EXACT = 1
MODE_FP32 = 2
MODE_BF16 = 3
MODE_BF16R = 4
MODE_FP8R = 5
FORM_SQUARE = 0
FORM_CONDENSED = 1

MODES = {"auto": MODE_AUTO, "exact": MODE_EXACT, "fp32": MODE_FP32, "bf16": MODE_BF16, "bf16r": MODE_BF16R,
         "fp8r": MODE_FP8R}
FLAG_FALLBACK_ARITH = 1
FLAG_NO_SEED = 2
# rounding of the restated distance: "avx" = fuel's AVX+FMA kernels
# (src/distance/squared_euclidean.rs:13-67, x86 hosts with AVX2+FMA),
# "fallback" = its portable kernel (:99-133, every other host)
ARITH = {"avx": 0, "fallback": FLAG_FALLBACK_ARITH}

# every symbol include/fuel_amd.h declares
EXPORTED = [
    "fuel_amd_version", "fuel_amd_device_count", "fuel_amd_last_error",
    "fuel_amd_request_cancel", "fuel_amd_clear_cancel",
    "fuel_amd_index_create", "fuel_amd_index_create_device", "fuel_amd_index_destroy",
    "fuel_amd_knn", "fuel_amd_knn_device", "fuel_amd_knn_device_noalloc", "fuel_amd_range", "fuel_amd_knn_ties", "fuel_amd_csr_free",
    "fuel_amd_pairwise", "fuel_amd_index_pairwise", "fuel_amd_lof", "fuel_amd_knn_outlier", "fuel_amd_last_stats",
    "fuel_amd_last_fallback_rows",
    "fuel_amd_dbscan", "fuel_amd_optics", "fuel_amd_pairwise_sqdist", "fuel_amd_pairwise_sqdist_device",
    "fuel_amd_timing_enable", "fuel_amd_timing_read",
    "fuel_amd_server_create", "fuel_amd_server_submit", "fuel_amd_server_wait", "fuel_amd_server_stats",
    "fuel_amd_server_destroy",
    "fuel_amd_batch_create", "fuel_amd_batch_run", "fuel_amd_batch_describe", "fuel_amd_batch_destroy",
    "fuel_amd_priority_create", "fuel_amd_priority_destroy", "fuel_amd_priority_reset",
    "fuel_amd_priority_decrease_cutoff", "fuel_amd_priority_next", "fuel_amd_priority_all_lower_bound",
]

SKIP_POINT = ctypes.CFUNCTYPE(ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64)


class Opts(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("widen", ctypes.c_int32), ("device", ctypes.c_int32),
                ("flags", ctypes.c_int32), ("reserved", ctypes.c_int64 * 6)]


class Csr(ctypes.Structure):
    _fields_ = [("offsets", ctypes.POINTER(ctypes.c_int64)), ("index", ctypes.POINTER(ctypes.c_int64)),
                ("dist", ctypes.c_void_p), ("nnz", ctypes.c_int64), ("m", ctypes.c_int64)]


class Stats(ctypes.Structure):
    _fields_ = [("queries", ctypes.c_int64), ("fallback_queries", ctypes.c_int64),
                ("mode_used", ctypes.c_int32), ("splits", ctypes.c_int32), ("kp", ctypes.c_int32),
                ("exact_queries", ctypes.c_int32), ("ms_total", ctypes.c_double)]


class ServerOpts(ctypes.Structure):
    _fields_ = [("workers", ctypes.c_int32), ("reserved0", ctypes.c_int32), ("max_batch", ctypes.c_int64),
                ("hbm_budget", ctypes.c_int64), ("reserved", ctypes.c_int64 * 5)]


class ServeStats(ctypes.Structure):
    _fields_ = [("requests", ctypes.c_int64), ("queries", ctypes.c_int64), ("batches", ctypes.c_int64),
                ("max_batch_queries", ctypes.c_int64), ("failed", ctypes.c_int64), ("p50_ms", ctypes.c_double),
                ("p99_ms", ctypes.c_double), ("max_ms", ctypes.c_double), ("mean_ms", ctypes.c_double),
                ("busy_s", ctypes.c_double)]


class Problem(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("n", ctypes.c_int64), ("ld", ctypes.c_int64), ("d", ctypes.c_int32),
                ("k", ctypes.c_int32), ("out_idx", ctypes.c_void_p), ("out_dist", ctypes.c_void_p)]


class BatchOpts(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("widen", ctypes.c_int32), ("device", ctypes.c_int32),
                ("no_graph", ctypes.c_int32), ("tune_cache", ctypes.c_char_p), ("reserved", ctypes.c_int64 * 3)]


class FuelAmdError(RuntimeError):
    """Raised when the C ABI returns an error status."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


_lib = None


def load():
    """Load the in-tree library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m fuel_amd._build`")
    # One HIP runtime per process: torch bundles its own libamdhip64 (same soname,
    # libamdhip64.so.7).  Loading torch first lets the dynamic loader satisfy this
    # library's dependency with the copy torch already mapped, so torch device
    # buffers and streams can be passed to the *_device entry points.
    try:
        import torch  # noqa: F401
    except Exception:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    I32 = ctypes.c_int32
    I64 = ctypes.c_int64
    lib.fuel_amd_version.restype = ctypes.c_char_p
    lib.fuel_amd_device_count.restype = I32
    lib.fuel_amd_last_error.restype = ctypes.c_char_p
    lib.fuel_amd_index_create.argtypes = [P, I64, I32, I32, I32, ctypes.POINTER(Opts), ctypes.POINTER(P)]
    lib.fuel_amd_index_create.restype = I32
    lib.fuel_amd_index_create_device.argtypes = [P, I64, I32, I64, I32, I32, ctypes.POINTER(Opts),
                                                 ctypes.POINTER(P)]
    lib.fuel_amd_index_create_device.restype = I32
    lib.fuel_amd_index_destroy.argtypes = [P]
    lib.fuel_amd_index_destroy.restype = I32
    lib.fuel_amd_knn.argtypes = [P, P, I64, I32, I32, P, P]
    lib.fuel_amd_knn.restype = I32
    lib.fuel_amd_knn_device.argtypes = [P, P, I64, I64, I32, I32, P, P, P]
    lib.fuel_amd_knn_device.restype = I32
    lib.fuel_amd_range.argtypes = [P, P, I64, ctypes.c_double, I32, ctypes.POINTER(Csr)]
    lib.fuel_amd_range.restype = I32
    lib.fuel_amd_pairwise_sqdist.argtypes = [P, I64, P, I64, I32, I32, P]
    lib.fuel_amd_pairwise_sqdist.restype = I32
    lib.fuel_amd_pairwise_sqdist_device.argtypes = [P, I64, I64, P, I64, I64, I32, I32, P, I64, P]
    lib.fuel_amd_pairwise_sqdist_device.restype = I32
    lib.fuel_amd_dbscan.argtypes = [P, ctypes.c_double, I64, I32, P]
    lib.fuel_amd_dbscan.restype = I32
    lib.fuel_amd_optics.argtypes = [P, ctypes.c_double, I64, P, P, P, P, P]
    lib.fuel_amd_optics.restype = I32
    lib.fuel_amd_knn_ties.argtypes = [P, P, I64, I32, ctypes.POINTER(Csr)]
    lib.fuel_amd_knn_ties.restype = I32
    lib.fuel_amd_csr_free.argtypes = [ctypes.POINTER(Csr)]
    lib.fuel_amd_pairwise.argtypes = [P, I64, I32, I32, I32, I32, P]
    lib.fuel_amd_pairwise.restype = I32
    lib.fuel_amd_index_pairwise.argtypes = [P, I32, P]
    lib.fuel_amd_index_pairwise.restype = I32
    lib.fuel_amd_lof.argtypes = [P, I32, P]
    lib.fuel_amd_lof.restype = I32
    lib.fuel_amd_knn_outlier.argtypes = [P, I32, P]
    lib.fuel_amd_knn_outlier.restype = I32
    lib.fuel_amd_last_stats.argtypes = [P, ctypes.POINTER(Stats)]
    lib.fuel_amd_last_stats.restype = I32
    lib.fuel_amd_last_fallback_rows.argtypes = [P, ctypes.POINTER(I64), I64, ctypes.POINTER(I64)]
    lib.fuel_amd_last_fallback_rows.restype = I32
    lib.fuel_amd_timing_enable.argtypes = [P, I32]
    lib.fuel_amd_timing_enable.restype = I32
    lib.fuel_amd_timing_read.argtypes = [P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(I64),
                                         ctypes.c_char_p, I32]
    lib.fuel_amd_timing_read.restype = I32
    lib.fuel_amd_server_create.argtypes = [P, ctypes.POINTER(ServerOpts), ctypes.POINTER(P)]
    lib.fuel_amd_server_create.restype = I32
    lib.fuel_amd_server_submit.argtypes = [P, P, I64, I32, ctypes.POINTER(ctypes.c_uint64)]
    lib.fuel_amd_server_submit.restype = I32
    lib.fuel_amd_server_wait.argtypes = [P, ctypes.c_uint64, P, P, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(I64)]
    lib.fuel_amd_server_wait.restype = I32
    lib.fuel_amd_server_stats.argtypes = [P, I32, ctypes.POINTER(ServeStats)]
    lib.fuel_amd_server_stats.restype = I32
    lib.fuel_amd_server_destroy.argtypes = [P]
    lib.fuel_amd_server_destroy.restype = I32
    lib.fuel_amd_batch_create.argtypes = [ctypes.POINTER(Problem), I32, I32, I32, ctypes.POINTER(BatchOpts),
                                          ctypes.POINTER(P)]
    lib.fuel_amd_batch_create.restype = I32
    lib.fuel_amd_batch_run.argtypes = [P]
    lib.fuel_amd_batch_run.restype = I32
    lib.fuel_amd_batch_describe.argtypes = [P, ctypes.c_char_p, I32]
    lib.fuel_amd_batch_describe.restype = I32
    lib.fuel_amd_batch_destroy.argtypes = [P]
    lib.fuel_amd_batch_destroy.restype = I32
    lib.fuel_amd_priority_create.argtypes = [P, P, I64, ctypes.POINTER(P)]
    lib.fuel_amd_priority_create.restype = I32
    lib.fuel_amd_priority_destroy.argtypes = [P]
    lib.fuel_amd_priority_destroy.restype = I32
    lib.fuel_amd_priority_reset.argtypes = [P, I64, ctypes.c_double, ctypes.c_double]
    lib.fuel_amd_priority_reset.restype = I32
    lib.fuel_amd_priority_decrease_cutoff.argtypes = [P, I64, ctypes.c_double]
    lib.fuel_amd_priority_decrease_cutoff.restype = I32
    lib.fuel_amd_priority_next.argtypes = [P, I64, SKIP_POINT, P, ctypes.POINTER(I64), P]
    lib.fuel_amd_priority_next.restype = I32
    lib.fuel_amd_priority_all_lower_bound.argtypes = [P, I64, P]
    lib.fuel_amd_priority_all_lower_bound.restype = I32
    _lib = lib
    return lib


def check(rc):
    if rc != OK:
        msg = load().fuel_amd_last_error().decode(errors="replace")
        if rc == ERR_INVALID:
            raise ValueError(msg)
        if rc == ERR_INTERRUPTED:
            raise KeyboardInterrupt(msg)
        if rc == ERR_UNSUPPORTED:
            raise NotImplementedError(msg)
        raise FuelAmdError(rc, msg)


def dtype_code(dt):
    dt = np.dtype(dt)
    if dt == np.float32:
        return F32
    if dt == np.float64:
        return F64
    raise ValueError(f"unsupported dtype {dt}")


def metric_code(name):
    """fuel's parse_distance_fn names for the two metrics this engine serves
    (python/mod.rs:42-71); other names are outside the linear-scan engine."""
    key = name.lower()
    if key in ("euclidean", "l2"):
        return EUCLIDEAN
    if key in ("sqeuclidean", "squared_euclidean"):
        return SQEUCLIDEAN
    known = {"manhattan", "l1", "cityblock", "chebyshev", "linf", "chessboard", "cosine", "arccosine",
             "angular", "canberra", "braycurtis", "bray_curtis", "hellinger", "clark", "chi", "chi_squared",
             "chisquared", "chi2", "jensen_shannon", "jensenshannon", "js", "jeffrey", "jeffreys",
             "histogram_intersection", "intersection"}
    if key in known:
        raise NotImplementedError(f"distance '{name}' is not served by the GPU linear-scan engine")
    raise ValueError(
        f"unknown distance '{name}', valid options are: euclidean, sqeuclidean, manhattan, "
        "chebyshev, cosine, arccosine, canberra, braycurtis, hellinger, clark, chi, "
        "chi_squared, jensen_shannon, jeffrey, histogram_intersection")


class Index:
    """Owner of a fuel_amd_index handle."""

    def __init__(self, data, metric="euclidean", mode="auto", widen=0, device=0, arith="avx", seed=True):
        lib = load()
        data = np.ascontiguousarray(data)
        if data.ndim != 2:
            raise ValueError("data must be a 2D array")
        self.dtype = data.dtype
        self.n, self.d = data.shape
        opts = Opts(MODES[mode] if isinstance(mode, str) else int(mode), int(widen), int(device),
                    ARITH[arith] | (0 if seed else FLAG_NO_SEED))
        h = ctypes.c_void_p()
        check(lib.fuel_amd_index_create(data.ctypes.data, self.n, self.d, dtype_code(data.dtype),
                                        metric_code(metric), ctypes.byref(opts), ctypes.byref(h)))
        self.handle = h

    @classmethod
    def from_device(cls, ptr, n, d, ld, dtype, metric="euclidean", mode="auto", widen=0, device=0, arith="avx",
                    seed=True):
        lib = load()
        self = cls.__new__(cls)
        self.dtype = np.dtype(dtype)
        self.n, self.d = n, d
        opts = Opts(MODES[mode] if isinstance(mode, str) else int(mode), int(widen), int(device),
                    ARITH[arith] | (0 if seed else FLAG_NO_SEED))
        h = ctypes.c_void_p()
        check(lib.fuel_amd_index_create_device(ctypes.c_void_p(ptr), n, d, ld, dtype_code(dtype),
                                               metric_code(metric), ctypes.byref(opts), ctypes.byref(h)))
        self.handle = h
        return self

    def close(self):
        if getattr(self, "handle", None):
            load().fuel_amd_index_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def knn(self, query, k, exclude=EXCLUDE_NONE):
        query = np.ascontiguousarray(query, dtype=self.dtype)
        if query.ndim != 2 or query.shape[1] != self.d:
            raise ValueError("query shape mismatch")
        m = query.shape[0]
        idx = np.empty((m, k), np.int64)
        dist = np.empty((m, k), self.dtype)
        check(load().fuel_amd_knn(self.handle, query.ctypes.data, m, int(k), int(exclude), idx.ctypes.data,
                                  dist.ctypes.data))
        return idx, dist

    def knn_device(self, qptr, m, ldq, k, exclude, idx_ptr, dist_ptr, stream=0):
        check(load().fuel_amd_knn_device(self.handle, ctypes.c_void_p(qptr), m, ldq, int(k), int(exclude),
                                         ctypes.c_void_p(idx_ptr), ctypes.c_void_p(dist_ptr),
                                         ctypes.c_void_p(stream)))

    def range(self, query, radius, exclude_self=False):
        """All points within `radius` (inclusive) of each query: CSR (offsets, idx, dist)."""
        query = np.ascontiguousarray(query, dtype=self.dtype)
        if query.ndim != 2 or query.shape[1] != self.d:
            raise ValueError("query shape mismatch")
        c = Csr()
        check(load().fuel_amd_range(self.handle, query.ctypes.data, query.shape[0], float(radius),
                                    int(bool(exclude_self)), ctypes.byref(c)))
        return self._take_csr(c, query.shape[0])

    def knn_ties(self, query, k):
        """kNN with fuel's tie semantics (every point at or below the k-th distance):
        CSR (offsets, idx, dist) sorted by (distance, index)."""
        query = np.ascontiguousarray(query, dtype=self.dtype)
        if query.ndim != 2 or query.shape[1] != self.d:
            raise ValueError("query shape mismatch")
        c = Csr()
        check(load().fuel_amd_knn_ties(self.handle, query.ctypes.data, query.shape[0], int(k), ctypes.byref(c)))
        return self._take_csr(c, query.shape[0])

    def dbscan(self, eps, min_points, variant=0):
        """DBSCAN labels of the index rows (variant 0 sequential, 1 parallel)."""
        labels = np.empty(self.n, np.int64)
        check(load().fuel_amd_dbscan(self.handle, float(eps), int(min_points), int(variant), labels.ctypes.data))
        return labels

    def optics(self, max_eps, min_points):
        """(ordering, reachability, core_distance, predecessor, labels) of the index rows."""
        n = self.n
        ordering = np.empty(n, np.int64)
        reach = np.empty(n, self.dtype)
        core = np.empty(n, self.dtype)
        pred = np.empty(n, np.int64)
        labels = np.empty(n, np.int64)
        check(load().fuel_amd_optics(self.handle, float(max_eps), int(min_points), ordering.ctypes.data,
                                     reach.ctypes.data, core.ctypes.data, pred.ctypes.data, labels.ctypes.data))
        return ordering, reach, core, pred, labels

    def _take_csr(self, c, m):
        lib = load()
        try:
            offs = np.ctypeslib.as_array(c.offsets, shape=(m + 1,)).copy()
            nnz = int(c.nnz)
            idx = np.ctypeslib.as_array(c.index, shape=(max(1, nnz),))[:nnz].copy()
            cty = ctypes.c_float if self.dtype == np.float32 else ctypes.c_double
            dist = np.ctypeslib.as_array(ctypes.cast(c.dist, ctypes.POINTER(cty)), shape=(max(1, nnz),))[:nnz].copy()
        finally:
            lib.fuel_amd_csr_free(ctypes.byref(c))
        return offs, idx, dist

    def priority(self, query):
        """fuel's LinearScanPrioritySearcher for every query row (slot q = row q)."""
        return PrioritySearchers(self, query)

    def pairwise(self, form="square"):
        """Distance matrix of the index rows (square, or fuel's condensed order)."""
        n = self.n
        out = np.empty((n, n) if form == "square" else (n * (n - 1) // 2,), self.dtype)
        check(load().fuel_amd_index_pairwise(self.handle, FORM_SQUARE if form == "square" else FORM_CONDENSED,
                                             out.ctypes.data))
        return out

    def lof(self, k):
        out = np.empty(self.n, self.dtype)
        check(load().fuel_amd_lof(self.handle, int(k), out.ctypes.data))
        return out

    def knn_outlier(self, k):
        out = np.empty(self.n, self.dtype)
        check(load().fuel_amd_knn_outlier(self.handle, int(k), out.ctypes.data))
        return out

    def timing_enable(self, on=True):
        check(load().fuel_amd_timing_enable(self.handle, int(bool(on))))

    def timing_read(self):
        """(summed kernel ms, launches, kernel name) since the last read."""
        ms = ctypes.c_double()
        cnt = ctypes.c_int64()
        name = ctypes.create_string_buffer(128)
        check(load().fuel_amd_timing_read(self.handle, ctypes.byref(ms), ctypes.byref(cnt), name, 128))
        return ms.value, cnt.value, name.value.decode()

    def fallback_rows(self):
        """Query rows of the last kNN call whose first-level certificate failed."""
        lib = load()
        cnt = ctypes.c_int64()
        check(lib.fuel_amd_last_fallback_rows(self.handle, None, 0, ctypes.byref(cnt)))
        rows = np.empty(cnt.value, np.int64)
        if cnt.value:
            check(lib.fuel_amd_last_fallback_rows(self.handle, rows.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                                   cnt.value, ctypes.byref(cnt)))
        return rows

    def stats(self):
        s = Stats()
        check(load().fuel_amd_last_stats(self.handle, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in Stats._fields_}


def pairwise(data, metric="euclidean", form="square"):
    data = np.ascontiguousarray(data)
    n, d = data.shape
    out = np.empty((n, n) if form == "square" else (n * (n - 1) // 2,), data.dtype)
    check(load().fuel_amd_pairwise(data.ctypes.data, n, d, dtype_code(data.dtype), metric_code(metric),
                                   FORM_SQUARE if form == "square" else FORM_CONDENSED, out.ctypes.data))
    return out


def device_count():
    return int(load().fuel_amd_device_count())


def request_cancel():
    load().fuel_amd_request_cancel()


def clear_cancel():
    load().fuel_amd_clear_cancel()


def pairwise_sqdist(points1, points2):
    """fuel's math::pairwise_sqdist on the GPU: (nrows, ncols) squared distances
    in the reference's k-means rounding order."""
    p1 = np.asarray(points1)
    dt = np.float64 if p1.dtype == np.float64 else np.float32
    p1 = np.ascontiguousarray(p1, dt)
    p2 = np.ascontiguousarray(points2, dt)
    if p1.ndim != 2 or p2.ndim != 2 or p1.shape[1] != p2.shape[1]:
        raise ValueError("point dimensionality mismatch")
    out = np.empty((p1.shape[0], p2.shape[0]), dt)
    check(load().fuel_amd_pairwise_sqdist(p1.ctypes.data, p1.shape[0], p2.ctypes.data, p2.shape[0], p1.shape[1],
                                          dtype_code(dt), out.ctypes.data))
    return out


class PrioritySearchers:
    """One LinearScanPrioritySearcher (src/search/linear_scan.rs:70-235) per
    query row; slot q answers for query row q.  Methods mirror the
    PrioritySearcher trait (src/api/search.rs:216-229), with the slot first."""

    def __init__(self, index, query):
        query = np.ascontiguousarray(query, dtype=index.dtype)
        if query.ndim != 2 or query.shape[1] != index.d:
            raise ValueError("query shape mismatch")
        self.index = index  # keeps the index alive
        self.m = query.shape[0]
        self._cty = ctypes.c_float if index.dtype == np.float32 else ctypes.c_double
        h = ctypes.c_void_p()
        check(load().fuel_amd_priority_create(index.handle, query.ctypes.data, self.m, ctypes.byref(h)))
        self.handle = h
        self._no_filter = ctypes.cast(None, SKIP_POINT)

    def close(self):
        if getattr(self, "handle", None):
            load().fuel_amd_priority_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self, q):
        check(load().fuel_amd_priority_reset(self.handle, int(q), float("inf"), 0.0))

    def reset_with_limits(self, q, cutoff, skip):
        check(load().fuel_amd_priority_reset(self.handle, int(q), float(cutoff), float(skip)))

    def decrease_cutoff(self, q, threshold):
        check(load().fuel_amd_priority_decrease_cutoff(self.handle, int(q), float(threshold)))

    def next(self, q, skip_point=None):
        """(index, distance) of the next neighbour of slot q, or None."""
        idx = ctypes.c_int64()
        dist = self._cty()
        cb = self._no_filter if skip_point is None else SKIP_POINT(lambda _c, i: 1 if skip_point(int(i)) else 0)
        check(load().fuel_amd_priority_next(self.handle, int(q), cb, None, ctypes.byref(idx), ctypes.byref(dist)))
        if idx.value < 0:
            return None
        return int(idx.value), np.array(dist.value, dtype=self.index.dtype)[()]

    def all_lower_bound(self, q):
        v = self._cty()
        check(load().fuel_amd_priority_all_lower_bound(self.handle, int(q), ctypes.byref(v)))
        return np.array(v.value, dtype=self.index.dtype)[()]
