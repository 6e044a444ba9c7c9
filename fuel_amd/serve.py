"""Serving: continuous batching of kNN requests over one resident index.

fuel itself answers one batch per ``SearchIndex.knn`` call
(python/search.rs:560-569); a server keeps the index resident in HBM and
answers requests submitted from any thread.  Worker threads (one stream and
one set of query workspaces each) coalesce every request waiting when they
become free into one launch, within ``max_batch`` queries and the HBM budget,
so requests arriving while other batches run join the next launch and overlap
with the running ones on the device.  Every answer equals
``SearchIndex.knn(query, k)`` for that request alone.
"""
import ctypes

import numpy as np

from . import _lib


class Server:
    """Request loop over a resident index (a ``SearchIndex`` or ``_lib.Index``)."""

    def __init__(self, index, workers=2, max_batch=4096, hbm_budget=4 << 30):
        ix = getattr(index, "index", index)
        if not isinstance(ix, _lib.Index):
            raise TypeError("index must be a SearchIndex or fuel_amd._lib.Index")
        self._ix = ix  # keeps the index alive while the server runs
        self.max_batch = int(max_batch)
        opts = _lib.ServerOpts(int(workers), 0, self.max_batch, int(hbm_budget))
        h = ctypes.c_void_p()
        _lib.check(_lib.load().fuel_amd_server_create(ix.handle, ctypes.byref(opts), ctypes.byref(h)))
        self.handle = h
        self._k = {}

    def submit(self, query, k):
        """Queue the rows of `query` (m x d); returns a ticket."""
        q = np.ascontiguousarray(query, dtype=self._ix.dtype)
        if q.ndim != 2 or q.shape[1] != self._ix.d:
            raise ValueError("query shape mismatch")
        t = ctypes.c_uint64()
        _lib.check(_lib.load().fuel_amd_server_submit(self.handle, q.ctypes.data, q.shape[0], int(k),
                                                       ctypes.byref(t)))
        self._k[t.value] = (q.shape[0], int(k))
        return t.value

    def wait(self, ticket):
        """Block until the request is answered: (indices, distances, latency_ms,
        queries in the launch that answered it)."""
        if ticket not in self._k:
            raise _lib.FuelAmdError(_lib.ERR_INVALID, "unknown ticket")
        m, k = self._k.pop(ticket)
        idx = np.empty((m, k), np.int64)
        dist = np.empty((m, k), self._ix.dtype)
        lat = ctypes.c_double()
        bq = ctypes.c_int64()
        _lib.check(_lib.load().fuel_amd_server_wait(self.handle, ticket, idx.ctypes.data, dist.ctypes.data,
                                                     ctypes.byref(lat), ctypes.byref(bq)))
        return idx, dist, lat.value, bq.value

    def stats(self, reset=False):
        st = _lib.ServeStats()
        _lib.check(_lib.load().fuel_amd_server_stats(self.handle, 1 if reset else 0, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in _lib.ServeStats._fields_}

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().fuel_amd_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
