"""Many independent self-kNN problems in one batched pass.

Each problem is a dataset searched against itself with fuel's semantics
(by default the loop of for_each_knn, src/outlier/common.rs:12-49: k + 1
searched, self dropped, truncated to k).  Problems of one dimension share one
persistent MFMA scan; the whole level-0 pass is a captured hipGraph replayed
by every ``run``; the autotuner (``mode="auto"``) picks, per dimension, the
MFMA precision, candidate widening and split size by timing them on a sample
of that dimension's problems, then its four leading choices on all of them.
Results equal ``SearchIndex.knn`` /
``fuel_amd_knn`` on each problem alone.
"""
import ctypes
import json

import numpy as np

from . import _lib


class BatchKnn:
    def __init__(self, datasets, k, metric="euclidean", exclude=_lib.EXCLUDE_FOR_EACH_KNN, mode="auto", widen=4,
                 graph=True, device=0, tune_cache=None):
        """``tune_cache``: optional path of a text file that keeps the autotuner's
        choices across processes (read here, appended to for every group tuned)."""
        import torch
        self._dev = torch.device(f"cuda:{device}")
        self._data, self._idx, self._dist, probs = [], [], [], []
        ks = list(k) if np.ndim(k) else [int(k)] * len(datasets)
        for X, kk in zip(datasets, ks):
            t = torch.as_tensor(X, device=self._dev)
            if t.dtype != torch.float32 or t.ndim != 2:
                raise ValueError("datasets must be 2-D float32 arrays")
            n, d = t.shape
            ld = max(8, (d + 7) // 8 * 8)
            if t.stride(1) != 1 or t.stride(0) != ld or t.data_ptr() % 32:
                buf = torch.zeros((n, ld), dtype=torch.float32, device=self._dev)
                buf[:, :d] = t
                t = buf
            oi = torch.empty((n, kk), dtype=torch.int64, device=self._dev)
            od = torch.empty((n, kk), dtype=torch.float32, device=self._dev)
            self._data.append(t)
            self._idx.append(oi)
            self._dist.append(od)
            probs.append(_lib.Problem(t.data_ptr(), n, ld, d, kk, oi.data_ptr(), od.data_ptr()))
        arr = (_lib.Problem * max(1, len(probs)))(*probs)
        self._tune_cache = None if tune_cache is None else str(tune_cache).encode()
        opts = _lib.BatchOpts(_lib.MODES[mode], int(widen), int(device), 0 if graph else 1, self._tune_cache)
        h = ctypes.c_void_p()
        _lib.check(_lib.load().fuel_amd_batch_create(arr, len(probs), _lib.metric_code(metric), int(exclude),
                                                      ctypes.byref(opts), ctypes.byref(h)))
        self.handle = h

    def run(self):
        _lib.check(_lib.load().fuel_amd_batch_run(self.handle))

    def results(self):
        """(indices, distances) per problem, as host arrays."""
        return [(i.cpu().numpy(), d.cpu().numpy()) for i, d in zip(self._idx, self._dist)]

    def device_results(self):
        return list(zip(self._idx, self._dist))

    def describe(self):
        buf = ctypes.create_string_buffer(1 << 20)
        _lib.check(_lib.load().fuel_amd_batch_describe(self.handle, buf, len(buf)))
        return json.loads(buf.value.decode())

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().fuel_amd_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
