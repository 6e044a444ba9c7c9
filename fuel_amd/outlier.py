"""kNN-based outlier scores on the GPU linear scan, with the API of fuel's
``fuel.outlier`` module for LOF and the kNN outlier score
(python/fuel/outlier/__init__.py:407-431, :588-610)."""
from collections import namedtuple as _namedtuple

import numpy as _np

from . import _lib
from .search import SearchIndex, _ensure_float

OutlierResult = _namedtuple('OutlierResult', ['scores', 'metadata'])


def _meta(scores, label, ascending, baseline, tmin, tmax):
    """OutlierMetadata as built by make_outlier_result (src/outlier/common.rs:143-195)."""
    if scores.size:
        finite_min = _np.min(scores)
        finite_max = _np.max(scores)
        mn, mx = float(finite_min), float(finite_max)
        if _np.isinf(mn) and _np.isinf(mx):
            mn = mx = 0.0
    else:
        mn = mx = 0.0
    return {'label': label, 'ascending': ascending, 'baseline': baseline, 'minimum': mn, 'maximum': mx,
            'theoretical_minimum': tmin, 'theoretical_maximum': tmax}


def _index_for(data, distance, index):
    if index is None or isinstance(index, str):
        return SearchIndex(data, distance=distance, tree=index).index
    if isinstance(index, SearchIndex):
        if index.dtype != data.dtype or index.shape != data.shape:
            raise ValueError("SearchIndex must be built from the same data shape and dtype as the outlier input")
        return index.index
    raise TypeError("index must be a SearchIndex instance or string name")


def local_outlier_factor(data, k, *, distance='euclidean', index=None):
    """Local Outlier Factor (src/outlier/lof.rs:14-81) over the GPU self-kNN."""
    data = _ensure_float(data)
    if k <= 0:
        raise ValueError("k must be greater than 0")
    ix = _index_for(data, distance, index)
    scores = ix.lof(int(k))
    n = data.shape[0]
    k_eff = min(k, max(n - 1, 0))
    baseline = 1.0 if k_eff == 0 else 0.0
    return OutlierResult(scores, _meta(scores, "LOF", False, baseline, 0.0, float('inf')))


def k_nearest_neighbors_outlier(data, k, *, distance='euclidean', index=None):
    """Distance to the k-th neighbour, self excluded (src/outlier/knn.rs:14-32)."""
    data = _ensure_float(data)
    if k <= 0:
        raise ValueError("k must be greater than 0")
    ix = _index_for(data, distance, index)
    scores = ix.knn_outlier(int(k))
    return OutlierResult(scores, _meta(scores, "kNN Outlier Score", False, 0.0, 0.0, float('inf')))
