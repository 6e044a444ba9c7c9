"""GPU linear-scan search with the API of fuel's ``fuel.search`` module.

Mirrors ``python/fuel/search.py`` (SearchIndex, knn_search, range_search,
pairwise_distances) for the linear-scan path, backed by libfuel_amd.so.  Results
are identical to fuel's CPU linear scan: same f32/f64 distance values (fuel's
AVX2 kernel, src/distance/squared_euclidean.rs:13-67, sqrt'd for 'euclidean'),
same (distance, index) order, same -1/+inf padding and exclusion semantics.

Differences in scope (documented, not silent):
  * only the 'euclidean'/'l2' and 'sqeuclidean'/'squared_euclidean' distances
    run here; other fuel distance names raise NotImplementedError;
  * tree='auto' (or None) selects the linear scan; the exact tree indexes
    ('vp', 'kd', 'cover') are not part of this engine and raise
    NotImplementedError.
"""
import math as _math

import numpy as _np

from . import _lib

_TREE_NAME_MAP = {
    'auto': 'auto', 'vp': 'vp', 'vptree': 'vp', 'cover': 'cover', 'covertree': 'cover', 'ct': 'cover',
    'ball': 'cover', 'balltree': 'cover', 'kd': 'kd', 'kdtree': 'kd', 'linear': 'linear', 'brute': 'linear',
    'bruteforce': 'linear',
}


def _ensure_float(data):
    """float32/float64 pass through, everything else becomes float64
    (python/fuel/_dispatch.py:8-12)."""
    data = _np.asarray(data)
    if data.dtype not in (_np.float32, _np.float64):
        return data.astype(_np.float64)
    return data


def _normalize_search_tree_name(tree):
    """python/fuel/search.py:37-47."""
    if tree is None:
        return 'auto'
    if not isinstance(tree, str):
        raise TypeError("tree must be a string or None")
    key = tree.lower().replace('_', '').replace('-', '')
    if key in _TREE_NAME_MAP:
        return _TREE_NAME_MAP[key]
    raise ValueError(
        f"unknown tree '{tree}', valid values are 'auto', 'vp', 'kd', 'cover', 'linear', 'vptree', 'ct', "
        "'covertree', 'brute', or 'brute_force'")


def _resolve_tree(tree):
    tree = _normalize_search_tree_name(tree)
    if tree == 'auto':
        return 'linear'
    if tree != 'linear':
        raise NotImplementedError(f"tree '{tree}' is not served by the GPU linear-scan engine")
    return tree


def _condensed_length_to_n(length):
    """python/fuel/search.py:50-58."""
    if length < 0:
        raise ValueError("condensed distance vector length must be non-negative")
    n = (1 + _math.isqrt(1 + 8 * length)) // 2
    if n * (n - 1) // 2 != length:
        raise ValueError("condensed distance vector length must be n*(n-1)/2 for some integer n")
    return n


def _square_to_condensed(square):
    if square.ndim != 2 or square.shape[0] != square.shape[1]:
        raise ValueError("precomputed distance matrix must be square")
    n = square.shape[0]
    return square[_np.tril_indices(n, -1)].copy()


def _condensed_to_square(condensed):
    if condensed.ndim != 1:
        raise ValueError("precomputed condensed distance vector must be one-dimensional")
    n = _condensed_length_to_n(condensed.shape[0])
    square = _np.zeros((n, n), dtype=condensed.dtype)
    tril = _np.tril_indices(n, -1)
    square[tril] = condensed
    square[(tril[1], tril[0])] = condensed
    return square


def _scipy_to_internal_condensed(data):
    """scipy's upper-triangular pdist order -> fuel's internal lower-triangular
    order i(i-1)/2 + j (python/fuel/search.py:77-89)."""
    n = _condensed_length_to_n(data.shape[0])
    tril_i, tril_j = _np.tril_indices(n, -1)
    scipy_idx = tril_j * (2 * n - tril_j - 1) // 2 + (tril_i - tril_j - 1)
    return data[scipy_idx]


def pairwise_distances(data, distance='euclidean', form='condensed'):
    """Pairwise distances (python/fuel/search.py:92-150).  The condensed form is
    fuel's internal lower-triangular order, pair (i, j), i > j, at i(i-1)/2 + j."""
    data = _ensure_float(data)
    distance = distance.lower()
    form = form.lower()
    if form not in {'condensed', 'square'}:
        raise ValueError("form must be 'condensed' or 'square'")
    if distance == 'precomputed':
        if data.ndim == 1:
            internal = _scipy_to_internal_condensed(data)
            return internal if form == 'condensed' else _condensed_to_square(internal)
        if data.ndim == 2:
            if data.shape[0] != data.shape[1]:
                raise ValueError("precomputed distance matrix must be square")
            return data if form == 'square' else _square_to_condensed(data)
        raise ValueError("precomputed data must be a square matrix or condensed distance vector")
    if data.ndim != 2:
        raise ValueError("data must be a 2D array of observations when distance is not 'precomputed'")
    return _lib.pairwise(_np.ascontiguousarray(data), distance, form)


class SearchIndex:
    """Persistent GPU linear-scan index (python/fuel/search.py:205-313).

    The rows are copied to HBM once; ``knn`` and ``radius_search`` then stream
    only the queries.  ``mode`` selects how candidates are generated ('auto',
    'exact', 'fp32', 'bf16', 'bf16r' -- bf16 on residuals against coarse
    centroids, for clustered data whose norms are large next to the neighbour
    distances -- or 'fp8r', the same with an FP8 E4M3 resident copy of one byte
    per coordinate and a power-of-two scale per row); every mode returns the
    exact fuel answer.  ``widen`` (0: the mode's default, 4, or 16 for 'fp8r')
    sets the candidates kept per query, widen * k.

    ``precompute=max_k`` fills fuel's precomputed kNN table (every point's
    neighbours with ties up to max_k) with one GPU call; self-queries with
    k <= max_k are then served from it.  fuel's precomputed linear index answers
    queries from another array with point 0's table row (its query index is
    never set); this index answers them with a scan instead.
    """

    def __init__(self, data, *, distance='euclidean', tree=None, seed=None, precompute=None, mode='auto',
                 widen=0, device=0):
        data = _ensure_float(data)
        if data.ndim != 2:
            raise ValueError("data must be a 2D array")
        self.data = data
        self.distance = distance
        self.seed = seed
        self.precompute = precompute
        if precompute is not None:
            precompute = int(precompute)
            if precompute < 0:
                raise OverflowError("can't convert negative int to unsigned")
        self.dtype = data.dtype
        self.shape = data.shape
        self.tree = _resolve_tree(tree)
        self._source_shape = data.shape
        self._source_strides = data.strides
        self._source_ptr = data.__array_interface__['data'][0]
        self.index = _lib.Index(_np.ascontiguousarray(data), distance, mode=mode, widen=widen, device=device)
        # PrecomputedKnnSearcher::new (src/search/precomputed.rs:36-56): every data
        # point's kNN with ties up to max_k = min(precompute, n), self included,
        # filled by one GPU call.
        self._table = None
        self._max_k = 0
        if precompute is not None:
            n = data.shape[0]
            self._max_k = min(precompute, n)
            if n > 0 and self._max_k > 0:
                self._table = self.index.knn_ties(_np.ascontiguousarray(data), self._max_k)

    def _table_knn(self, m, k, exclude_self):
        """SearchIndex.knn on the precomputed arm for self-queries
        (python/search.rs:571-589 with batch_knn_with_query :1154-1173): search
        k (+1 with exclude_self) through subset_with_ties (precomputed.rs:60-67),
        drop the query's own index, pack to (m, k) with -1 / +inf."""
        offs, idx, dist = self._table
        query_k = k + 1 if exclude_self else k
        out_i = _np.full((m, k), -1, _np.int64)
        out_d = _np.full((m, k), _np.inf, self.dtype)
        for i in range(m):
            row_i = idx[offs[i]:offs[i + 1]]
            row_d = dist[offs[i]:offs[i + 1]]
            if query_k == 0 or row_i.size == 0:
                continue
            kth = row_d[min(query_k, row_i.size) - 1]
            cut = int(_np.searchsorted(row_d, kth, side='right'))
            ri, rd = row_i[:cut], row_d[:cut]
            if exclude_self:
                keep = ri != i
                ri, rd = ri[keep], rd[keep]
            c = min(k, ri.size)
            out_i[i, :c] = ri[:c]
            out_d[i, :c] = rd[:c]
        return out_i, out_d

    def _same_source(self, query):
        return (query.shape == self._source_shape and query.strides == self._source_strides
                and query.__array_interface__['data'][0] == self._source_ptr)

    def _check_query(self, query):
        query = _ensure_float(query)
        if query.dtype != self.dtype:
            query = query.astype(self.dtype)
        if query.ndim != 2 or query.shape[1] != self.shape[1]:
            raise ValueError(
                f"query shape {query.shape} is not compatible with index data shape {self.shape}")
        return query

    def knn(self, query, k, *, exclude_self=None):
        """k nearest neighbours as (int64[m, k] padded -1, float[m, k] padded +inf).

        With exclude_self the reference's linear-scan path searches k neighbours
        and then drops the query's own index (python/search.rs:485-488), so a
        self-query yields k-1 neighbours plus padding unless the k-th distance is
        tied; this is reproduced exactly."""
        query_in = query
        query = self._check_query(query)
        if exclude_self is None:
            exclude_self = self._same_source(query_in if isinstance(query_in, _np.ndarray) else query)
        elif exclude_self and not self._same_source(query_in if isinstance(query_in, _np.ndarray) else query):
            raise ValueError(
                "exclude_self=True is only supported when query refers to the same underlying array as data")
        if k < 0:
            raise OverflowError("can't convert negative int to unsigned")
        k = int(k)
        if self.precompute is not None and self._same_source(query_in if isinstance(query_in, _np.ndarray) else query):
            # precomputed arm: k (+1 with exclude_self) with ties, then drop self and
            # truncate -- the search_knn + retain + pack of the reference; within
            # max_k served from the table, beyond it by the scan (same result)
            if (k + 1 if exclude_self else k) <= self._max_k and self._table is not None:
                return self._table_knn(query.shape[0], k, exclude_self)
            return self.index.knn(query, k, _lib.EXCLUDE_FOR_EACH_KNN if exclude_self else _lib.EXCLUDE_NONE)
        return self.index.knn(query, k, _lib.EXCLUDE_PY_LINEAR if exclude_self else _lib.EXCLUDE_NONE)

    def radius_search(self, query, radius, *, exclude_self=None):
        """All neighbours within radius as (list of int64 arrays, list of float
        arrays), each sorted by (distance, index) (python/search.rs:1218-1235)."""
        query_in = query
        query = self._check_query(query)
        src = query_in if isinstance(query_in, _np.ndarray) else query
        if exclude_self is None:
            exclude_self = self._same_source(src)
        elif exclude_self and not self._same_source(src):
            raise ValueError(
                "exclude_self=True is only supported when query refers to the same underlying array as data")
        offs, idx, dist = self.index.range(query, radius, exclude_self)
        idx_list = [idx[offs[i]:offs[i + 1]].copy() for i in range(query.shape[0])]
        dist_list = [dist[offs[i]:offs[i + 1]].copy() for i in range(query.shape[0])]
        return idx_list, dist_list

    def __repr__(self):
        return f"<SearchIndex tree={self.tree!r} dtype={self.dtype!r} shape={self.shape!r}>"


def knn_search(data, query, k, *, exclude_self=None, distance='euclidean', tree='auto', seed=None):
    """python/fuel/search.py:315-361."""
    data = _ensure_float(data)
    query = _ensure_float(query)
    index = SearchIndex(data, distance=distance, tree=tree, seed=seed)
    return index.knn(query, k, exclude_self=exclude_self)


def range_search(data, query, radius, *, exclude_self=None, distance='euclidean', tree='auto', seed=None):
    """python/fuel/search.py:364-408."""
    data = _ensure_float(data)
    query = _ensure_float(query)
    index = SearchIndex(data, distance=distance, tree=tree, seed=seed)
    return index.radius_search(query, radius, exclude_self=exclude_self)
