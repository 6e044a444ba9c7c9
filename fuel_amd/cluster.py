"""Density clustering on the GPU range search, with the API of fuel's
``fuel.cluster`` DBSCAN and OPTICS (python/fuel/cluster/_dbscan.py).

One GPU call computes every point's eps-neighbourhood exactly (fuel's restated
distance, eps inclusive, sorted by (distance, index)); the label and ordering
walks follow src/cluster/dbscan.rs, src/cluster/parallel_dbscan.rs and
src/cluster/optics.rs on the host.  Results equal fuel's with a linear-scan
index (tree indexes return range results in another order, which only the
parallel variant's border points can see).
"""
import math

import numpy as _np

from . import _lib
from .search import SearchIndex, _ensure_float


def _index_for(data, distance, index):
    if isinstance(index, SearchIndex):
        return index.index
    if index not in (None, 'linear', 'brute', 'brute_force', 'auto'):
        raise NotImplementedError(f"index {index!r}: only the linear scan runs on the GPU engine")
    return _lib.Index(_np.ascontiguousarray(data), distance)


def dbscan(data, eps, min_points, *, distance="euclidean", variant="dbscan", index=None):
    """DBSCAN labels (int64, -1 = noise), python/fuel/cluster/_dbscan.py:6-38.

    variant 'dbscan' follows src/cluster/dbscan.rs:14-83, 'parallel'
    src/cluster/parallel_dbscan.rs:16-113.  eps is cast to the data dtype."""
    data = _ensure_float(data)
    if variant not in ('dbscan', 'parallel'):
        raise ValueError("unsupported variant: {}".format(variant))
    if eps < 0:
        raise ValueError("eps must be non-negative")
    if min_points < 0:
        raise OverflowError("can't convert negative int to unsigned")
    if variant == 'parallel' and min_points == 0:
        raise ValueError("min_points must be greater than 0")
    ix = _index_for(data, distance, index)
    return ix.dbscan(eps, int(min_points), 0 if variant == 'dbscan' else 1)


class OpticsResult:
    """OPTICS output with the accessors of fuel's OpticsResultF32/F64
    (python/dbscan.rs:96-150)."""

    def __init__(self, ordering, reachability, core_distance, predecessor, labels):
        self._ordering = ordering
        self._reachability = reachability
        self._core_distance = core_distance
        self._predecessor = predecessor
        self._labels = labels

    def ordering(self):
        return self._ordering.copy()

    def reachability(self):
        return self._reachability.copy()

    def core_distance(self):
        """Core distance per point, NaN where the point is not a core point."""
        return self._core_distance.copy()

    def predecessor(self):
        """Predecessor per point, -1 where there is none."""
        return self._predecessor.copy()

    def labels(self):
        """DBSCAN-style labels at max_eps (src/cluster/optics.rs:345-377)."""
        return self._labels.copy()

    def extract_xi(self, xi, min_points):
        """Xi cluster extraction, src/cluster/optics.rs:148-304."""
        return extract_xi_labels(self._ordering, self._reachability, self._predecessor, xi, min_points)


def optics(data, max_eps, min_points, *, distance="euclidean", index=None):
    """OPTICS ordering and reachability, python/fuel/cluster/_dbscan.py:41-72
    (src/cluster/optics.rs:56-133)."""
    data = _ensure_float(data)
    if max_eps < 0:
        raise ValueError("eps must be non-negative")
    if min_points <= 0:
        raise ValueError("min_points must be greater than 0")
    ix = _index_for(data, distance, index)
    return OpticsResult(*ix.optics(max_eps, int(min_points)))


def extract_xi_labels(ordering, reachability, predecessor, xi, min_points):
    """Flat Xi labels (-1 = noise) from an OPTICS ordering.

    Steep-down areas (a drop by the factor 1 - xi to the next point) are kept in
    a list and filtered by the running maximum in between; every steep-up area
    closes candidate intervals against them, corrected by the predecessor
    relation (an interval ends where a point's predecessor lies inside it), and
    intervals shorter than min_points are dropped.  Smaller intervals claim
    their points first, so nested clusters keep their points (the same rules
    as src/cluster/optics.rs:148-304, :379-453)."""
    if not (0.0 < xi < 1.0):
        raise ValueError("xi must be in (0, 1)")
    if min_points <= 0:
        raise ValueError("min_points must be greater than 0")
    ordering = _np.asarray(ordering, _np.int64)
    reach_pts = _np.asarray(reachability)
    n = ordering.shape[0]
    labels = _np.full(n, -1, _np.int64)
    if n == 0:
        return labels
    ftype = reach_pts.dtype.type
    ixi = ftype(1) - ftype(xi)
    one = ftype(1)
    r = [ftype(v) for v in reach_pts[ordering]]
    pred = _np.asarray(predecessor, _np.int64)
    pos = _np.empty(n, _np.int64)
    pos[ordering] = _np.arange(n)
    inf = ftype(math.inf)

    def finite(v):
        return math.isfinite(float(v))

    def steep_up(i, f):
        return finite(r[i]) and (i + 1 >= n or r[i] <= r[i + 1] * f)

    def steep_down(i, f):
        return i + 1 < n and finite(r[i + 1]) and r[i + 1] <= r[i] * f

    def next_r(i):
        return r[i + 1] if i + 1 < n else inf

    def pred_filter(cstart, cend):
        if cend >= n:
            return n - 1
        startval = r[cstart]
        while cend > cstart:
            if r[cend] < startval:
                break
            p = pred[ordering[cend]]
            if p >= 0 and cstart <= pos[p] < cend:
                break
            cend -= 1
        return cend

    mib = ftype(0)
    areas = []  # [start, maximum, mib]
    intervals = []

    def refresh(m):
        kept = []
        for a in areas:
            if a[1] * ixi <= m:
                continue
            if m > a[2]:
                a[2] = m
            kept.append(a)
        areas[:] = kept

    scan = 0
    while scan < n:
        mib = max(mib, r[scan])
        if steep_down(scan, ixi):
            refresh(mib)
            startval = r[scan]
            mib = ftype(0)
            start = end = scan
            scan += 1
            while scan < n:
                if steep_down(scan, ixi):
                    end = scan
                elif not steep_down(scan, one) or scan - end > min_points:
                    break
                scan += 1
            areas.append([start, startval, ftype(0)])
            continue
        if steep_up(scan, ixi):
            refresh(mib)
            end = scan
            mib = r[scan]
            esuccr = next_r(scan)
            while finite(esuccr) and scan < n:
                scan += 1
                if scan >= n:
                    break
                if steep_up(scan, ixi):
                    end = scan
                    mib = r[scan]
                    esuccr = next_r(scan)
                elif not steep_up(scan, one) or scan - end > min_points:
                    break
            if not finite(esuccr) and scan < n:
                scan += 1
            for a in reversed(areas):
                cstart = a[0]
                cend = pred_filter(cstart, end)
                e_u = r[cend + 1] if cend + 1 < n else inf
                if a[2] > min(a[1], e_u) * ixi:
                    continue
                if a[1] * ixi >= e_u:
                    while cstart < cend and r[cstart + 1] * ixi > e_u:
                        cstart += 1
                elif e_u * ixi >= a[1]:
                    while cend > cstart and r[cend] * ixi > a[1]:
                        cend -= 1
                cend = pred_filter(cstart, cend)
                if cend + 1 - cstart < min_points:
                    continue
                intervals.append((cstart, cend))
            continue
        scan += 1

    intervals.sort(key=lambda iv: (iv[1] - iv[0], iv[0]))
    claimed = _np.zeros(n, bool)
    cid = 0
    for cstart, cend in intervals:
        pts = ordering[cstart:cend + 1]
        free = pts[~claimed[pts]]
        if free.size:
            claimed[free] = True
            labels[free] = cid
            cid += 1
    return labels
