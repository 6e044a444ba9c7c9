"""fuel_amd — MI355X-native exact linear-scan kNN / range / pairwise engine.

A drop-in for fuel's linear-scan search path (SearchIndex knn / radius_search,
pairwise_distances, LOF and the kNN outlier score, and DBSCAN / OPTICS on the
range search) running hand-written HIP kernels for gfx950 behind the C ABI in
include/fuel_amd.h.
"""
from . import _lib
from ._lib import pairwise_sqdist
from .search import SearchIndex, knn_search, pairwise_distances, range_search
from .outlier import OutlierResult, k_nearest_neighbors_outlier, local_outlier_factor
from .cluster import OpticsResult, dbscan, extract_xi_labels, optics
from .serve import Server
from .batch import BatchKnn

__all__ = ["SearchIndex", "knn_search", "range_search", "pairwise_distances", "local_outlier_factor",
           "k_nearest_neighbors_outlier", "OutlierResult", "dbscan", "optics", "OpticsResult",
           "extract_xi_labels", "pairwise_sqdist", "Server", "BatchKnn"]
