"""Merge trees in scipy's linkage format and flat cuts of them.

The names, ids, algorithm rules and error messages are the reference's
interface (python/cuhc/_linkage.py); the code is written here.
"""

import numpy as np
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

from . import _lib
from ._distances import (METRIC_IDS, as_points, check_distance_range, input_points,
                         validate_metric)

LINKAGE_IDS = {"single": 0, "complete": 1, "average": 2, "ward": 3}
ALGORITHM_IDS = {"dense": 0, "mst": 1, "nn_chain": 2, "rnn": 3}
# Which linkages each algorithm builds, and the one "auto" picks per linkage.
ALGORITHM_LINKAGES = {
    "dense": set(LINKAGE_IDS), "mst": {"single"}, "nn_chain": {"ward"}, "rnn": {"ward"},
}
_AUTO = {"single": "mst", "ward": "rnn", "complete": "dense", "average": "dense"}
# Linkages defined for euclidean distances only.
EUCLIDEAN_ONLY = {"ward"}


def validate_linkage(method, metric):
    if method not in LINKAGE_IDS:
        raise ValueError(f"unknown linkage method {method!r}; expected one of {list(LINKAGE_IDS)}")
    if method in EUCLIDEAN_ONLY and METRIC_IDS.get(metric) != METRIC_IDS["euclidean"]:
        raise ValueError(f"{method} linkage requires metric='euclidean', got {metric!r}")


def resolve_algorithm(algorithm, method, num_points=None):
    """The algorithm to run; "auto" is mst for single, rnn for ward, dense otherwise."""
    if algorithm == "auto":
        return _AUTO[method]
    if algorithm not in ALGORITHM_IDS:
        raise ValueError(
            f"unknown algorithm {algorithm!r}; expected 'auto' or one of {list(ALGORITHM_IDS)}"
        )
    supported = ALGORITHM_LINKAGES[algorithm]
    if method not in supported:
        raise ValueError(
            f"algorithm={algorithm!r} supports linkage {sorted(supported)}, not {method!r}")
    return algorithm


def linkage(X, method="single", metric="euclidean", p=2.0, *, algorithm="auto"):
    """Full merge tree of the rows of X (scipy.cluster.hierarchy.linkage format).

    Every algorithm gives the exact tree: "dense" (condensed float32 matrix in
    GPU memory, any linkage), "mst" (Boruvka, single linkage, O(n) memory),
    "rnn" (reciprocal nearest-neighbour rounds, Ward), "nn_chain" (Ward).

    Returns Z of shape (n - 1, 4), float64: rows [cluster_a, cluster_b, height,
    size]; row i creates cluster n + i.

    >>> import numpy as np, cuhc_amd as cuhc
    >>> cuhc.linkage(np.array([[0.0], [1.0], [5.0], [7.0]]), method="single")
    array([[0., 1., 1., 2.],
           [2., 3., 2., 2.],
           [4., 5., 4., 4.]])
    """
    # Arguments are checked before the points are converted.
    validate_linkage(method, metric)
    chosen = resolve_algorithm(algorithm, method)
    validate_metric(metric, p)
    points, device = input_points(X, "X")
    source = points if device is None else device
    try:
        n, d = source.shape
        # Ward's squared heights reach 2 a b / (a + b) <= n times a squared point distance.
        check_distance_range(source, metric, p, scale=float(n) if method == "ward" else 1.0)
        Z = np.empty((max(n - 1, 0), 4), dtype=np.float64)
        if n < 2:
            return Z
        args = (n, d, METRIC_IDS[metric], float(p), LINKAGE_IDS[method], ALGORITHM_IDS[chosen],
                _lib.double_ptr(Z))
        if device is None:
            status = _lib.lib.cuhc_linkage(_lib.float_ptr(points), *args)
        else:
            # Points read in place from GPU memory (no copy to the host).
            import ctypes
            status = _lib.lib.cuhc_linkage_device(ctypes.c_void_p(device.ptr), *args)
            if status == _lib.NON_FINITE:
                raise ValueError("X contains NaN or infinity")
    finally:
        if device is not None:
            device.close()
    if chosen == "dense" and status == _lib.HIP_ERROR_OUT_OF_MEMORY:
        raise MemoryError(_dense_memory_message(n, method))
    _lib.check(status)
    return Z


def _dense_memory_message(n, method):
    gigabytes = 2 * n * (n - 1) / 1e9  # the condensed float32 matrix
    message = (f"cuhc.linkage: the dense distance matrix for {n} points needs {gigabytes:.1f} GB "
               "of GPU memory, which isn't available")
    if method == "ward":
        message += " (ward linkage can use algorithm='rnn', which doesn't need it)"
    return message


def _ward_rnn_with_options(X, min_progress, dense_below=0, allow_dense=True):
    """Test hook: Ward via rnn rounds with explicit hand-over settings."""
    points = as_points(X, "X")
    n, d = points.shape
    Z = np.empty((max(n - 1, 0), 4), dtype=np.float64)
    _lib.check(_lib.lib.cuhc_ward_rnn_with_options(
        _lib.float_ptr(points), n, d, float(min_progress), int(dense_below),
        int(bool(allow_dense)), _lib.double_ptr(Z)))
    return Z


def cut_tree(Z, num_merges):
    """Labels 0..k-1 of the leaves after the first num_merges merges of Z.

    Labels follow the order of each flat cluster's top node id (the largest id
    in it, since a merge's id exceeds its children's).
    """
    Z = np.asarray(Z)
    n = len(Z) + 1
    nodes = 2 * n - 1
    # The first num_merges merges as edges child -> new cluster; the flat
    # clusters are the connected components of that forest.
    made = n + np.arange(num_merges)
    kids = Z[:num_merges, :2].astype(np.int64)
    edges = coo_matrix((np.ones(2 * num_merges, dtype=np.int8),
                        (kids.T.ravel(), np.concatenate([made, made]))), shape=(nodes, nodes))
    _, component = connected_components(edges, directed=False)
    top = np.zeros(component.max() + 1, dtype=np.int64)
    np.maximum.at(top, component, np.arange(nodes))
    return np.unique(top[component[:n]], return_inverse=True)[1]


def fcluster(Z, t, criterion="maxclust"):
    """Flat clusters of a merge tree (scipy.cluster.hierarchy.fcluster for
    "maxclust" and "distance"); labels start at 1.

    >>> import numpy as np, cuhc_amd as cuhc
    >>> cuhc.fcluster(cuhc.linkage(np.array([[0.0], [1.0], [5.0], [7.0]])), 2)
    array([1, 1, 2, 2])
    """
    Z = np.asarray(Z)
    n = len(Z) + 1
    if criterion == "distance":
        merges = int((Z[:, 2] <= t).sum())
    elif criterion == "maxclust":
        if t < 1:
            raise ValueError(f"maxclust needs t >= 1, got {t}")
        merges = n - min(int(t), n)
    else:
        raise ValueError(f"unknown criterion {criterion!r}; expected 'maxclust' or 'distance'")
    return cut_tree(Z, merges) + 1
