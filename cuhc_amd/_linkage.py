"""Merge trees in scipy's linkage format and flat cuts of them.

Same surface, algorithm choices and messages as reference python/cuhc/_linkage.py.
"""

import numpy as np

from . import _lib
from ._distances import METRIC_IDS, as_points, check_distance_range, validate_metric

LINKAGE_IDS = {"single": 0, "complete": 1, "average": 2, "ward": 3}
ALGORITHM_IDS = {"dense": 0, "mst": 1, "nn_chain": 2, "rnn": 3}
ALGORITHM_LINKAGES = {
    "dense": set(LINKAGE_IDS), "mst": {"single"}, "nn_chain": {"ward"}, "rnn": {"ward"},
}
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
        return {"single": "mst", "ward": "rnn"}.get(method, "dense")
    if algorithm not in ALGORITHM_IDS:
        raise ValueError(
            f"unknown algorithm {algorithm!r}; expected 'auto' or one of {list(ALGORITHM_IDS)}"
        )
    if method not in ALGORITHM_LINKAGES[algorithm]:
        raise ValueError(
            f"algorithm={algorithm!r} supports linkage {sorted(ALGORITHM_LINKAGES[algorithm])}, "
            f"not {method!r}"
        )
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
    validate_linkage(method, metric)
    resolve_algorithm(algorithm, method)
    validate_metric(metric, p)
    X = as_points(X, "X")
    n = X.shape[0]
    # Ward distances grow with cluster size: 2 n_a n_b / (n_a + n_b) <= n.
    check_distance_range(X, metric, p, scale=n if method == "ward" else 1.0)
    algorithm = resolve_algorithm(algorithm, method, n)
    tree = np.empty((max(n - 1, 0), 4), dtype=np.float64)
    if n == 1:
        return tree
    status = _lib.lib.cuhc_linkage(
        _lib.float_ptr(X), n, X.shape[1], METRIC_IDS[metric], float(p), LINKAGE_IDS[method],
        ALGORITHM_IDS[algorithm], _lib.double_ptr(tree))
    if status == _lib.HIP_ERROR_OUT_OF_MEMORY and algorithm == "dense":
        hint = " (ward linkage can use algorithm='rnn', which doesn't need it)" \
            if method == "ward" else ""
        raise MemoryError(
            f"cuhc.linkage: the dense distance matrix for {n} points needs "
            f"{2 * n * (n - 1) / 1e9:.1f} GB of GPU memory, which isn't available{hint}")
    _lib.check(status)
    return tree


def _ward_rnn_with_options(X, min_progress, dense_below=0, allow_dense=True):
    """Test hook: Ward via rnn rounds with explicit hand-over settings."""
    X = as_points(X, "X")
    n = X.shape[0]
    tree = np.empty((max(n - 1, 0), 4), dtype=np.float64)
    _lib.check(_lib.lib.cuhc_ward_rnn_with_options(
        _lib.float_ptr(X), n, X.shape[1], float(min_progress), int(dense_below),
        int(bool(allow_dense)), _lib.double_ptr(tree)))
    return tree


def cut_tree(Z, num_merges):
    """Labels 0..k-1 of the leaves after the first num_merges merges of Z.

    Labels follow the order of each flat cluster's root id (np.unique order).
    """
    Z = np.asarray(Z)
    n = Z.shape[0] + 1
    up = np.arange(2 * n - 1)
    kids = Z[:num_merges, :2].astype(np.intp)
    new_ids = n + np.arange(num_merges)
    up[kids[:, 0]] = new_ids
    up[kids[:, 1]] = new_ids
    # Pointer doubling until every node points at its flat cluster's root.
    while True:
        nxt = up[up]
        if np.array_equal(nxt, up):
            break
        up = nxt
    return np.unique(up[:n], return_inverse=True)[1]


def fcluster(Z, t, criterion="maxclust"):
    """Flat clusters of a merge tree (scipy.cluster.hierarchy.fcluster for
    "maxclust" and "distance"); labels start at 1.

    >>> import numpy as np, cuhc_amd as cuhc
    >>> cuhc.fcluster(cuhc.linkage(np.array([[0.0], [1.0], [5.0], [7.0]])), 2)
    array([1, 1, 2, 2])
    """
    Z = np.asarray(Z)
    n = Z.shape[0] + 1
    if criterion == "maxclust":
        if t < 1:
            raise ValueError(f"maxclust needs t >= 1, got {t}")
        num_merges = n - min(int(t), n)
    elif criterion == "distance":
        num_merges = int(np.count_nonzero(Z[:, 2] <= t))
    else:
        raise ValueError(f"unknown criterion {criterion!r}; expected 'maxclust' or 'distance'")
    return cut_tree(Z, num_merges) + 1
