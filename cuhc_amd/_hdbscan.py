"""HDBSCAN: the GPU single-linkage tree of mutual reachability, then cluster
extraction in libcuhc's host code.

The extraction gives scikit-learn's results (sklearn/cluster/_hdbscan/_tree.pyx,
tree_to_labels), as the reference's python/cuhc/_hdbscan.py does: the
condensed tree, stabilities, excess-of-mass or leaf selection, labels and
membership probabilities run in C++ (cuhc_condense_tree, cuhc_cluster_stability,
cuhc_select_clusters, cuhc_label_points, cuhc_membership_probabilities), each
linear in the tree size. Only the cluster_selection_epsilon rule, which walks
the small tree of clusters in scikit-learn's set order, runs here.
"""

import numbers

import numpy as np
from sklearn.base import BaseEstimator, ClusterMixin

from . import _lib
from ._device import gpu_shape
from ._distances import (METRIC_IDS, as_points, check_distance_range, input_points,
                         validate_metric)

NOISE = -1
CORE_METHOD_IDS = {"auto": 0, "fused": 1, "radix": 2, "learned": 3}
SELECTION_METHOD_IDS = {"eom": 0, "leaf": 1}


def hdbscan_tree(X, min_samples, metric="euclidean", alpha=1.0, p=2.0, core_method="auto"):
    """(single-linkage tree of mutual reachability (n-1, 4), core distances (n,)).
    X: host points, or a float32 matrix in GPU memory (read in place)."""
    points, device = input_points(X, "X")
    source = points if device is None else device
    try:
        check_distance_range(source, metric, p, scale=1.0 / alpha)
        n, d = source.shape
        tree = np.empty((max(n - 1, 0), 4), dtype=np.float64)
        core = np.empty(n, dtype=np.float32)
        args = (n, d, METRIC_IDS[metric], float(p), int(min_samples), float(alpha),
                CORE_METHOD_IDS[core_method], _lib.double_ptr(tree), _lib.float_ptr(core))
        if device is None:
            status = _lib.lib.cuhc_hdbscan_tree(_lib.float_ptr(points), *args)
        else:
            import ctypes
            status = _lib.lib.cuhc_hdbscan_tree_device(ctypes.c_void_p(device.ptr), *args)
            if status == _lib.NON_FINITE:
                raise ValueError("X contains NaN or infinity")
        _lib.check(status)
    finally:
        if device is not None:
            device.close()
    return tree, core


# ---------------------------------------------------------------------------
# Condensed trees: four columns (parent cluster, child point or cluster,
# lambda = 1 / distance, size); the root cluster is numbered n.


class _Columns:
    """A condensed tree as contiguous int64 / float64 columns for the C calls."""

    def __init__(self, condensed):
        parent, child, lam, size = condensed
        self.parent = np.ascontiguousarray(parent, dtype=np.int64)
        self.child = np.ascontiguousarray(child, dtype=np.int64)
        self.lam = np.ascontiguousarray(lam, dtype=np.float64)
        self.size = np.ascontiguousarray(size, dtype=np.int64)
        self.rows = len(self.parent)
        self.root = int(self.parent.min()) if self.rows else 0
        self.top = int(self.parent.max()) if self.rows else -1

    def args(self):
        return (_lib.int64_ptr(self.parent), _lib.int64_ptr(self.child),
                _lib.double_ptr(self.lam), _lib.int64_ptr(self.size), self.rows)

    @property
    def num_clusters(self):
        return self.top - self.root + 1


def condense_tree(Z, min_cluster_size):
    """Condensed tree of a scipy-format single-linkage tree (C++ on the host),
    rows in scikit-learn's _condense_tree order."""
    Z = np.ascontiguousarray(Z, dtype=np.float64)
    n = len(Z) + 1
    room = n + 2 * (n // max(int(min_cluster_size), 1)) + 2
    while True:
        cols = (np.empty(room, np.int64), np.empty(room, np.int64), np.empty(room, np.float64),
                np.empty(room, np.int64))
        rows = _lib.lib.cuhc_condense_tree(
            _lib.double_ptr(Z), n, int(min_cluster_size), room, _lib.int64_ptr(cols[0]),
            _lib.int64_ptr(cols[1]), _lib.double_ptr(cols[2]), _lib.int64_ptr(cols[3]))
        if rows >= 0:
            break
        room = -rows
    return (cols[0][:rows].astype(np.intp), cols[1][:rows].astype(np.intp), cols[2][:rows],
            cols[3][:rows].astype(np.intp))


def condense_tree_reference(Z, min_cluster_size):
    """The condensed tree by a second, independent method (NumPy, one tree level
    at a time); the same rows in the same order as condense_tree."""
    Z = np.asarray(Z, dtype=np.float64)
    n = len(Z) + 1
    kids = Z[:, :2].astype(np.intp)
    height = Z[:, 2]
    leaves = np.ones(2 * n - 1, dtype=np.intp)
    leaves[n:] = Z[:, 3].astype(np.intp)

    def points_below(node):
        found, level = [], np.array([node], dtype=np.intp)
        while level.size:
            found.append(level[level < n])
            level = kids[level[level >= n] - n].ravel()
        return np.concatenate(found)

    out = []
    owner = {2 * n - 2: n}  # merge-tree node -> condensed cluster still growing
    fresh = n + 1
    level = [2 * n - 2]
    while level:
        deeper = []
        for node in level:
            if node < n or node not in owner:
                continue
            a, b = kids[node - n]
            dist = height[node - n]
            lam = np.inf if dist <= 0.0 else 1.0 / dist
            big = [leaves[a] >= min_cluster_size, leaves[b] >= min_cluster_size]
            me = owner[node]
            for side, keep in zip((a, b), big):
                if all(big):
                    owner[side] = fresh
                    out.append(([me], [fresh], [lam], [leaves[side]]))
                    fresh += 1
                elif keep:
                    owner[side] = me
                else:
                    pts = points_below(side)
                    out.append((np.full(pts.size, me), pts, np.full(pts.size, lam),
                                np.ones(pts.size, dtype=np.intp)))
            deeper.extend((a, b))
        level = deeper
    if not out:
        empty = np.empty(0, dtype=np.intp)
        return empty, empty.copy(), np.empty(0), empty.copy()
    return tuple(np.concatenate([np.asarray(part[i]) for part in out]).astype(dtype)
                 for i, dtype in enumerate((np.intp, np.intp, np.float64, np.intp)))


def _stability_array(cols):
    out = np.zeros(max(cols.num_clusters, 0))
    _lib.check(_lib.lib.cuhc_cluster_stability(*cols.args(), _lib.double_ptr(out)))
    return out


def compute_stability(condensed):
    """{cluster id: stability}: the sum over the cluster's rows, in row order, of
    (lambda - the cluster's birth lambda) * size."""
    cols = _Columns(condensed)
    return dict(zip(range(cols.root, cols.top + 1), _stability_array(cols).tolist()))


def _epsilon_rule(cols, candidates, epsilon, allow_single_cluster):
    """scikit-learn's epsilon search over the tree of clusters: a candidate that
    split off below distance epsilon is replaced by its nearest ancestor that
    split off above it (the root's child next to the root, or the root when a
    single cluster is allowed). Candidates are visited in the iteration order of
    the set scikit-learn builds from the same list, which decides overlaps."""
    is_cluster_row = cols.size > 1
    up = dict(zip(cols.child[is_cluster_row].tolist(), cols.parent[is_cluster_row].tolist()))
    born = dict(zip(cols.child[is_cluster_row].tolist(), cols.lam[is_cluster_row].tolist()))
    below = {}
    for c, p in up.items():
        below.setdefault(p, []).append(c)

    def climb(c):
        while True:
            p = up[c]
            if p == cols.root:
                return p if allow_single_cluster else c
            if 1.0 / born[p] > epsilon:
                return p
            c = p

    def subtree(c):
        stack, out = [c], []
        while stack:
            x = stack.pop()
            out.append(x)
            stack.extend(below.get(x, ()))
        return out

    chosen, covered = [], set()
    for c in set(candidates):
        if 1.0 / born[c] >= epsilon:
            chosen.append(c)
        elif c not in covered:
            top = climb(c)
            chosen.append(top)
            covered.update(x for x in subtree(top) if x != top)
    return set(chosen)


def _cluster_tree_leaves(cols):
    """Leaves of the tree of clusters in depth-first order (children in row order)."""
    rows = cols.size > 1
    if not rows.any():
        return []
    below = {}
    for p, c in zip(cols.parent[rows].tolist(), cols.child[rows].tolist()):
        below.setdefault(p, []).append(c)
    out, stack = [], [cols.root]
    while stack:
        c = stack.pop()
        kids = below.get(c)
        if kids:
            stack.extend(reversed(kids))
        else:
            out.append(c)
    return out


def _select(cols, stab, method, allow_single_cluster, epsilon, max_cluster_size):
    if cols.rows == 0:
        return set()
    stab = np.array(stab, dtype=np.float64)  # updated in place by excess of mass
    chosen = np.zeros(cols.num_clusters, dtype=np.uint8)
    _lib.check(_lib.lib.cuhc_select_clusters(
        *cols.args(), _lib.double_ptr(stab), SELECTION_METHOD_IDS[method],
        int(bool(allow_single_cluster)), -1 if max_cluster_size is None else int(max_cluster_size),
        chosen.ctypes.data_as(_lib._u8p)))
    picked = (np.flatnonzero(chosen) + cols.root).tolist()
    if epsilon == 0.0 or not (cols.size > 1).any():
        return set(picked)
    # The epsilon rule starts from the candidates in scikit-learn's list order:
    # descending ids for excess of mass, depth-first leaves for leaf selection.
    first = cols.root if allow_single_cluster else cols.root + 1
    if method == "eom":
        candidates = sorted(picked, reverse=True)
        if candidates == [cols.root]:
            return set(candidates) if allow_single_cluster else set()
    else:
        candidates = _cluster_tree_leaves(cols)
    chosen_eps = _epsilon_rule(cols, candidates, epsilon, allow_single_cluster)
    return {c for c in chosen_eps if first <= c <= cols.top}


def select_clusters(condensed, stability, cluster_selection_method="eom",
                    allow_single_cluster=False, cluster_selection_epsilon=0.0,
                    max_cluster_size=None):
    """The selected cluster ids (scikit-learn's _get_clusters before labelling);
    stability is {cluster id: value} as compute_stability returns it."""
    cols = _Columns(condensed)
    stab = [stability[c] for c in range(cols.root, cols.top + 1)]
    return _select(cols, stab, cluster_selection_method, allow_single_cluster,
                   cluster_selection_epsilon, max_cluster_size)


def _selected_mask(cols, clusters):
    mask = np.zeros(max(cols.num_clusters, 0), dtype=np.uint8)
    for c in clusters:
        mask[c - cols.root] = 1
    return mask


def label_points(condensed, clusters, allow_single_cluster, cluster_selection_epsilon):
    """Label of every point: the rank of its selected cluster among the sorted
    selected ids, or -1 (noise)."""
    cols = _Columns(condensed)
    labels = np.empty(cols.root, dtype=np.int64)
    _lib.check(_lib.lib.cuhc_label_points(
        *cols.args(), _selected_mask(cols, clusters).ctypes.data_as(_lib._u8p),
        int(bool(allow_single_cluster)), float(cluster_selection_epsilon),
        _lib.int64_ptr(labels)))
    return labels.astype(np.intp)


def membership_probabilities(condensed, clusters, labels):
    """Membership strength of each point in its cluster, 0 for noise."""
    cols = _Columns(condensed)
    labels = np.ascontiguousarray(labels, dtype=np.int64)
    out = np.zeros(cols.root, dtype=np.float64)
    _lib.check(_lib.lib.cuhc_membership_probabilities(
        *cols.args(), _selected_mask(cols, clusters).ctypes.data_as(_lib._u8p),
        _lib.int64_ptr(labels), _lib.double_ptr(out)))
    return out


def tree_to_labels(Z, min_cluster_size, cluster_selection_method="eom",
                   allow_single_cluster=False, cluster_selection_epsilon=0.0,
                   max_cluster_size=None):
    """(labels, probabilities) from a single-linkage tree, as scikit-learn's
    tree_to_labels."""
    cols = _Columns(condense_tree(Z, min_cluster_size))
    clusters = _select(cols, _stability_array(cols), cluster_selection_method,
                       allow_single_cluster, cluster_selection_epsilon, max_cluster_size)
    mask = _selected_mask(cols, clusters).ctypes.data_as(_lib._u8p)
    labels = np.empty(cols.root, dtype=np.int64)
    probabilities = np.zeros(cols.root, dtype=np.float64)
    _lib.check(_lib.lib.cuhc_label_points(
        *cols.args(), mask, int(bool(allow_single_cluster)), float(cluster_selection_epsilon),
        _lib.int64_ptr(labels)))
    _lib.check(_lib.lib.cuhc_membership_probabilities(
        *cols.args(), mask, _lib.int64_ptr(labels), _lib.double_ptr(probabilities)))
    return labels.astype(np.intp), probabilities


def _require(ok, message):
    if not ok:
        raise ValueError(message)


def _is_int(value, lowest):
    return isinstance(value, numbers.Integral) and value >= lowest


class HDBSCAN(ClusterMixin, BaseEstimator):
    """Density-based clustering on the GPU; parameters and results as sklearn.cluster.HDBSCAN.

    Parameters: min_cluster_size (5), min_samples (None = min_cluster_size, the
    point itself counted), cluster_selection_epsilon (0.0), max_cluster_size
    (None), metric ("euclidean"), alpha (1.0; mutual reachability is
    max(core(a), core(b), d(a, b) / alpha)), cluster_selection_method ("eom" or
    "leaf"), allow_single_cluster (False). Fitted: labels_ (-1 = noise),
    probabilities_, core_distances_, n_features_in_.

    >>> import numpy as np, cuhc_amd as cuhc
    >>> rng = np.random.default_rng(0)
    >>> X = np.vstack([rng.normal(0, 0.1, (50, 2)), rng.normal(5, 0.1, (50, 2))])
    >>> sorted(set(cuhc.HDBSCAN(min_cluster_size=10).fit_predict(X).tolist()))
    [0, 1]
    """

    def __init__(self, min_cluster_size=5, min_samples=None, cluster_selection_epsilon=0.0,
                 max_cluster_size=None, metric="euclidean", alpha=1.0,
                 cluster_selection_method="eom", allow_single_cluster=False):
        self.min_cluster_size = min_cluster_size
        self.min_samples = min_samples
        self.cluster_selection_epsilon = cluster_selection_epsilon
        self.max_cluster_size = max_cluster_size
        self.metric = metric
        self.alpha = alpha
        self.cluster_selection_method = cluster_selection_method
        self.allow_single_cluster = allow_single_cluster

    def _check_parameters(self):
        mcs, ms, top = self.min_cluster_size, self.min_samples, self.max_cluster_size
        _require(_is_int(mcs, 2), f"min_cluster_size must be an integer >= 2, got {mcs!r}")
        _require(ms is None or _is_int(ms, 1),
                 f"min_samples must be None or an integer >= 1, got {ms!r}")
        _require(top is None or _is_int(top, 1),
                 f"max_cluster_size must be None or an integer >= 1, got {top!r}")
        _require(self.cluster_selection_epsilon >= 0,
                 f"cluster_selection_epsilon must be >= 0, got {self.cluster_selection_epsilon!r}")
        _require(self.alpha > 0, f"alpha must be > 0, got {self.alpha!r}")
        _require(self.cluster_selection_method in SELECTION_METHOD_IDS,
                 "cluster_selection_method must be 'eom' or 'leaf', got "
                 f"{self.cluster_selection_method!r}")
        validate_metric(self.metric, 2.0)

    def fit(self, X, y=None):
        self._check_parameters()
        shape = gpu_shape(X)
        if shape is None:  # host data: converted and checked once
            X = as_points(X, "X")
            shape = X.shape
        k = self.min_samples if self.min_samples is not None else self.min_cluster_size
        _require(k <= shape[0],
                 f"min_samples ({k}) must be at most the number of samples in X ({shape[0]})")
        Z, core = hdbscan_tree(X, k, metric=self.metric, alpha=self.alpha)
        self.labels_, self.probabilities_ = tree_to_labels(
            Z, self.min_cluster_size, self.cluster_selection_method, self.allow_single_cluster,
            self.cluster_selection_epsilon, self.max_cluster_size)
        self._single_linkage_tree_ = Z
        self.core_distances_ = core
        self.n_features_in_ = shape[1]
        return self

