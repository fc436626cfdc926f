"""HDBSCAN: GPU single-linkage tree of mutual reachability, then cluster extraction.

The extraction reproduces sklearn's sklearn/cluster/_hdbscan/_tree.pyx
(tree_to_labels: _condense_tree, _compute_stability, _get_clusters,
_do_labelling, get_probabilities) result for result, as the reference does in
python/cuhc/_hdbscan.py; condensing runs in C++ (cuhc_condense_tree). Written
for n in the millions: children lists are precomputed, so selection is linear
in the number of condensed clusters.
"""

import numbers

import numpy as np
from sklearn.base import BaseEstimator, ClusterMixin

from . import _lib
from ._distances import METRIC_IDS, as_points, check_distance_range, validate_metric

NOISE = -1
CORE_METHOD_IDS = {"auto": 0, "fused": 1, "radix": 2, "learned": 3}


def hdbscan_tree(X, min_samples, metric="euclidean", alpha=1.0, p=2.0, core_method="auto"):
    """(single-linkage tree of mutual reachability (n-1, 4), core distances (n,))."""
    X = as_points(X, "X")
    check_distance_range(X, metric, p, scale=1.0 / alpha)
    n = X.shape[0]
    tree = np.empty((max(n - 1, 0), 4), dtype=np.float64)
    core = np.empty(n, dtype=np.float32)
    _lib.check(_lib.lib.cuhc_hdbscan_tree(
        _lib.float_ptr(X), n, X.shape[1], METRIC_IDS[metric], float(p), int(min_samples),
        float(alpha), CORE_METHOD_IDS[core_method], _lib.double_ptr(tree), _lib.float_ptr(core)))
    return tree, core


# ---------------------------------------------------------------------------
# Condensed tree: rows (parent cluster, child point or cluster, lambda, size),
# clusters numbered from n (the root), lambda = 1 / distance.


def condense_tree(Z, min_cluster_size):
    """Condensed tree of a scipy-format single-linkage tree (C++ on the host)."""
    Z = np.ascontiguousarray(Z, dtype=np.float64)
    n = len(Z) + 1
    capacity = n + 2 * (n // max(int(min_cluster_size), 1)) + 2
    parent, child = np.empty(capacity, np.int64), np.empty(capacity, np.int64)
    lam, size = np.empty(capacity, np.float64), np.empty(capacity, np.int64)
    rows = _lib.lib.cuhc_condense_tree(
        _lib.double_ptr(Z), n, int(min_cluster_size), capacity, _lib.int64_ptr(parent),
        _lib.int64_ptr(child), _lib.double_ptr(lam), _lib.int64_ptr(size))
    if rows < 0:
        raise RuntimeError(f"condensed tree needs {-rows} rows, only {capacity} allocated")
    return (parent[:rows].astype(np.intp), child[:rows].astype(np.intp), lam[:rows],
            size[:rows].astype(np.intp))


def condense_tree_numpy(Z, min_cluster_size):
    """The same condensed tree computed level by level with NumPy (a cross-check of
    the C++ version; same rows in the same order)."""
    Z = np.asarray(Z, dtype=np.float64)
    n = len(Z) + 1
    root = 2 * (n - 1)
    kids = Z[:, :2].astype(np.intp)
    height = Z[:, 2]
    count = np.ones(2 * n - 1, dtype=np.intp)
    count[n:] = Z[:, 3].astype(np.intp)

    def subtree_bfs(start):
        out, level = [], np.array([start], dtype=np.intp)
        while len(level):
            out.append(level)
            inner = level[level >= n] - n
            level = kids[inner].ravel()
        return np.concatenate(out)

    order = subtree_bfs(root)
    label = np.zeros(root + 1, dtype=np.intp)
    label[root] = n
    nxt = n + 1
    ignore = np.zeros(root + 1, dtype=bool)
    rows = []
    for node in order.tolist():
        if node < n or ignore[node]:
            continue
        left, right = kids[node - n]
        h = height[node - n]
        lam = 1.0 / h if h > 0.0 else np.inf
        lc, rc = count[left], count[right]
        big_l, big_r = lc >= min_cluster_size, rc >= min_cluster_size
        if big_l and big_r:
            for side, c in ((left, lc), (right, rc)):
                label[side] = nxt
                nxt += 1
                rows.append((np.array([label[node]]), np.array([label[side]]),
                             np.array([lam]), np.array([c])))
            continue
        drops = []
        if big_l and not big_r:
            label[left] = label[node]
            drops = [right]
        elif big_r and not big_l:
            label[right] = label[node]
            drops = [left]
        else:
            drops = [left, right]
        for side in drops:
            sub = subtree_bfs(side)
            ignore[sub] = True
            pts = sub[sub < n]
            rows.append((np.full(len(pts), label[node]), pts, np.full(len(pts), lam),
                         np.ones(len(pts), dtype=np.intp)))
    if not rows:
        empty = np.empty(0, dtype=np.intp)
        return empty, empty, np.empty(0), empty
    cols = [np.concatenate([r[i] for r in rows]) for i in range(4)]
    return (cols[0].astype(np.intp), cols[1].astype(np.intp), cols[2].astype(np.float64),
            cols[3].astype(np.intp))


def compute_stability(condensed):
    """{cluster: sum over its rows of (lambda - lambda_birth) * size}, summed in row
    order (the same floating-point order as sklearn)."""
    parents, children, lambdas, sizes = condensed
    smallest = int(parents.min())
    largest = max(int(children.max()), smallest)
    birth = np.full(largest + 1, np.nan)
    birth[children] = lambdas
    birth[smallest] = 0.0
    totals = np.zeros(int(parents.max()) - smallest + 1)
    np.add.at(totals, parents - smallest, (lambdas - birth[parents]) * sizes)
    return {smallest + i: float(v) for i, v in enumerate(totals)}


def _death_lambdas(condensed):
    """Per parent: the largest lambda of the last run of consecutive rows with that
    parent (sklearn's max_lambdas, quirk included)."""
    parents, _, lambdas, _ = condensed
    deaths = np.zeros(int(parents.max()) + 1)
    starts = np.concatenate([[0], np.flatnonzero(parents[1:] != parents[:-1]) + 1])
    run_max = np.maximum.reduceat(lambdas, starts)
    run_parent = parents[starts]
    # Later runs overwrite earlier ones: keep the last run of each parent.
    _, last = np.unique(run_parent[::-1], return_index=True)
    keep = len(starts) - 1 - last
    deaths[run_parent[keep]] = run_max[keep]
    return deaths


class _ClusterTree:
    """The cluster rows (size > 1) of a condensed tree with children lists."""

    def __init__(self, condensed):
        parents, children, lambdas, sizes = condensed
        mask = sizes > 1
        self.parents = parents[mask]
        self.children = children[mask]
        self.lambdas = lambdas[mask]
        self.sizes = sizes[mask]
        self.kids = {}
        for p, c in zip(self.parents.tolist(), self.children.tolist()):
            self.kids.setdefault(p, []).append(c)
        self.parent_of = dict(zip(self.children.tolist(), self.parents.tolist()))
        self.lambda_of = dict(zip(self.children.tolist(), self.lambdas.tolist()))
        self.root = int(self.parents.min()) if len(self.parents) else None

    def bfs(self, start):
        out, level = [], [start]
        while level:
            out.extend(level)
            level = [c for node in level for c in self.kids.get(node, ())]
        return out

    def leaves(self):
        if self.root is None:
            return []
        out, stack = [], [self.root]
        while stack:  # depth first, children in row order (sklearn's recursion)
            node = stack.pop()
            kids = self.kids.get(node)
            if not kids:
                out.append(node)
            else:
                stack.extend(reversed(kids))
        return out

    def climb(self, leaf, epsilon, allow_single_cluster):
        """sklearn's traverse_upwards."""
        while True:
            parent = self.parent_of[leaf]
            if parent == self.root:
                return parent if allow_single_cluster else leaf
            if 1.0 / self.lambda_of[parent] > epsilon:
                return parent
            leaf = parent

    def epsilon_search(self, leaves, epsilon, allow_single_cluster):
        selected, processed = [], set()
        for leaf in leaves:
            if 1.0 / self.lambda_of[leaf] < epsilon:
                if leaf not in processed:
                    top = self.climb(leaf, epsilon, allow_single_cluster)
                    selected.append(top)
                    processed.update(c for c in self.bfs(top) if c != top)
            else:
                selected.append(leaf)
        return set(selected)


def select_clusters(condensed, stability, cluster_selection_method="eom",
                    allow_single_cluster=False, cluster_selection_epsilon=0.0,
                    max_cluster_size=None):
    """The selected cluster ids (sklearn's _get_clusters without the labelling)."""
    parents, children, _, sizes = condensed
    stability = dict(stability)
    nodes = sorted(stability, reverse=True)  # children have larger ids: bottom-up
    if not allow_single_cluster:
        nodes = nodes[:-1]
    tree = _ClusterTree(condensed)
    is_cluster = {c: True for c in nodes}
    n_samples = int(children[sizes == 1].max()) + 1
    if max_cluster_size is None:
        max_cluster_size = n_samples + 1
    cluster_sizes = dict(zip(tree.children.tolist(), tree.sizes.tolist()))
    if allow_single_cluster:
        cluster_sizes[nodes[-1]] = int(tree.sizes[tree.parents == nodes[-1]].sum())

    if cluster_selection_method == "eom":
        for node in nodes:
            kids = tree.kids.get(node, [])
            below = float(np.sum([stability[c] for c in kids]))
            if below > stability[node] or cluster_sizes[node] > max_cluster_size:
                is_cluster[node] = False
                stability[node] = below
            else:
                for sub in tree.bfs(node):
                    if sub != node:
                        is_cluster[sub] = False
        if cluster_selection_epsilon != 0.0 and len(tree.children) > 0:
            eom = [c for c in is_cluster if is_cluster[c]]
            chosen = []
            if len(eom) == 1 and eom[0] == tree.root:
                if allow_single_cluster:
                    chosen = eom
            else:
                chosen = tree.epsilon_search(set(eom), cluster_selection_epsilon,
                                             allow_single_cluster)
            for c in is_cluster:
                is_cluster[c] = c in chosen
    elif cluster_selection_method == "leaf":
        leaves = set(tree.leaves())
        if not leaves:
            for c in is_cluster:
                is_cluster[c] = False
            is_cluster[int(parents.min())] = True
        if cluster_selection_epsilon != 0.0:
            chosen = tree.epsilon_search(leaves, cluster_selection_epsilon, allow_single_cluster)
        else:
            chosen = leaves
        for c in is_cluster:
            is_cluster[c] = c in chosen
    return {c for c in is_cluster if is_cluster[c]}


def label_points(condensed, clusters, allow_single_cluster, cluster_selection_epsilon):
    """Label of every point: its nearest selected ancestor cluster (labels in sorted
    cluster order), else noise. Equivalent to sklearn's union-find labelling."""
    parents, children, lambdas, sizes = condensed
    root = int(parents.min())
    n = root
    label_of = {c: i for i, c in enumerate(sorted(clusters))}
    owner = {root: root}
    for p, c in zip(parents[sizes > 1].tolist(), children[sizes > 1].tolist()):
        owner[c] = c if c in clusters else owner[p]
    labels = np.full(n, NOISE, dtype=np.intp)
    point_rows = np.flatnonzero(children < n)
    point_parents = parents[point_rows]
    uniq, inv = np.unique(point_parents, return_inverse=True)
    owner_of_row = np.array([owner[p] for p in uniq.tolist()], dtype=np.intp)[inv]
    inside = owner_of_row != root
    labels[children[point_rows[inside]]] = np.array(
        [label_of[c] for c in owner_of_row[inside].tolist()], dtype=np.intp)
    if len(clusters) == 1 and allow_single_cluster and root in clusters:
        if cluster_selection_epsilon != 0.0:
            threshold = 1.0 / cluster_selection_epsilon
        else:
            threshold = lambdas[parents == root].max()
        at_root = point_rows[~inside]
        keep = lambdas[at_root] >= threshold
        labels[children[at_root[keep]]] = label_of[root]
    return labels


def membership_probabilities(condensed, clusters, labels):
    """min(lambda_point, lambda_max) / lambda_max of the point's cluster (1 when
    lambda_max is 0 or the point's lambda is infinite), 0 for noise."""
    parents, children, lambdas, _ = condensed
    deaths = _death_lambdas(condensed)
    ordered = np.array(sorted(clusters), dtype=np.intp)
    out = np.zeros(len(labels))
    rows = np.flatnonzero(children < parents.min())
    pts = children[rows]
    labelled = labels[pts] != NOISE
    rows, pts = rows[labelled], pts[labelled]
    if len(rows) == 0:
        return out
    top = deaths[ordered[labels[pts]]]
    lam = lambdas[rows]
    certain = (top == 0.0) | np.isinf(lam)
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = np.minimum(lam, top) / top
    out[pts] = np.where(certain, 1.0, ratio)
    return out


def tree_to_labels(Z, min_cluster_size, cluster_selection_method="eom",
                   allow_single_cluster=False, cluster_selection_epsilon=0.0,
                   max_cluster_size=None):
    """(labels, probabilities) from a single-linkage tree, as sklearn's tree_to_labels."""
    condensed = condense_tree(Z, min_cluster_size)
    clusters = select_clusters(condensed, compute_stability(condensed), cluster_selection_method,
                               allow_single_cluster, cluster_selection_epsilon, max_cluster_size)
    labels = label_points(condensed, clusters, allow_single_cluster, cluster_selection_epsilon)
    return labels, membership_probabilities(condensed, clusters, labels)


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

    def _validate_params(self):
        def integer_at_least(value, lowest):
            return isinstance(value, numbers.Integral) and value >= lowest

        if not integer_at_least(self.min_cluster_size, 2):
            raise ValueError(
                f"min_cluster_size must be an integer >= 2, got {self.min_cluster_size!r}")
        if self.min_samples is not None and not integer_at_least(self.min_samples, 1):
            raise ValueError(
                f"min_samples must be None or an integer >= 1, got {self.min_samples!r}")
        if self.max_cluster_size is not None and not integer_at_least(self.max_cluster_size, 1):
            raise ValueError(
                f"max_cluster_size must be None or an integer >= 1, got {self.max_cluster_size!r}")
        if not self.cluster_selection_epsilon >= 0:
            raise ValueError("cluster_selection_epsilon must be >= 0, got "
                             f"{self.cluster_selection_epsilon!r}")
        if not self.alpha > 0:
            raise ValueError(f"alpha must be > 0, got {self.alpha!r}")
        if self.cluster_selection_method not in ("eom", "leaf"):
            raise ValueError("cluster_selection_method must be 'eom' or 'leaf', got "
                             f"{self.cluster_selection_method!r}")
        validate_metric(self.metric, 2.0)

    def fit(self, X, y=None):
        self._validate_params()
        X = as_points(X, "X")
        n = X.shape[0]
        k = self.min_cluster_size if self.min_samples is None else self.min_samples
        if k > n:
            raise ValueError(f"min_samples ({k}) must be at most the number of samples in X ({n})")
        Z, core = hdbscan_tree(X, k, metric=self.metric, alpha=self.alpha)
        self.labels_, self.probabilities_ = tree_to_labels(
            Z, self.min_cluster_size, self.cluster_selection_method, self.allow_single_cluster,
            self.cluster_selection_epsilon, self.max_cluster_size)
        self._single_linkage_tree_ = Z
        self.core_distances_ = core
        self.n_features_in_ = X.shape[1]
        return self
