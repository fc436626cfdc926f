"""scikit-learn compatible AgglomerativeClustering on the GPU.

Parameters, defaults, attributes and messages are those of reference
python/cuhc/_estimators.py (and sklearn.cluster.AgglomerativeClustering).
"""

import numbers

import numpy as np
from sklearn.base import BaseEstimator, ClusterMixin

from ._distances import as_points
from ._device import gpu_shape
from ._linkage import LINKAGE_IDS, cut_tree, linkage, resolve_algorithm, validate_linkage


def _argument_errors(est):
    """The first problem with the estimator's parameters, as a message, or None."""
    if (est.n_clusters is None) == (est.distance_threshold is None):
        return ("Exactly one of n_clusters and distance_threshold has to be set, "
                "and the other needs to be None.")
    k = est.n_clusters
    if k is not None and not (isinstance(k, numbers.Integral) and k >= 1):
        return f"n_clusters must be a positive integer, got {k!r}"
    if est.linkage not in LINKAGE_IDS:
        return f"linkage must be one of {list(LINKAGE_IDS)}, got {est.linkage!r}"
    return None


class AgglomerativeClustering(ClusterMixin, BaseEstimator):
    """Exact agglomerative clustering; attributes as in scikit-learn.

    Parameters: n_clusters (default 2), metric ("euclidean"), linkage ("ward"),
    distance_threshold (None), compute_distances (False), and algorithm
    ("auto"; as in cuhc_amd.linkage). Fitted: labels_, n_clusters_, children_,
    distances_ (if requested or thresholded), n_leaves_, n_features_in_,
    n_connected_components_.

    >>> import numpy as np, cuhc_amd as cuhc
    >>> X = np.array([[0.0, 0.0], [0.1, 0.0], [5.0, 5.0], [5.1, 5.0]])
    >>> cuhc.AgglomerativeClustering(n_clusters=2).fit_predict(X)
    array([1, 1, 0, 0])
    """

    def __init__(self, n_clusters=2, *, metric="euclidean", linkage="ward",
                 distance_threshold=None, compute_distances=False, algorithm="auto"):
        self.n_clusters = n_clusters
        self.metric = metric
        self.linkage = linkage
        self.distance_threshold = distance_threshold
        self.compute_distances = compute_distances
        self.algorithm = algorithm

    def fit(self, X, y=None):
        message = _argument_errors(self)
        if message:
            raise ValueError(message)
        validate_linkage(self.linkage, self.metric)
        resolve_algorithm(self.algorithm, self.linkage)
        shape = gpu_shape(X)
        points = X if shape is not None else as_points(X, "X")  # GPU data is read in place
        n = (shape or points.shape)[0]
        if self.n_clusters is not None and self.n_clusters > n:
            raise ValueError(f"Cannot extract more clusters ({self.n_clusters}) than samples ({n})")
        Z = linkage(points, method=self.linkage, metric=self.metric, algorithm=self.algorithm)
        heights = Z[:, 2]
        thresholded = self.distance_threshold is not None
        # scikit-learn's threshold rule: merges at or above it are not applied.
        k = int((heights >= self.distance_threshold).sum()) + 1 if thresholded \
            else int(self.n_clusters)
        self.n_clusters_ = k
        self.labels_ = cut_tree(Z, n - k)
        self.children_ = Z[:, :2].astype(np.intp)
        if thresholded or self.compute_distances:
            self.distances_ = heights.copy()
        self.n_leaves_ = n
        self.n_features_in_ = (shape or points.shape)[1]
        self.n_connected_components_ = 1
        return self
