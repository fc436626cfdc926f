"""Pairwise distances on the GPU and the input checks every entry point shares.

Same surface and messages as reference python/cuhc/_distances.py.
"""

import numpy as np

from . import _lib

# Metric names (sklearn's, with aliases) -> the C ABI metric id.
METRIC_IDS = {
    "euclidean": 0, "l2": 0,
    "sqeuclidean": 1,
    "manhattan": 2, "cityblock": 2, "l1": 2,
    "chebyshev": 3,
    "minkowski": 4,
    "cosine": 5,
}

FLOAT32_MAX = float(np.finfo(np.float32).max)


def validate_metric(metric, p):
    if metric not in METRIC_IDS:
        raise ValueError(f"unknown metric {metric!r}; expected one of {sorted(METRIC_IDS)}")
    if metric == "minkowski" and not p > 0:
        raise ValueError(f"Minkowski p must be positive, got {p}")


def as_points(array, name):
    """C-contiguous float32 (points x features), finite and non-empty."""
    out = np.ascontiguousarray(array, dtype=np.float32)
    if out.ndim != 2:
        raise ValueError(f"{name} must be 2D (points x features), got shape {out.shape}")
    if 0 in out.shape:
        raise ValueError(f"{name} must be non-empty, got shape {out.shape}")
    if not np.isfinite(out).all():
        raise ValueError(f"{name} contains NaN or infinity")
    return out


def _distance_bound(spread, magnitude, metric, p):
    kind = METRIC_IDS[metric]
    if kind in (0, 1):
        return float(np.sum(spread * spread))     # squared distances are accumulated
    if kind == 2:
        return float(np.sum(spread))
    if kind == 3:
        return float(np.max(spread))
    if kind == 4:
        return float(np.sum(spread ** p))         # sum of |x_k - y_k|^p before the root
    return float(np.sum(magnitude * magnitude))   # cosine: dot products and squared norms


def check_distance_range(X, metric, p=2.0, Y=None, scale=1.0):
    """Rejects inputs whose FP32 distances (times `scale`) could overflow.

    |x_k - y_k| never exceeds feature k's spread, so the bound is exact per
    feature; `scale` covers factors applied after the metric (Ward's cluster
    sizes, HDBSCAN's 1 / alpha).
    """
    values = X if Y is None else np.vstack([X, Y])
    with np.errstate(over="ignore", invalid="ignore"):
        hi = values.max(axis=0).astype(np.float64)
        lo = values.min(axis=0).astype(np.float64)
        spread = hi - lo
        magnitude = np.maximum(np.abs(hi), np.abs(lo))
        bound = _distance_bound(spread, magnitude, metric, p) * scale
    if not bound <= FLOAT32_MAX:
        raise ValueError(
            f"the data's range is too large for float32 {metric} distances (they could reach "
            f"{bound:.3g}, above float32's {FLOAT32_MAX:.3g}); rescale the data"
        )


def pairwise_distances(X, Y=None, metric="euclidean", p=2.0):
    """Distances between the rows of X and the rows of Y (Y = X if omitted).

    Same as sklearn.metrics.pairwise_distances; computed in float32 on the GPU.

    Returns an (n_samples_X, n_samples_Y) float32 array.

    >>> import numpy as np, cuhc_amd as cuhc
    >>> cuhc.pairwise_distances(np.array([[0.0, 0.0], [3.0, 4.0]]))
    array([[0., 5.],
           [5., 0.]], dtype=float32)
    """
    validate_metric(metric, p)
    X = as_points(X, "X")
    if Y is not None:
        Y = as_points(Y, "Y")
        if Y.shape[1] != X.shape[1]:
            raise ValueError(
                f"X and Y must have the same number of features, got {X.shape[1]} and {Y.shape[1]}"
            )
    check_distance_range(X, metric, p, Y)
    n_b = X.shape[0] if Y is None else Y.shape[0]
    out = np.empty((X.shape[0], n_b), dtype=np.float32)
    _lib.check(_lib.lib.cuhc_pairwise_distances(
        _lib.float_ptr(X), X.shape[0], None if Y is None else _lib.float_ptr(Y), n_b,
        X.shape[1], METRIC_IDS[metric], float(p), _lib.float_ptr(out)))
    return out
