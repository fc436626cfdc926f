"""Pairwise distances on the GPU and the input checks every entry point shares.

The names, metric ids, argument rules and error messages are the reference's
interface (python/cuhc/_distances.py); the checks are written here.
"""

import numpy as np

from . import _lib
from ._device import DeviceMatrix, device_matrix, feature_ranges

# Metric names (scikit-learn's, with aliases) -> the C ABI metric id.
METRIC_IDS = {
    "euclidean": 0, "l2": 0,
    "sqeuclidean": 1,
    "manhattan": 2, "cityblock": 2, "l1": 2,
    "chebyshev": 3,
    "minkowski": 4,
    "cosine": 5,
}

FLOAT32_MAX = float(np.finfo(np.float32).max)

# Largest value a metric accumulates in FP32 before any final root, from the
# per-feature spread s_k = max_k - min_k and magnitude a_k = max |x_k|:
# |x_k - y_k| <= s_k, and cosine accumulates dot products and squared norms.
_ACCUMULATED_BOUND = {
    0: lambda s, a, p: (s * s).sum(),
    1: lambda s, a, p: (s * s).sum(),
    2: lambda s, a, p: s.sum(),
    3: lambda s, a, p: s.max(),
    4: lambda s, a, p: (s ** p).sum(),
    5: lambda s, a, p: (a * a).sum(),
}


def validate_metric(metric, p):
    if metric not in METRIC_IDS:
        raise ValueError(f"unknown metric {metric!r}; expected one of {sorted(METRIC_IDS)}")
    if metric == "minkowski" and not p > 0:
        raise ValueError(f"Minkowski p must be positive, got {p}")


def as_points(array, name):
    """The rows of `array` as a C-contiguous float32 (points x features) array,
    checked to be 2-D, non-empty and finite."""
    points = np.require(np.asarray(array, dtype=np.float32), requirements="C")
    problem = None
    if points.ndim != 2:
        problem = f"must be 2D (points x features), got shape {points.shape}"
    elif min(points.shape) == 0:
        problem = f"must be non-empty, got shape {points.shape}"
    elif not np.isfinite(points).all():
        problem = "contains NaN or infinity"
    if problem:
        raise ValueError(f"{name} {problem}")
    return points


def input_points(array, name):
    """(host points, None) for host data (as_points), or (None, DeviceMatrix)
    for a float32 matrix already in GPU memory (CUDA array interface or a ROCm
    DLPack tensor), which is read in place."""
    dm = device_matrix(array, name)
    if dm is not None:
        return None, dm
    return as_points(array, name), None


def ranges_of(X):
    """Per-feature (min, max) of host points or of a DeviceMatrix (on the GPU)."""
    if isinstance(X, DeviceMatrix):
        return feature_ranges(_lib.lib, X)
    return X.min(axis=0), X.max(axis=0)


def check_distance_range(X, metric, p=2.0, Y=None, scale=1.0):
    """Raises ValueError when some FP32 distance (times `scale`: Ward's cluster
    sizes, HDBSCAN's 1 / alpha) could overflow float32. The bound is exact per
    feature, so data far from the origin passes as long as its spread fits.
    X and Y: host points or DeviceMatrix views."""
    lo, hi = ranges_of(X)
    if Y is not None:
        ylo, yhi = ranges_of(Y)
        lo, hi = np.minimum(lo, ylo), np.maximum(hi, yhi)
    if not (np.isfinite(lo).all() and np.isfinite(hi).all()):
        raise ValueError("X contains NaN or infinity")
    lo, hi = lo.astype(np.float64), hi.astype(np.float64)
    with np.errstate(over="ignore", invalid="ignore"):
        bound = float(_ACCUMULATED_BOUND[METRIC_IDS[metric]](
            hi - lo, np.maximum(np.abs(lo), np.abs(hi)), p)) * scale
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
    A, dA = input_points(X, "X")
    if dA is not None:
        return _pairwise_on_device(dA, Y, metric, p)
    B = None if Y is None else as_points(Y, "Y")
    if B is not None and B.shape[1] != A.shape[1]:
        raise ValueError(
            f"X and Y must have the same number of features, got {A.shape[1]} and {B.shape[1]}"
        )
    check_distance_range(A, metric, p, B)
    cols = A if B is None else B
    out = np.empty((len(A), len(cols)), dtype=np.float32)
    _lib.check(_lib.lib.cuhc_pairwise_distances(
        _lib.float_ptr(A), len(A), None if B is None else _lib.float_ptr(B), len(cols),
        A.shape[1], METRIC_IDS[metric], float(p), _lib.float_ptr(out)))
    return out


def _pairwise_on_device(dA, Y, metric, p):
    """pairwise_distances for X in GPU memory: Y must be in GPU memory too (or
    None), and the matrix stays there (a DeviceArray)."""
    with dA:
        dB = None
        if Y is not None:
            dB = device_matrix(Y, "Y")
            if dB is None:
                raise ValueError("Y must be in GPU memory when X is")
        try:
            if dB is not None and dB.shape[1] != dA.shape[1]:
                raise ValueError(f"X and Y must have the same number of features, got "
                                 f"{dA.shape[1]} and {dB.shape[1]}")
            check_distance_range(dA, metric, p, dB)
            cols = dA if dB is None else dB
            out = _lib.DeviceArray.empty((dA.shape[0], cols.shape[0]), np.float32)
            import ctypes
            _lib.check(_lib.lib.cuhc_pairwise_distances_device(
                ctypes.c_void_p(dA.ptr), dA.shape[0],
                None if dB is None else ctypes.c_void_p(dB.ptr), cols.shape[0], dA.shape[1],
                METRIC_IDS[metric], float(p), out.ptr))
            return out
        finally:
            if dB is not None:
                dB.close()
