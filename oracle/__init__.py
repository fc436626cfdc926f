"""Python wrapper of the CPU oracle (oracle.c). Test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
package, as the checker; the cuhc_amd product never imports it.
"""

import ctypes
import pathlib

import numpy as np

_PATH = pathlib.Path(__file__).resolve().parent / "_build" / "liboracle.so"
_lib = None

METRICS = {"euclidean": 0, "sqeuclidean": 1, "manhattan": 2, "chebyshev": 3, "minkowski": 4,
           "cosine": 5}
LINKAGES = {"single": 0, "complete": 1, "average": 2, "ward": 3}


def lib():
    global _lib
    if _lib is None:
        if not _PATH.exists():
            raise ImportError(f"{_PATH} not built; run `python -m oracle.build`")
        L = ctypes.CDLL(str(_PATH))
        f32p, f64p = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double)
        i32p = ctypes.POINTER(ctypes.c_int)
        c_int, c_float = ctypes.c_int, ctypes.c_float
        sigs = {
            "oracle_pairwise": [f32p, c_int, f32p, c_int, c_int, c_int, c_float, f32p],
            "oracle_mst": [f32p, c_int, c_int, c_int, c_float, f32p, c_float, i32p, i32p, f32p],
            "oracle_single_linkage": [f32p, c_int, c_int, c_int, c_float, f64p],
            "oracle_core_distances": [f32p, c_int, c_int, c_int, c_float, c_int, f32p],
            "oracle_hdbscan_tree": [f32p, c_int, c_int, c_int, c_float, c_int, c_float, f64p,
                                    f32p],
            "oracle_dense_on_matrix": [f32p, c_int, c_int, f64p],
            "oracle_dense_linkage": [f32p, c_int, c_int, c_int, c_float, c_int, f64p],
            "oracle_mlp_forward": [f32p, f32p, c_int, f64p],
            "oracle_mlp_gradient": [f32p, f32p, f32p, c_int, f64p, f64p],
        }
        for name, args in sigs.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = c_int
        _lib = L
    return _lib


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _check(status):
    if status != 0:
        raise RuntimeError(f"oracle status {status}")


def pairwise(X, Y=None, metric="euclidean", p=2.0):
    X, xp = _f32(X)
    Y, yp = (X, xp) if Y is None else _f32(Y)
    out = np.empty((len(X), len(Y)), dtype=np.float32)
    _check(lib().oracle_pairwise(xp, len(X), yp, len(Y), X.shape[1], METRICS[metric], p,
                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
    return out


def mst(X, metric="euclidean", p=2.0, core=None, alpha=1.0):
    """(u, v, w) of the MST under the order (w, u, v); u < v."""
    X, xp = _f32(X)
    n = len(X)
    eu, ev = np.zeros(max(n - 1, 0), np.int32), np.zeros(max(n - 1, 0), np.int32)
    ew = np.zeros(max(n - 1, 0), np.float32)
    cp = None
    if core is not None:
        core, cp = _f32(core)
    _check(lib().oracle_mst(xp, n, X.shape[1], METRICS[metric], p, cp, alpha,
                            eu.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                            ev.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                            ew.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
    return eu, ev, ew


def single_linkage(X, metric="euclidean", p=2.0):
    X, xp = _f32(X)
    tree = np.empty((max(len(X) - 1, 0), 4), dtype=np.float64)
    _check(lib().oracle_single_linkage(xp, len(X), X.shape[1], METRICS[metric], p,
                                       tree.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return tree


def core_distances(X, k, metric="euclidean", p=2.0):
    X, xp = _f32(X)
    core = np.empty(len(X), dtype=np.float32)
    _check(lib().oracle_core_distances(xp, len(X), X.shape[1], METRICS[metric], p, k,
                                       core.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
    return core


def hdbscan_tree(X, k, metric="euclidean", alpha=1.0, p=2.0):
    X, xp = _f32(X)
    n = len(X)
    tree = np.empty((max(n - 1, 0), 4), dtype=np.float64)
    core = np.empty(n, dtype=np.float32)
    _check(lib().oracle_hdbscan_tree(xp, n, X.shape[1], METRICS[metric], p, k, alpha,
                                     tree.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                     core.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
    return tree, core


def dense_linkage(X, method, metric="euclidean", p=2.0):
    X, xp = _f32(X)
    tree = np.empty((max(len(X) - 1, 0), 4), dtype=np.float64)
    _check(lib().oracle_dense_linkage(xp, len(X), X.shape[1], METRICS[metric], p,
                                      LINKAGES[method],
                                      tree.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return tree


def dense_on_matrix(D, method):
    D, dp = _f32(np.array(D, dtype=np.float32, copy=True))
    tree = np.empty((max(len(D) - 1, 0), 4), dtype=np.float64)
    _check(lib().oracle_dense_on_matrix(dp, len(D), LINKAGES[method],
                                        tree.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return tree


MLP_PARAMS = 32 * 4 + 32 + 32 * 32 + 32 + 32 + 1


def mlp_forward(params, X):
    """FP64 outputs of the radius network (layout of cuhc_mlp_*)."""
    params, pp = _f32(params)
    X, xp = _f32(X)
    y = np.empty(len(X), dtype=np.float64)
    _check(lib().oracle_mlp_forward(pp, xp, len(X),
                                    y.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return y


def mlp_gradient(params, X, t):
    """(FP64 gradient of the mean 0.5 (y - t)^2, loss)."""
    params, pp = _f32(params)
    X, xp = _f32(X)
    t, tp = _f32(t)
    g = np.empty(MLP_PARAMS, dtype=np.float64)
    loss = ctypes.c_double()
    _check(lib().oracle_mlp_gradient(pp, xp, tp, len(X),
                                     g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                     ctypes.byref(loss)))
    return g, loss.value
