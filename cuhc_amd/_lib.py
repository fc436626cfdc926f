"""ctypes binding of libcuhc.so (the gfx950 HIP library built by cuhc_amd.build).

Mirrors the reference binding (reference python/cuhc/_lib.py): the library is
loaded from this package's directory and importing fails loudly when it is not
built. There is no CPU fallback.
"""

import ctypes
import pathlib

_LIB_PATH = pathlib.Path(__file__).resolve().parent / "libcuhc.so"

if not _LIB_PATH.exists():
    raise ImportError(
        f"{_LIB_PATH} not found; build it with `python cuhc_amd/build.py` "
        "(hipcc, gfx950) before importing cuhc_amd"
    )

lib = ctypes.CDLL(str(_LIB_PATH))

_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u64p = ctypes.POINTER(ctypes.c_ulonglong)
_int, _float, _i64 = ctypes.c_int, ctypes.c_float, ctypes.c_int64


def _declare(name, restype, *argtypes):
    fn = getattr(lib, name)
    fn.restype = restype
    fn.argtypes = list(argtypes)
    return fn


# The eight reference entry points (same signatures).
_declare("cuhc_pairwise_distances", _int, _f32p, _int, _f32p, _int, _int, _int, _float, _f32p)
_declare("cuhc_linkage", _int, _f32p, _int, _int, _int, _float, _int, _int, _f64p)
_declare("cuhc_ward_rnn_with_options", _int, _f32p, _int, _int, _float, _int, _int, _f64p)
_declare("cuhc_hdbscan_tree", _int, _f32p, _int, _int, _int, _float, _int, _float, _int, _f64p,
         _f32p)
_declare("cuhc_condense_tree", _i64, _f64p, _i64, _i64, _i64, _i64p, _i64p, _f64p, _i64p)
_declare("cuhc_set_pruning", _int, _int)
_declare("cuhc_prune_stats", None, _u64p, _u64p)
_declare("cuhc_error_string", ctypes.c_char_p, _int)

# Extensions (no counterpart in the reference library).
_u8p = ctypes.POINTER(ctypes.c_uint8)
_CONDENSED = (_i64p, _i64p, _f64p, _i64p, _i64)
_declare("cuhc_cluster_stability", _int, *_CONDENSED, _f64p)
_declare("cuhc_select_clusters", _int, *_CONDENSED, _f64p, _int, _int, _i64, _u8p)
_declare("cuhc_label_points", _int, *_CONDENSED, _u8p, _int, ctypes.c_double, _i64p)
_declare("cuhc_membership_probabilities", _int, *_CONDENSED, _u8p, _i64p, _f64p)
_declare("cuhc_set_precision", _int, _int)
_declare("cuhc_linkage_device", _int, ctypes.c_void_p, _int, _int, _int, _float, _int, _int,
         _f64p)
_declare("cuhc_linkage_device_noalloc", _int, ctypes.c_void_p, _int, _int, _int, _float, _int,
         _int, _f64p)
_declare("cuhc_exclusive_scan", _int, ctypes.c_void_p, ctypes.c_void_p, _int)
_declare("cuhc_radix_sort_pairs_u32", _int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
         ctypes.c_void_p, _int, _int)
_declare("cuhc_radix_sort_pairs_u64", _int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
         ctypes.c_void_p, _int, _int)
_declare("cuhc_pairwise_distances_device", _int, ctypes.c_void_p, _int, ctypes.c_void_p, _int,
         _int, _int, _float, ctypes.c_void_p)
_declare("cuhc_feature_ranges_device", _int, ctypes.c_void_p, _int, _int, _f32p, _f32p)
_declare("cuhc_hdbscan_tree_device", _int, ctypes.c_void_p, _int, _int, _int, _float, _int,
         _float, _int, _f64p, _f32p)
_declare("cuhc_device_alloc", _int, ctypes.c_ulonglong, ctypes.POINTER(ctypes.c_void_p))
_declare("cuhc_device_free", _int, ctypes.c_void_p)
_declare("cuhc_copy_to_device", _int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_ulonglong)
_declare("cuhc_copy_device", _int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_ulonglong)
_declare("cuhc_copy_to_host", _int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_ulonglong)
_declare("cuhc_mfma_probe", _int, _int, _int, _int, ctypes.POINTER(ctypes.c_float),
         ctypes.POINTER(ctypes.c_double))
_declare("cuhc_set_timing", _int, _int)
_declare("cuhc_read_timings", None, _f64p, _i64p, _i64p, _int)
_declare("cuhc_reset_timings", None)
_declare("cuhc_release", None)
_declare("cuhc_mixed_stats", None, _u64p, _u64p, _u64p, _u64p)
_declare("cuhc_learned_stats", None, _u64p, _u64p, _u64p, _u64p, _u64p, _u64p, _u64p)
_declare("cuhc_set_learned_training", _int, _int)
_declare("cuhc_mlp_param_count", _int)
_declare("cuhc_mlp_gradient", _int, _f32p, _f32p, _f32p, _int, _f32p, _f32p)
_declare("cuhc_mlp_forward", _int, _f32p, _f32p, _int, _f32p)
_declare("cuhc_mlp_train", _int, _f32p, _f32p, _f32p, _int, _int, _float)

PRUNE_MODES = {"auto": 0, "always": 1, "never": 2}
PRECISION_MODES = {"auto": 0, "fp32": 1, "mixed": 2}
PHASES = ["distance_tiles", "mfma_candidates", "recheck", "core", "contract", "dense", "rnn",
          "mlp", "dense_select", "dense_update", "dense_rescan", "dense_barrier", "host_tree"]

# The library's status for NaN or infinite input (reference src/mst.cuh:80-81).
NON_FINITE = -6
# hipErrorOutOfMemory, the status the dense path returns when its matrix does not fit.
HIP_ERROR_OUT_OF_MEMORY = 2


def check(status):
    """Raises RuntimeError for a non-zero status of a C ABI call."""
    if status != 0:
        raise RuntimeError(f"cuhc: {lib.cuhc_error_string(status).decode()} (code {status})")


def set_pruning(mode):
    """Spatial pruning for mst, rnn and HDBSCAN: "auto", "always" or "never" (same results)."""
    check(lib.cuhc_set_pruning(PRUNE_MODES[mode]))


def prune_stats():
    """(tiles computed, tiles in total) of the pruned passes since the last call."""
    computed, total = ctypes.c_ulonglong(), ctypes.c_ulonglong()
    lib.cuhc_prune_stats(ctypes.byref(computed), ctypes.byref(total))
    return computed.value, total.value


def set_precision(mode):
    """Boruvka pass precision: "auto", "fp32" (plain tiles) or "mixed" (same trees)."""
    check(lib.cuhc_set_precision(PRECISION_MODES[mode]))


def mixed_stats():
    """Counters of the mixed-precision passes since the last call (then reset)."""
    vals = [ctypes.c_ulonglong() for _ in range(4)]
    lib.cuhc_mixed_stats(*[ctypes.byref(v) for v in vals])
    return dict(zip(["recorded", "rechecked_tiles", "rechecked_pairs", "overflow_rows"],
                    [v.value for v in vals]))


def learned_stats():
    """Counters of the learned-radius core distances since the last call (then reset)."""
    vals = [ctypes.c_ulonglong() for _ in range(7)]
    lib.cuhc_learned_stats(*[ctypes.byref(v) for v in vals])
    return dict(zip(["proposals", "hits", "fallbacks", "candidates", "tiles", "settled_rows",
                     "offered_rows"], [v.value for v in vals]))


def set_timing(on):
    lib.cuhc_set_timing(1 if on else 0)


def reset_timings():
    lib.cuhc_reset_timings()


def read_timings():
    """{phase: (total ms, launches, algorithmic units)} since the last reset."""
    n = len(PHASES)
    ms = (ctypes.c_double * n)()
    launches = (ctypes.c_int64 * n)()
    units = (ctypes.c_int64 * n)()
    lib.cuhc_read_timings(ms, launches, units, n)
    return {name: (ms[i], launches[i], units[i]) for i, name in enumerate(PHASES)}


class DeviceArray:
    """A C-contiguous array in HBM, allocated with the library's allocation
    entry point: a copy of a host array, or uninitialised (DeviceArray.empty).
    .ptr is the device address the *_device entry points take; the array
    exports the CUDA array interface (version 3), so cuhc_amd's functions
    read it in place, and .to_host() copies it back."""

    def __init__(self, host=None, *, shape=None, dtype=None):
        import numpy as np
        if host is not None:
            self.host = np.ascontiguousarray(host)
            self.shape, self.dtype = self.host.shape, self.host.dtype
            nbytes = self.host.nbytes
        else:
            self.host = None
            self.shape, self.dtype = tuple(shape), np.dtype(dtype)
            nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.nbytes = nbytes
        self.ptr = ctypes.c_void_p()
        check(lib.cuhc_device_alloc(max(nbytes, 1), ctypes.byref(self.ptr)))
        if host is not None:
            check(lib.cuhc_copy_to_device(self.ptr, self.host.ctypes.data_as(ctypes.c_void_p),
                                          nbytes))

    @classmethod
    def empty(cls, shape, dtype):
        return cls(shape=shape, dtype=dtype)

    @property
    def __cuda_array_interface__(self):
        return {"shape": tuple(self.shape), "typestr": self.dtype.str, "strides": None,
                "data": (int(self.ptr.value or 0), False), "version": 3}

    def to_host(self):
        import numpy as np
        out = np.empty(self.shape, self.dtype)
        check(lib.cuhc_copy_to_host(out.ctypes.data_as(ctypes.c_void_p), self.ptr, self.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib.cuhc_device_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        self.free()


def float_ptr(array):
    return array.ctypes.data_as(_f32p)


def double_ptr(array):
    return array.ctypes.data_as(_f64p)


def int64_ptr(array):
    return array.ctypes.data_as(_i64p)


def set_learned_training(steps):
    """Adam steps that train the radius network per learned-core call (0 = untrained)."""
    check(lib.cuhc_set_learned_training(int(steps)))


MLP_PARAMS = lib.cuhc_mlp_param_count()


def _mlp_inputs(params, X):
    import numpy as np
    params = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32)
    if params.shape != (MLP_PARAMS,) or X.ndim != 2 or X.shape[1] != 4:
        raise ValueError(f"params must have {MLP_PARAMS} entries and X shape (m, 4)")
    return params, X


def mlp_forward(params, X):
    """Outputs of the radius network for the rows of X (m x 4), computed on the GPU."""
    import numpy as np
    params, X = _mlp_inputs(params, X)
    y = np.empty(len(X), dtype=np.float32)
    check(lib.cuhc_mlp_forward(float_ptr(params), float_ptr(X), len(X), float_ptr(y)))
    return y


def mlp_gradient(params, X, t):
    """(gradient, loss) of the mean 0.5 (y - t)^2, from the GPU backward kernels."""
    import numpy as np
    params, X = _mlp_inputs(params, X)
    t = np.ascontiguousarray(t, dtype=np.float32)
    grad = np.empty(MLP_PARAMS, dtype=np.float32)
    loss = ctypes.c_float()
    check(lib.cuhc_mlp_gradient(float_ptr(params), float_ptr(X), float_ptr(t), len(X),
                                float_ptr(grad), ctypes.byref(loss)))
    return grad, loss.value


def mlp_train(params, X, t, steps, lr=0.01):
    """Parameters after `steps` full-batch Adam updates on the GPU (input unchanged)."""
    params, X = _mlp_inputs(params, X)
    params = params.copy()
    import numpy as np
    t = np.ascontiguousarray(t, dtype=np.float32)
    check(lib.cuhc_mlp_train(float_ptr(params), float_ptr(X), float_ptr(t), len(X), int(steps),
                             float(lr)))
    return params
