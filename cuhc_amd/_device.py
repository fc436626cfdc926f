"""Inputs already in GPU memory: the CUDA array interface (version 3, as HIP
device pointers) and DLPack capsules of ROCm device tensors, read without a
copy to the host (the reference's planned zero-copy binding,
docs/architecture.md:273-279).

Only C-contiguous float32 matrices are taken as they are; anything else must be
converted by its owner first (no device-side conversion happens here).
"""

import ctypes

import numpy as np

# DLPack device types (dlpack.h): 2 CUDA, 10 ROCm.
DL_ROCM = 10
_DL_FLOAT = 2


class _DLDevice(ctypes.Structure):
    _fields_ = [("device_type", ctypes.c_int32), ("device_id", ctypes.c_int32)]


class _DLDataType(ctypes.Structure):
    _fields_ = [("code", ctypes.c_uint8), ("bits", ctypes.c_uint8), ("lanes", ctypes.c_uint16)]


class _DLTensor(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("device", _DLDevice), ("ndim", ctypes.c_int32),
                ("dtype", _DLDataType), ("shape", ctypes.POINTER(ctypes.c_int64)),
                ("strides", ctypes.POINTER(ctypes.c_int64)), ("byte_offset", ctypes.c_uint64)]


class _DLManagedTensor(ctypes.Structure):
    pass


_DELETER = ctypes.CFUNCTYPE(None, ctypes.POINTER(_DLManagedTensor))
_DLManagedTensor._fields_ = [("dl_tensor", _DLTensor), ("manager_ctx", ctypes.c_void_p),
                             ("deleter", _DELETER)]

_api = ctypes.pythonapi
_api.PyCapsule_GetPointer.restype = ctypes.c_void_p
_api.PyCapsule_GetPointer.argtypes = [ctypes.py_object, ctypes.c_char_p]
_api.PyCapsule_IsValid.restype = ctypes.c_int
_api.PyCapsule_IsValid.argtypes = [ctypes.py_object, ctypes.c_char_p]
_api.PyCapsule_SetName.restype = ctypes.c_int
_api.PyCapsule_SetName.argtypes = [ctypes.py_object, ctypes.c_char_p]


class DeviceMatrix:
    """A borrowed view of an n x d float32 matrix in GPU memory: .ptr (int
    device address) and .shape. Close it (or use it as a context manager)
    to hand a consumed DLPack tensor back to its producer."""

    def __init__(self, ptr, shape, release=None, owner=None):
        self.ptr, self.shape = ptr, shape
        self._release, self._owner = release, owner

    def close(self):
        if self._release is not None:
            release, self._release = self._release, None
            release()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _check(shape, typestr_ok, contiguous, name):
    problem = None
    if len(shape) != 2:
        problem = f"must be 2D (points x features), got shape {tuple(shape)}"
    elif min(shape) == 0:
        problem = f"must be non-empty, got shape {tuple(shape)}"
    elif not typestr_ok:
        problem = "must hold float32 values (convert it on the device first)"
    elif not contiguous:
        problem = "must be C-contiguous (copy it on the device first)"
    if problem:
        raise ValueError(f"{name} {problem}")


def _from_cai(obj, name):
    cai = obj.__cuda_array_interface__
    shape = tuple(int(s) for s in cai["shape"])
    typestr = np.dtype(cai["typestr"])
    strides = cai.get("strides")
    contiguous = strides is None or (len(shape) == 2 and
                                     tuple(strides) == (shape[1] * 4, 4))
    _check(shape, typestr == np.dtype("<f4"), contiguous, name)
    ptr = int(cai["data"][0])
    return DeviceMatrix(ptr, shape, owner=obj)


def _from_dlpack(obj, name):
    capsule = obj.__dlpack__()
    if not _api.PyCapsule_IsValid(capsule, b"dltensor"):
        raise ValueError(f"{name}: not an unconsumed DLPack capsule")
    managed_ptr = _api.PyCapsule_GetPointer(capsule, b"dltensor")
    managed = ctypes.cast(managed_ptr, ctypes.POINTER(_DLManagedTensor))
    t = managed.contents.dl_tensor
    shape = tuple(t.shape[i] for i in range(t.ndim))
    if t.strides:
        strides = tuple(t.strides[i] for i in range(t.ndim))
        contiguous = len(shape) != 2 or strides == (shape[1], 1) or shape[0] == 1
    else:
        contiguous = True
    dtype_ok = t.dtype.code == _DL_FLOAT and t.dtype.bits == 32 and t.dtype.lanes == 1
    # The capsule is consumed: its destructor must no longer free the tensor;
    # the deleter runs when the view is closed.
    _api.PyCapsule_SetName(capsule, b"used_dltensor")

    def release():
        if managed.contents.deleter:
            managed.contents.deleter(managed)

    try:
        _check(shape, dtype_ok, contiguous, name)
    except ValueError:
        release()
        raise
    return DeviceMatrix(int(t.data or 0) + int(t.byte_offset), shape, release, capsule)


def device_matrix(obj, name):
    """A DeviceMatrix when obj is a GPU array (CUDA array interface, or DLPack
    on a ROCm device), else None (host data: the caller converts it)."""
    if hasattr(obj, "__cuda_array_interface__"):
        return _from_cai(obj, name)
    if hasattr(obj, "__dlpack_device__") and hasattr(obj, "__dlpack__"):
        device_type, _ = obj.__dlpack_device__()
        if int(device_type) == DL_ROCM:
            return _from_dlpack(obj, name)
    return None


def gpu_shape(obj):
    """The shape of a GPU array (CUDA array interface, or DLPack on a ROCm
    device), without consuming it; None for host data."""
    if hasattr(obj, "__cuda_array_interface__"):
        return tuple(int(s) for s in obj.__cuda_array_interface__["shape"])
    if hasattr(obj, "__dlpack_device__") and hasattr(obj, "__dlpack__"):
        if int(obj.__dlpack_device__()[0]) == DL_ROCM:
            return tuple(int(s) for s in obj.shape)
    return None


def feature_ranges(lib, dm):
    """Per-feature (min, max) of the device points, computed on the GPU."""
    from . import _lib
    n, d = dm.shape
    lo = np.empty(d, np.float32)
    hi = np.empty(d, np.float32)
    _lib.check(lib.cuhc_feature_ranges_device(ctypes.c_void_p(dm.ptr), n, d, _lib.float_ptr(lo),
                                              _lib.float_ptr(hi)))
    return lo, hi
