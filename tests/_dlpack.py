"""A minimal DLPack producer for tests: wraps an existing buffer (a device
address from the library's allocator, or any address for parsing tests) in a
"dltensor" capsule with the given shape, dtype and device, and records when
its deleter runs. It owns nothing: the buffer's lifetime is the caller's."""

import ctypes

from cuhc_amd._device import _DELETER, _DLManagedTensor

_api = ctypes.pythonapi
_api.PyCapsule_New.restype = ctypes.py_object
_api.PyCapsule_New.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]
_NAME = b"dltensor"


class Producer:
    def __init__(self, ptr, shape, device_type=10, code=2, bits=32, strides=None, offset=0):
        self.ptr, self.shape, self.device_type = ptr, tuple(shape), device_type
        self.code, self.bits, self.strides, self.offset = code, bits, strides, offset
        self.deleted = 0
        self._keep = []

    def __dlpack_device__(self):
        return (self.device_type, 0)

    def __dlpack__(self, stream=None):
        mt = _DLManagedTensor()
        shape = (ctypes.c_int64 * len(self.shape))(*self.shape)
        t = mt.dl_tensor
        t.data = ctypes.c_void_p(self.ptr)
        t.device.device_type, t.device.device_id = self.device_type, 0
        t.ndim = len(self.shape)
        t.dtype.code, t.dtype.bits, t.dtype.lanes = self.code, self.bits, 1
        t.shape = shape
        strides = None
        if self.strides is not None:
            strides = (ctypes.c_int64 * len(self.strides))(*self.strides)
            t.strides = strides
        t.byte_offset = self.offset

        def deleter(_):
            self.deleted += 1

        callback = _DELETER(deleter)
        mt.deleter = callback
        self._keep.append((mt, shape, strides, callback))
        return _api.PyCapsule_New(ctypes.addressof(mt), _NAME, None)
