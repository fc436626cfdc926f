"""Parsing of GPU arrays handed to the Python surface (no GPU needed: nothing
here dereferences the addresses): the CUDA array interface and ROCm DLPack
capsules become (device address, shape) views; other dtypes, layouts and
devices are refused or left to the host path; a consumed DLPack tensor is
handed back to its producer exactly once."""

import numpy as np
import pytest

from _dlpack import Producer
from cuhc_amd._device import device_matrix, gpu_shape


class CAI:
    def __init__(self, ptr, shape, typestr="<f4", strides=None):
        self.__cuda_array_interface__ = {"shape": shape, "typestr": typestr, "strides": strides,
                                         "data": (ptr, False), "version": 3}


def test_array_interface_view():
    dm = device_matrix(CAI(0x7000, (10, 3)), "X")
    assert (dm.ptr, dm.shape) == (0x7000, (10, 3))
    assert gpu_shape(CAI(0x7000, (10, 3))) == (10, 3)
    assert device_matrix(CAI(0x7000, (10, 3), strides=(12, 4)), "X").shape == (10, 3)


@pytest.mark.parametrize("kwargs, message", [
    ({"typestr": "<f8"}, "float32"),
    ({"strides": (4, 40)}, "C-contiguous"),
    ({"shape": (10,)}, "2D"),
    ({"shape": (0, 3)}, "non-empty"),
])
def test_array_interface_refusals(kwargs, message):
    args = {"shape": (10, 3), **kwargs}
    with pytest.raises(ValueError, match=message):
        device_matrix(CAI(0x7000, **args), "X")


def test_dlpack_rocm_view_and_release():
    prod = Producer(0x9000, (6, 4), offset=64)
    assert gpu_shape(prod) == (6, 4)  # read without consuming a capsule
    assert prod.deleted == 0
    dm = device_matrix(prod, "X")
    assert (dm.ptr, dm.shape) == (0x9000 + 64, (6, 4))
    assert prod.deleted == 0
    dm.close()
    dm.close()
    assert prod.deleted == 1


@pytest.mark.parametrize("kwargs, message", [
    ({"bits": 64}, "float32"),
    ({"code": 0}, "float32"),
    ({"strides": (1, 6)}, "C-contiguous"),
])
def test_dlpack_refusals_release_the_tensor(kwargs, message):
    prod = Producer(0x9000, (6, 4), **kwargs)
    with pytest.raises(ValueError, match=message):
        device_matrix(prod, "X")
    assert prod.deleted == 1


def test_host_data_is_not_a_device_view():
    assert device_matrix(np.zeros((3, 2), np.float32), "X") is None
    assert device_matrix(Producer(0x9000, (3, 2), device_type=1), "X") is None  # kDLCPU
    assert device_matrix([[0.0, 1.0]], "X") is None
