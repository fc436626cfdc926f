"""cuhc_linkage_device_noalloc through the C ABI: the same tree as
cuhc_linkage_device on the same device points, call after call (the fixed
workspace keeps the previous call's contents), in every precision and pruning
mode, at the largest accepted size, and a clear error beyond it."""

import numpy as np
import pytest

from cuhc_amd import _lib

pytestmark = pytest.mark.gpu
WORKSPACE_ERROR = -7


def trees(dev, n, d):
    a = np.full((n - 1, 4), np.nan)
    b = np.full((n - 1, 4), np.nan)
    _lib.check(_lib.lib.cuhc_linkage_device(dev.ptr, n, d, 0, 2.0, 0, 1, _lib.double_ptr(a)))
    _lib.check(_lib.lib.cuhc_linkage_device_noalloc(dev.ptr, n, d, 0, 2.0, 0, 1,
                                                    _lib.double_ptr(b)))
    return a, b


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
@pytest.mark.parametrize("pruning", ["never", "always"])
def test_noalloc_equals_linkage_device_twice(precision, pruning):
    rng = np.random.default_rng(7)
    shapes = [(20_000, 32), (5_000, 16)]  # the second call is smaller
    try:
        _lib.set_precision(precision)
        _lib.set_pruning(pruning)
        for n, d in shapes:
            print(f"[noalloc] {precision} {pruning} n={n} d={d}", flush=True)
            X = rng.standard_normal((n, d), dtype=np.float32)
            dev = _lib.DeviceArray(X)
            try:
                ref, got = trees(dev, n, d)
            finally:
                dev.free()
            np.testing.assert_array_equal(got, ref)
    finally:
        _lib.set_precision("auto")
        _lib.set_pruning("auto")


def test_noalloc_largest_accepted_size_and_rejections():
    n, d = 65_536, 64
    X = np.random.default_rng(8).standard_normal((n, d), dtype=np.float32)
    dev = _lib.DeviceArray(X)
    try:
        for precision in ["mixed", "fp32"]:
            print(f"[noalloc] largest size, {precision}", flush=True)
            _lib.set_precision(precision)
            ref, got = trees(dev, n, d)
            np.testing.assert_array_equal(got, ref)
        _lib.set_precision("auto")
        tree = np.zeros((n, 4))
        L = _lib.lib
        assert L.cuhc_linkage_device_noalloc(dev.ptr, n + 1, d - 1, 0, 2.0, 0, 1,
                                             _lib.double_ptr(tree)) == WORKSPACE_ERROR
        assert L.cuhc_linkage_device_noalloc(dev.ptr, n // 2, d + 1, 0, 2.0, 0, 1,
                                             _lib.double_ptr(tree)) == WORKSPACE_ERROR
        assert L.cuhc_linkage_device_noalloc(dev.ptr, 100, d, 0, 2.0, 3, 3,
                                             _lib.double_ptr(tree)) == -3
        assert L.cuhc_linkage_device_noalloc(dev.ptr, 100, d, 0, 2.0, 0, 0,
                                             _lib.double_ptr(tree)) == -4
        assert not tree.any()
        assert "workspace" in L.cuhc_error_string(WORKSPACE_ERROR).decode()
    finally:
        _lib.set_precision("auto")
        dev.free()
