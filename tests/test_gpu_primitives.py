"""The library's own device scan and radix sorts (they replace CUB's in the
contraction, the mixed pass's component order and the edge sort), against
numpy: exclusive sums equal np.cumsum shifted by one, sorted pairs equal a
stable argsort of the keys (ties keep input order), at tile edges (4,096
items per workgroup), with few and many distinct keys."""

import ctypes

import numpy as np
import pytest

from cuhc_amd import _lib

pytestmark = pytest.mark.gpu


def device_out(nbytes):
    ptr = ctypes.c_void_p()
    _lib.check(_lib.lib.cuhc_device_alloc(max(nbytes, 1), ctypes.byref(ptr)))
    return ptr


def to_host(ptr, like):
    out = np.empty_like(like)
    _lib.check(_lib.lib.cuhc_copy_to_host(out.ctypes.data_as(ctypes.c_void_p), ptr, out.nbytes))
    return out


SIZES = [1, 2, 255, 4095, 4096, 4097, 70_000, 1_048_579]


@pytest.mark.parametrize("count", SIZES + [16_777_217])
def test_exclusive_scan(count):
    rng = np.random.default_rng(count)
    x = rng.integers(0, 3, count).astype(np.int32)
    src = _lib.DeviceArray(x)
    out = device_out(x.nbytes)
    try:
        print(f"[primitives] scan count={count}", flush=True)
        _lib.check(_lib.lib.cuhc_exclusive_scan(src.ptr, out, count))
        got = to_host(out, x)
        _lib.check(_lib.lib.cuhc_exclusive_scan(src.ptr, src.ptr, count))  # in place
        inplace = to_host(src.ptr, x)
    finally:
        _lib.lib.cuhc_device_free(out)
        src.free()
    want = np.concatenate([[0], np.cumsum(x[:-1], dtype=np.int64)]).astype(np.int32)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(inplace, want)


def check_sort(keys, bits, wide):
    count = len(keys)
    values = np.arange(count, dtype=np.int32 if not wide else np.uint32)
    k_in, v_in = _lib.DeviceArray(keys), _lib.DeviceArray(values)
    k_out, v_out = device_out(keys.nbytes), device_out(values.nbytes)
    try:
        fn = _lib.lib.cuhc_radix_sort_pairs_u64 if wide else _lib.lib.cuhc_radix_sort_pairs_u32
        _lib.check(fn(k_in.ptr, k_out, v_in.ptr, v_out, count, bits))
        got_k, got_v = to_host(k_out, keys), to_host(v_out, values)
        same_in = to_host(k_in.ptr, keys)
    finally:
        for p in (k_out, v_out):
            _lib.lib.cuhc_device_free(p)
        k_in.free()
        v_in.free()
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(got_k, keys[order])
    np.testing.assert_array_equal(got_v, values[order])
    np.testing.assert_array_equal(same_in, keys)


@pytest.mark.parametrize("count", SIZES)
@pytest.mark.parametrize("bits", [1, 7, 8, 13, 20, 32])
def test_radix_sort_u32(count, bits):
    rng = np.random.default_rng(bits * 7 + count)
    high = 1 << bits
    distinct = min(high, max(2, count // 50))  # many ties
    pool = rng.integers(0, high, distinct, dtype=np.uint64)
    keys = rng.choice(pool, count).astype(np.uint32)
    print(f"[primitives] sort32 count={count} bits={bits}", flush=True)
    check_sort(keys, bits, wide=False)


@pytest.mark.parametrize("count", [1, 4097, 1_048_579])
@pytest.mark.parametrize("bits", [20, 52, 64])
def test_radix_sort_u64(count, bits):
    rng = np.random.default_rng(bits + count)
    keys = rng.integers(0, 1 << 62, count, dtype=np.uint64) >> np.uint64(max(0, 62 - bits))
    keys[: count // 3] = keys[0]  # a long run of equal keys
    rng.shuffle(keys)
    print(f"[primitives] sort64 count={count} bits={bits}", flush=True)
    check_sort(keys.astype(np.uint64), bits, wide=True)


def test_radix_sort_zero_bits_copies():
    keys = np.zeros(5000, np.uint32)
    check_sort(keys, 0, wide=False)


@pytest.mark.parametrize("shape", [0, 1, 2, 3, 4, 5, 6])
def test_mfma_probe_runs(shape):
    """The MFMA issue-rate probe runs each shape and counts its flops exactly
    (16 MFMAs per wave per step, 8 for shapes 4 and 5; 32,768 flops each,
    16,384 for the 16x16x32 loops; the symmetric shape 6 counts each value
    twice); a bad shape is refused."""
    ms, flops = ctypes.c_float(), ctypes.c_double()
    print(f"[primitives] mfma probe shape={shape}", flush=True)
    _lib.check(_lib.lib.cuhc_mfma_probe(shape, 8, 16, ctypes.byref(ms), ctypes.byref(flops)))
    assert ms.value > 0.0
    per_step = 8 if shape in (4, 5) else 16
    per_mfma = 16384.0 if shape in (3, 5) else 32768.0
    assert flops.value == 8 * 4 * 16 * per_step * per_mfma * (2 if shape == 6 else 1)
    assert _lib.lib.cuhc_mfma_probe(7, 8, 16, ctypes.byref(ms), ctypes.byref(flops)) != 0
