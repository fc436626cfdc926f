"""NaN and infinity through the raw C ABI (the Python layer rejects them
first; a C caller gets status -6 before any distance is computed, in every
precision mode and algorithm, host or device input)."""

import numpy as np
import pytest

from cuhc_amd import _lib

pytestmark = pytest.mark.gpu
NON_FINITE = -6


@pytest.fixture(params=["nan", "inf", "-inf"])
def bad_points(request):
    X = np.random.default_rng(4).standard_normal((3000, 32), dtype=np.float32)
    X[1234, 5] = float(request.param)  # inside the first 8 features too
    return X


@pytest.mark.parametrize("precision", ["auto", "fp32", "mixed"])
@pytest.mark.parametrize("pruning", ["auto", "always"])
def test_single_linkage_rejects_non_finite(bad_points, precision, pruning):
    X = bad_points
    n, d = X.shape
    tree = np.zeros((n - 1, 4))
    dev = _lib.DeviceArray(X)
    try:
        _lib.set_precision(precision)
        _lib.set_pruning(pruning)
        assert _lib.lib.cuhc_linkage(_lib.float_ptr(X), n, d, 0, 2.0, 0, 1,
                                     _lib.double_ptr(tree)) == NON_FINITE
        assert _lib.lib.cuhc_linkage_device(dev.ptr, n, d, 0, 2.0, 0, 1,
                                            _lib.double_ptr(tree)) == NON_FINITE
    finally:
        _lib.set_precision("auto")
        _lib.set_pruning("auto")
        dev.free()
    assert not tree.any()


@pytest.mark.parametrize("linkage_id, algorithm", [(3, 3), (3, 2), (2, 0), (1, 0)])
def test_other_algorithms_reject_non_finite(bad_points, linkage_id, algorithm):
    X = bad_points[:500].copy()
    X[7, 3] = np.nan
    n, d = X.shape
    tree = np.zeros((n - 1, 4))
    assert _lib.lib.cuhc_linkage(_lib.float_ptr(X), n, d, 0, 2.0, linkage_id, algorithm,
                                 _lib.double_ptr(tree)) == NON_FINITE
    assert _lib.lib.cuhc_ward_rnn_with_options(_lib.float_ptr(X), n, d, 0.01, 0, 1,
                                               _lib.double_ptr(tree)) == NON_FINITE


@pytest.mark.parametrize("core_method", [0, 1, 2, 3])
def test_hdbscan_tree_rejects_non_finite(bad_points, core_method):
    X = bad_points[:, :8].copy()
    n, d = X.shape
    tree = np.zeros((n - 1, 4))
    core = np.zeros(n, np.float32)
    dev = _lib.DeviceArray(X)
    try:
        assert _lib.lib.cuhc_hdbscan_tree(_lib.float_ptr(X), n, d, 0, 2.0, 10, 1.0, core_method,
                                          _lib.double_ptr(tree), _lib.float_ptr(core)) == NON_FINITE
        assert _lib.lib.cuhc_hdbscan_tree_device(dev.ptr, n, d, 0, 2.0, 10, 1.0, core_method,
                                                 _lib.double_ptr(tree),
                                                 _lib.float_ptr(core)) == NON_FINITE
    finally:
        dev.free()


def test_finite_input_still_runs_after_rejections():
    X = np.random.default_rng(5).standard_normal((2000, 32), dtype=np.float32)
    tree = np.zeros((1999, 4))
    _lib.check(_lib.lib.cuhc_linkage(_lib.float_ptr(X), 2000, 32, 0, 2.0, 0, 1,
                                     _lib.double_ptr(tree)))
    import oracle
    np.testing.assert_array_equal(tree, oracle.single_linkage(X))
