"""The Python surface on points already in GPU memory (read in place through
the CUDA array interface or a ROCm DLPack capsule): the same trees, labels
and distances as from host arrays, and the same argument errors."""

import numpy as np
import pytest

import cuhc_amd as cuhc
from _dlpack import Producer
from _helpers import blobs
from cuhc_amd import _lib
from cuhc_amd._hdbscan import hdbscan_tree

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def data():
    X = blobs(n_samples=3000, n_features=8, centers=10, noise_fraction=0.1, seed=3)
    dev = _lib.DeviceArray(X)
    yield X, dev
    dev.free()


@pytest.mark.parametrize("method", ["single", "ward", "average", "complete"])
def test_linkage_from_device_points(data, method):
    X, dev = data
    np.testing.assert_array_equal(cuhc.linkage(dev, method=method), cuhc.linkage(X, method=method))


def test_linkage_from_dlpack_capsule(data):
    X, dev = data
    prod = Producer(dev.ptr.value, X.shape)
    np.testing.assert_array_equal(cuhc.linkage(prod, method="single"),
                                  cuhc.linkage(X, method="single"))
    assert prod.deleted == 1


def test_hdbscan_from_device_points(data):
    X, dev = data
    Z, core = hdbscan_tree(dev, 10)
    Zh, ch = hdbscan_tree(X, 10)
    np.testing.assert_array_equal(core, ch)
    np.testing.assert_array_equal(Z, Zh)
    a = cuhc.HDBSCAN(min_cluster_size=15).fit(dev)
    b = cuhc.HDBSCAN(min_cluster_size=15).fit(X)
    np.testing.assert_array_equal(a.labels_, b.labels_)
    assert a.n_features_in_ == 8


def test_agglomerative_from_device_points(data):
    X, dev = data
    a = cuhc.AgglomerativeClustering(n_clusters=5).fit(dev)
    b = cuhc.AgglomerativeClustering(n_clusters=5).fit(X)
    np.testing.assert_array_equal(a.labels_, b.labels_)
    np.testing.assert_array_equal(a.children_, b.children_)


def test_pairwise_distances_stay_on_device(data):
    X, dev = data
    Y = X[:700].copy()
    devY = _lib.DeviceArray(Y)
    try:
        out = cuhc.pairwise_distances(dev, devY, metric="manhattan")
        assert isinstance(out, _lib.DeviceArray) and out.shape == (3000, 700)
        np.testing.assert_array_equal(out.to_host(),
                                      cuhc.pairwise_distances(X, Y, metric="manhattan"))
        out.free()
        same = cuhc.pairwise_distances(devY)
        np.testing.assert_array_equal(same.to_host(), cuhc.pairwise_distances(Y))
        same.free()
        with pytest.raises(ValueError, match="GPU memory"):
            cuhc.pairwise_distances(dev, Y)
    finally:
        devY.free()


def test_device_points_errors():
    bad = np.random.default_rng(0).standard_normal((500, 4)).astype(np.float32)
    bad[7, 2] = np.nan
    dev = _lib.DeviceArray(bad)
    try:
        with pytest.raises(ValueError, match="NaN"):
            cuhc.linkage(dev)
        with pytest.raises(ValueError, match="NaN"):
            hdbscan_tree(dev, 5)
    finally:
        dev.free()
    huge = np.zeros((10, 2), np.float32)
    huge[0, 0] = 3e38
    huge[1, 0] = -3e38
    dev = _lib.DeviceArray(huge)
    try:
        with pytest.raises(ValueError, match="too large for float32"):
            cuhc.linkage(dev)
    finally:
        dev.free()
    dev = _lib.DeviceArray(np.zeros((10, 2), np.float64))
    try:
        with pytest.raises(ValueError, match="float32"):
            cuhc.linkage(dev)
    finally:
        dev.free()
