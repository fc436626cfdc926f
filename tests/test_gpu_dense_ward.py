"""Dense Lance-Williams linkages and Ward (rnn rounds, tails, nn_chain) on the GPU
(reference tests/test_linkage.py, test_rnn_ward.py, test_nn_chain.py)."""

import pathlib

import numpy as np
import pytest
from scipy.cluster.hierarchy import fcluster as scipy_fcluster
from scipy.cluster.hierarchy import linkage as scipy_linkage
from scipy.spatial.distance import pdist

import cuhc_amd as cuhc
import oracle
from _helpers import (assert_same_dendrogram, assert_same_partition, assert_valid_tree, blobs,
                      max_rel_height_error)
from cuhc_amd._linkage import _ward_rnn_with_options

pytestmark = pytest.mark.gpu
GOLDEN = np.load(pathlib.Path(__file__).resolve().parent / "golden" / "golden.npz")


def scipy_ward(X):
    return scipy_linkage(X.astype(np.float64), "ward")


@pytest.mark.parametrize("method", ["single", "complete", "average", "ward"])
@pytest.mark.parametrize("metric", ["euclidean", "manhattan", "chebyshev", "cosine"])
@pytest.mark.parametrize("n", [2, 3, 130, 400])
def test_dense_matches_oracle(method, metric, n):
    if method == "ward" and metric != "euclidean":
        pytest.skip("ward is euclidean only")
    X = blobs(n_samples=n, seed=n)
    Z = cuhc.linkage(X, method=method, metric=metric, algorithm="dense")
    np.testing.assert_array_equal(Z, oracle.dense_linkage(X, method, metric))


@pytest.mark.parametrize("method, metric", [("single", "euclidean"), ("single", "cosine"),
                                            ("complete", "manhattan"), ("average", "cosine"),
                                            ("average", "euclidean"), ("ward", "euclidean")])
def test_dense_matches_scipy_golden(method, metric):
    X = GOLDEN["link_X"]
    Z = cuhc.linkage(X, method=method, metric=metric, algorithm="dense")
    ref = GOLDEN[f"link_{method}_{metric}"]
    np.testing.assert_array_equal(np.sort(Z[:, :2], 1), np.sort(ref[:, :2], 1))
    np.testing.assert_allclose(Z[:, 2], ref[:, 2], rtol=1e-4, atol=1e-5)


def test_dense_larger_and_duplicates():
    X = blobs(n_samples=2500, n_features=16, centers=12, seed=31)
    for method in ["complete", "average"]:
        np.testing.assert_array_equal(cuhc.linkage(X, method=method),
                                      oracle.dense_linkage(X, method))
    Y = np.repeat(np.random.default_rng(9).normal(size=(40, 3)), 5, axis=0).astype(np.float32)
    for method in ["single", "complete", "average", "ward"]:
        Z = cuhc.linkage(Y, method=method, algorithm="dense")
        assert_valid_tree(Z, 200)
        np.testing.assert_allclose(Z[:160, 2], 0.0, atol=1e-5)
        assert Z[160, 2] > 1e-5


@pytest.mark.parametrize("method", ["single", "complete", "average"])
def test_dense_star_many_rescans(method):
    # Random unit vectors and the origin last: every row's nearest is the
    # origin (distance about 1, all other pairs farther), so the first merge
    # leaves every row without its nearest: every slot range lists more rows
    # to rescan than an update record carries inline (1401 rows over at most
    # 256 workgroups is 6 per workgroup).
    m = 1400
    rng = np.random.default_rng(5)
    V = rng.normal(size=(m, 128))
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    X = np.vstack([V, np.zeros((1, 128))]).astype(np.float32)
    assert pdist(X[:m]).min() > np.linalg.norm(X[:m], axis=1).max()
    Z = cuhc.linkage(X, method=method, algorithm="dense")
    np.testing.assert_array_equal(Z, oracle.dense_linkage(X, method))


@pytest.mark.parametrize("n", [2, 3, 127, 128, 129, 600])
@pytest.mark.parametrize("d", [1, 7, 64])
def test_rnn_matches_scipy(n, d):
    X = blobs(n_samples=n, n_features=d, seed=n + d)
    Z = cuhc.linkage(X, method="ward", algorithm="rnn")
    assert_same_dendrogram(Z, scipy_ward(X), rtol=1e-5, atol=1e-6)


def test_rnn_many_points_and_error_bound():
    X = blobs(n_samples=6000, n_features=10, centers=25, seed=51)
    Z = cuhc.linkage(X, method="ward")
    ref = scipy_ward(X)
    assert_same_dendrogram(Z, ref, rtol=1e-5, atol=1e-6)
    assert max_rel_height_error(Z, ref) < 1e-5


@pytest.mark.parametrize("tail", ["dense", "chain"])
def test_rnn_tails(tail):
    X = blobs(n_samples=2500, n_features=8, seed=53)
    ref = scipy_ward(X)
    Z = _ward_rnn_with_options(X, min_progress=1.0, allow_dense=(tail == "dense"))
    assert_same_dendrogram(Z, ref, rtol=1e-5, atol=1e-6)
    if tail == "dense":
        assert_same_dendrogram(_ward_rnn_with_options(X, 0.0, dense_below=len(X)), ref,
                               rtol=1e-5, atol=1e-6)
        assert_same_dendrogram(_ward_rnn_with_options(X, 0.0, dense_below=500), ref,
                               rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("allow_dense", [True, False])
def test_rnn_worst_case_line(allow_dense):
    x = 1.01 ** np.arange(1500, dtype=np.float64)
    X = ((x - x[0]) / (x[-1] - x[0]))[:, None].astype(np.float32)
    Z = _ward_rnn_with_options(X, min_progress=0.01, allow_dense=allow_dense)
    assert_same_dendrogram(Z, scipy_ward(X), rtol=1e-4, atol=1e-5)


@pytest.mark.parametrize("n", [2, 3, 255, 256, 257, 600])
@pytest.mark.parametrize("d", [1, 7, 64])
def test_nn_chain_matches_scipy(n, d):
    X = blobs(n_samples=n, n_features=d, seed=n + d)
    Z = cuhc.linkage(X, method="ward", algorithm="nn_chain")
    assert_same_dendrogram(Z, scipy_ward(X), rtol=1e-5, atol=1e-6)


def test_ward_algorithms_agree_and_are_deterministic():
    X = blobs(n_samples=3000, n_features=16, seed=52)
    rnn = cuhc.linkage(X, method="ward", algorithm="rnn")
    for other in ["nn_chain", "dense"]:
        Z = cuhc.linkage(X, method="ward", algorithm=other)
        assert_same_dendrogram(rnn, Z, rtol=1e-5, atol=1e-6)
        for k in [2, 10, 100]:
            assert_same_partition(scipy_fcluster(rnn, k, "maxclust"),
                                  scipy_fcluster(Z, k, "maxclust"))
    Y = blobs(n_samples=4000, n_features=4, seed=55)
    first = cuhc.linkage(Y, method="ward")
    for _ in range(3):
        np.testing.assert_array_equal(cuhc.linkage(Y, method="ward"), first)


def test_ward_duplicates():
    X = np.repeat(np.random.default_rng(56).normal(size=(60, 3)), 5, axis=0).astype(np.float32)
    for algorithm in ["rnn", "nn_chain"]:
        Z = cuhc.linkage(X, method="ward", algorithm=algorithm)
        np.testing.assert_array_equal(Z[:240, 2], 0.0)
        assert_same_partition(scipy_fcluster(Z, 60, "maxclust"), np.repeat(np.arange(60), 5))


@pytest.mark.parametrize("linkage", ["single", "complete", "average", "ward"])
def test_agglomerative_matches_sklearn(linkage):
    from sklearn.cluster import AgglomerativeClustering as SkAgg
    X = blobs(n_samples=300, seed=5)
    ours = cuhc.AgglomerativeClustering(n_clusters=5, linkage=linkage,
                                        compute_distances=True).fit(X)
    ref = SkAgg(n_clusters=5, linkage=linkage, compute_distances=True).fit(X.astype(np.float64))
    assert_same_partition(ours.labels_, ref.labels_)
    np.testing.assert_array_equal(np.sort(ours.children_, 1), np.sort(ref.children_, 1))
    np.testing.assert_allclose(ours.distances_, ref.distances_, rtol=1e-4, atol=1e-5)


@pytest.mark.parametrize("n_features", [2, 8, 20])
def test_rnn_pruned_equals_unpruned(n_features):
    # Reference tests/test_spatial_pruning.py:55-59: the pruned rnn passes
    # (points in k-d order, cluster tiles walked with box bounds) give the
    # unpruned dendrogram, and both give scipy's.
    X = blobs(n_samples=6000, n_features=n_features, centers=30, seed=10 + n_features)
    trees, stats = {}, {}
    try:
        for mode in ["always", "never"]:
            from cuhc_amd import _lib
            _lib.set_pruning(mode)
            _lib.prune_stats()
            trees[mode] = cuhc.linkage(X, method="ward", algorithm="rnn")
            stats[mode] = _lib.prune_stats()
    finally:
        _lib.set_pruning("auto")
    assert_same_dendrogram(trees["always"], trees["never"], rtol=1e-5, atol=1e-6)
    assert_same_dendrogram(trees["always"], scipy_ward(X), rtol=1e-5, atol=1e-6)
    computed, total = stats["always"]
    print(n_features, stats)
    assert 0 < computed < total  # some tile pairs were skipped
    assert stats["never"] == (0, 0)
