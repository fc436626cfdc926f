"""HDBSCAN on the GPU: core distances and the mutual-reachability tree bit for
bit against the oracle; labels exactly sklearn's extraction of our tree, and
sklearn end to end (reference tests/test_hdbscan.py, tests/test_spatial_pruning.py)."""

import pathlib

import numpy as np
import pytest
from scipy.cluster.hierarchy import cophenet
from sklearn.cluster import HDBSCAN as SkHDBSCAN
from sklearn.cluster._hdbscan._linkage import HIERARCHY_dtype
from sklearn.cluster._hdbscan._tree import tree_to_labels as sk_tree_to_labels
from sklearn.metrics import adjusted_rand_score

import cuhc_amd as cuhc
import oracle
from _helpers import assert_same_partition, blobs
from cuhc_amd import _lib
from cuhc_amd._hdbscan import hdbscan_tree

pytestmark = pytest.mark.gpu
GOLDEN = np.load(pathlib.Path(__file__).resolve().parent / "golden" / "golden.npz")


def structured(Z):
    tree = np.empty(len(Z), dtype=HIERARCHY_dtype)
    tree["left_node"], tree["right_node"] = Z[:, 0], Z[:, 1]
    tree["value"], tree["cluster_size"] = Z[:, 2], Z[:, 3]
    return tree


@pytest.mark.parametrize("metric", ["euclidean", "manhattan", "chebyshev", "cosine"])
@pytest.mark.parametrize("k", [1, 2, 7, 25, 40])
def test_tree_and_core_match_oracle(metric, k):
    X = blobs(n_samples=600, n_features=4, centers=6, noise_fraction=0.1, seed=k)
    Z, core = hdbscan_tree(X, k, metric=metric)
    Zo, co = oracle.hdbscan_tree(X, k, metric=metric)
    np.testing.assert_array_equal(core, co)
    np.testing.assert_array_equal(Z, Zo)


@pytest.mark.parametrize("k", [1, 5, 64, 65, 200])
@pytest.mark.parametrize("metric", ["euclidean", "cosine"])
def test_core_methods_agree(k, metric):
    X = blobs(n_samples=1000, n_features=6, seed=k)
    _, radix = hdbscan_tree(X, k, metric=metric, core_method="radix")
    np.testing.assert_array_equal(radix, oracle.core_distances(X, k, metric))
    if k <= 64:
        _, fused = hdbscan_tree(X, k, metric=metric, core_method="fused")
        np.testing.assert_array_equal(fused, radix)
    else:
        with pytest.raises(RuntimeError, match="algorithm"):
            hdbscan_tree(X, k, metric=metric, core_method="fused")


@pytest.mark.parametrize("n_features", [2, 8])
@pytest.mark.parametrize("k", [1, 5, 40])
@pytest.mark.parametrize("alpha", [1.0, 0.7, 1.7])
def test_pruned_tree_is_identical(n_features, k, alpha):
    X = blobs(n_samples=3000, n_features=n_features, centers=30, seed=20 + n_features)
    out = {}
    for mode in ["always", "never"]:
        _lib.set_pruning(mode)
        try:
            out[mode] = hdbscan_tree(X, k, alpha=alpha)
        finally:
            _lib.set_pruning("auto")
    np.testing.assert_array_equal(out["always"][1], out["never"][1])
    np.testing.assert_array_equal(out["always"][0], out["never"][0])
    Zo, co = oracle.hdbscan_tree(X, k, alpha=alpha)
    np.testing.assert_array_equal(out["never"][0], Zo)


def test_labels_are_sklearn_extraction_of_our_tree():
    X = blobs(n_samples=2000, n_features=8, centers=8, noise_fraction=0.1, seed=7)
    for params in [dict(min_cluster_size=25), dict(min_cluster_size=10,
                                                   cluster_selection_method="leaf"),
                   dict(min_cluster_size=15, min_samples=5, max_cluster_size=300)]:
        ours = cuhc.HDBSCAN(**params).fit(X)
        ext = {k: v for k, v in params.items() if k != "min_samples"}
        labels, probs = sk_tree_to_labels(structured(ours._single_linkage_tree_), **ext)
        np.testing.assert_array_equal(ours.labels_, labels)
        np.testing.assert_allclose(ours.probabilities_, probs, rtol=0, atol=1e-12)


def test_end_to_end_against_sklearn_golden():
    X = GOLDEN["hdb_X"]
    ours = cuhc.HDBSCAN(min_cluster_size=20, min_samples=1).fit(X)
    assert_same_partition(ours.labels_, GOLDEN["hdb_labels"])
    np.testing.assert_allclose(ours.probabilities_, GOLDEN["hdb_probabilities"], atol=1e-3)


@pytest.mark.parametrize("metric", ["euclidean", "manhattan", "cosine"])
def test_end_to_end_with_ties_against_sklearn(metric):
    X = blobs(n_samples=600, n_features=4, centers=6, noise_fraction=0.1, seed=15)
    ours = cuhc.HDBSCAN(min_cluster_size=15, metric=metric).fit(X)
    ref = SkHDBSCAN(min_cluster_size=15, metric=metric, copy=True).fit(X.astype(np.float64))
    np.testing.assert_allclose(cophenet(ours._single_linkage_tree_),
                               cophenet(np.column_stack([
                                   ref._single_linkage_tree_["left_node"],
                                   ref._single_linkage_tree_["right_node"],
                                   ref._single_linkage_tree_["value"],
                                   ref._single_linkage_tree_["cluster_size"]]).astype(float)),
                               rtol=1e-5, atol=1e-6)
    assert adjusted_rand_score(ref.labels_, ours.labels_) >= 0.98


@pytest.mark.parametrize("k, alpha", [(3, 1.0), (8, 1.0), (8, 0.6)])
def test_pruned_ties_identical(k, alpha):
    # A lattice (every point has several neighbours at exactly the same
    # distance) plus duplicated points, shuffled: mutual-reachability weights
    # tie at the core distances, so the four keys each row keeps in the first
    # round, and the rows later rounds settle from them, meet ties broken by
    # original index. Pruned runs (kept keys, open-row walks) against the
    # unpruned pass and the oracle.
    rng = np.random.default_rng(70 + k)
    g = np.stack(np.meshgrid(np.arange(14), np.arange(14), np.arange(8)), axis=-1).reshape(-1, 3)
    dup = np.repeat(rng.uniform(0, 14, size=(120, 3)), 3, axis=0)
    X = np.concatenate([g, dup]).astype(np.float32)[rng.permutation(len(g) + len(dup))]
    print(f"[hdbscan] pruned ties n={len(X)} k={k} alpha={alpha}", flush=True)
    out = {}
    for mode in ["always", "never"]:
        _lib.set_pruning(mode)
        try:
            out[mode] = hdbscan_tree(X, k, alpha=alpha)
        finally:
            _lib.set_pruning("auto")
    np.testing.assert_array_equal(out["always"][1], out["never"][1])
    np.testing.assert_array_equal(out["always"][0], out["never"][0])
    Zo, co = oracle.hdbscan_tree(X, k, alpha=alpha)
    np.testing.assert_array_equal(out["never"][1], co)
    np.testing.assert_array_equal(out["never"][0], Zo)


def test_duplicates_and_single_point():
    X = np.repeat(np.random.default_rng(3).normal(size=(50, 3)), 4, axis=0).astype(np.float32)
    np.testing.assert_array_equal(cuhc.HDBSCAN(min_cluster_size=5, min_samples=4)
                                  .fit(X).core_distances_, 0.0)
    assert np.all(cuhc.HDBSCAN(min_cluster_size=5, min_samples=5).fit(X).core_distances_ > 0)


@pytest.mark.parametrize("metric", ["euclidean", "manhattan"])
@pytest.mark.parametrize("k, alpha", [(1, 1.0), (10, 1.0), (25, 0.8)])
def test_pruned_large_blobs_and_noise_identical(metric, k, alpha):
    # 40,000 points: 12 blobs of unequal spread plus 15% uniform noise, so the
    # pruned passes go through many rounds with a few large components, where
    # rows stop at their component's best known edge (and the points are
    # reordered by core-distance class); the unpruned symmetric pass is the
    # reference.
    rng = np.random.default_rng(40 + k)
    centres = rng.uniform(-20, 20, (12, 5))
    parts = [c + rng.normal(scale=s, size=(2_833, 5)) for c, s in
             zip(centres, rng.uniform(0.3, 2.5, 12))]
    X = np.concatenate(parts)
    X = np.concatenate([X, rng.uniform(X.min(0), X.max(0), (6_000, 5))]).astype(np.float32)
    out = {}
    for mode in ["always", "never"]:
        _lib.set_pruning(mode)
        try:
            out[mode] = hdbscan_tree(X, k, metric=metric, alpha=alpha)
        finally:
            _lib.set_pruning("auto")
    np.testing.assert_array_equal(out["always"][1], out["never"][1])
    np.testing.assert_array_equal(out["always"][0], out["never"][0])
