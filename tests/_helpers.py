"""Shared test data and tree/partition assertions (the reference's test
strategy, reference tests/_helpers.py: tie-free blobs, partition bijections,
dendrogram comparison as sets of clusters)."""

import numpy as np
from sklearn.datasets import make_blobs


def blobs(n_samples=400, n_features=8, centers=5, noise_fraction=0.0, seed=0):
    """Clustered float32 points (continuous, so ties are vanishingly unlikely),
    with an optional share of uniform background noise."""
    rng = np.random.default_rng(seed)
    n_noise = int(n_samples * noise_fraction)
    X, _ = make_blobs(n_samples=n_samples - n_noise, n_features=n_features, centers=centers,
                      cluster_std=rng.uniform(0.5, 2.0, size=centers), random_state=seed)
    if n_noise:
        X = np.vstack([X, rng.uniform(X.min(0), X.max(0), size=(n_noise, n_features))])
    return X.astype(np.float32)


def subtree_sizes(children, n):
    sizes = np.ones(2 * n - 1, dtype=np.int64)
    for i, (a, b) in enumerate(np.asarray(children, dtype=np.int64)):
        sizes[n + i] = sizes[a] + sizes[b]
    return sizes[n:]


def assert_same_partition(ours, ref):
    """Same clusters up to relabelling; noise (-1) must match exactly."""
    ours, ref = np.asarray(ours), np.asarray(ref)
    assert ours.shape == ref.shape
    np.testing.assert_array_equal(ours == -1, ref == -1, err_msg="noise points differ")
    keep = ref != -1
    pairs = set(zip(ours[keep].tolist(), ref[keep].tolist()))
    assert len(pairs) == len({a for a, _ in pairs}) == len({b for _, b in pairs}), \
        "partitions differ"


def tree_clusters(Z):
    """{(smallest point, size): height} for every cluster a merge tree creates."""
    Z = np.asarray(Z)
    n = len(Z) + 1
    smallest = np.empty(2 * n - 1, dtype=np.int64)
    smallest[:n] = np.arange(n)
    out = {}
    for i, (a, b, h, s) in enumerate(Z):
        smallest[n + i] = min(smallest[int(a)], smallest[int(b)])
        out[(int(smallest[n + i]), int(s))] = float(h)
    return out


def assert_same_dendrogram(Z, ref_Z, rtol=1e-4, atol=1e-5):
    """Same clusters at the same heights; tied merges may come in any order."""
    ours, ref = tree_clusters(Z), tree_clusters(ref_Z)
    assert ours.keys() == ref.keys(), f"{len(ours.keys() - ref.keys())} clusters differ"
    keys = sorted(ref)
    np.testing.assert_allclose([ours[k] for k in keys], [ref[k] for k in keys], rtol=rtol,
                               atol=atol)


def assert_valid_tree(Z, n):
    ids = np.asarray(Z)[:, :2].astype(np.int64)
    assert np.all(ids[:, 0] < ids[:, 1])
    np.testing.assert_array_equal(np.sort(ids.ravel()), np.arange(2 * n - 2))
    assert np.all(ids < n + np.arange(n - 1)[:, None])
    np.testing.assert_array_equal(np.asarray(Z)[:, 3], subtree_sizes(ids, n))


def max_rel_height_error(Z, ref_Z):
    """Largest relative height difference between same-cluster merges."""
    ours, ref = tree_clusters(Z), tree_clusters(ref_Z)
    common = sorted(ours.keys() & ref.keys())
    a = np.array([ours[k] for k in common])
    b = np.array([ref[k] for k in common])
    with np.errstate(divide="ignore", invalid="ignore"):
        rel = np.where(b != 0, np.abs(a - b) / np.abs(b), np.abs(a - b))
    return float(rel.max()) if len(rel) else 0.0
