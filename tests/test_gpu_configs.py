"""The five benchmark workloads at their full sizes (reference pattern
tests/test_mst.py:33-45: a realistic size checked against an independent
computation). Where an oracle run of the whole workload is out of reach,
size-independent properties and FP64 recomputations of samples are checked,
and the largest FP64 error is printed for each result.

C1 (single, n = 20,000, d = 32) runs bit-exact against the oracle in
test_gpu_single.py; this file covers the other four.
"""

import time

import numpy as np
import pytest
from scipy.cluster.hierarchy import linkage as scipy_linkage

import cuhc_amd as cuhc
from _helpers import max_rel_height_error, tree_clusters
from bench import make_c3, make_points
from cuhc_amd import _lib
from cuhc_amd._hdbscan import hdbscan_tree, tree_to_labels

pytestmark = pytest.mark.gpu


def note(msg):
    print(f"[configs] {msg}", flush=True)


def check_tree_invariants(Z, n):
    """Every id used once, children before parents, sizes add up, heights
    non-decreasing."""
    ids = Z[:, :2].astype(np.int64)
    assert np.all(ids[:, 0] < ids[:, 1])
    np.testing.assert_array_equal(np.sort(ids.ravel()), np.arange(2 * n - 2))
    assert np.all(ids[:, 1] < n + np.arange(n - 1))
    assert np.all(np.diff(Z[:, 2]) >= 0)
    size = [1] * n + [0] * (n - 1)
    for i, (a, b) in enumerate(ids.tolist()):
        size[n + i] = size[a] + size[b]
    np.testing.assert_array_equal(Z[:, 3], size[n:])


def first_merge_height(Z, n):
    """Height at which each point first joins another cluster."""
    out = np.empty(n)
    ids = Z[:, :2].astype(np.int64)
    for col in (0, 1):
        leaf = ids[:, col] < n
        out[ids[leaf, col]] = Z[leaf, 2]
    return out


def fp64_kth_distances(X, rows, k):
    """k-th smallest FP64 euclidean distance from X[rows] to every point
    (itself counted first, at 0), in chunks of the point set."""
    S = X[rows].astype(np.float64)
    ss = (S * S).sum(1)
    best = np.full((len(rows), 0), np.inf)
    for start in range(0, len(X), 100_000):
        C = X[start:start + 100_000].astype(np.float64)
        sq = np.maximum(ss[:, None] + (C * C).sum(1)[None, :] - 2.0 * S @ C.T, 0.0)
        own = (rows >= start) & (rows < start + len(C))
        sq[np.flatnonzero(own), rows[own] - start] = 0.0
        part = np.partition(sq, k - 1, axis=1)[:, :k] if sq.shape[1] >= k else sq
        best = np.partition(np.hstack([best, part]), k - 1, axis=1)[:, :k]
    return np.sqrt(np.sort(best, axis=1)[:, k - 1])


def test_c4_single_linkage_1M_mixed_equals_plain_and_fp64_nearest():
    # C4: n = 1,000,000, d = 32, N(0,1), seed 3. The mixed form (FP16 MFMA
    # candidates + certified FP32 recheck) must give the plain FP32 tree bit
    # for bit. Each point's first merge height is its FP32 nearest-neighbour
    # distance (the nearest edge is in the MST): 1,000 sampled points are
    # compared with an FP64 brute force over all 10^6 points.
    n, d = 1_000_000, 32
    X = make_points(n, d, 3)
    try:
        trees = {}
        for mode in ["mixed", "fp32"]:
            _lib.set_precision(mode)
            t0 = time.perf_counter()
            trees[mode] = cuhc.linkage(X, method="single")
            note(f"C4 {mode}: {time.perf_counter() - t0:.2f} s")
    finally:
        _lib.set_precision("auto")
    np.testing.assert_array_equal(trees["mixed"], trees["fp32"])
    Z = trees["mixed"]
    check_tree_invariants(Z, n)
    rows = np.random.default_rng(30).choice(n, 1000, replace=False)
    ref = fp64_kth_distances(X, rows, 2)  # the 2nd smallest: the nearest other point
    got = first_merge_height(Z, n)[rows]
    err = np.abs(got - ref) / ref
    note(f"C4 sampled nearest-neighbour heights vs FP64: max relative error {err.max():.3e}")
    assert err.max() < 1e-5


def test_c3_hdbscan_1M_methods_agree_fp64_core_and_sklearn_labels():
    # C3: n = 1,000,000, d = 8, 20 blobs + 10% noise, min_samples = mcs = 25.
    # The pruned auto path, the learned-radius core distances and the unpruned
    # search give identical core distances and trees; 1,000 sampled core
    # distances match FP64; the labels equal scikit-learn's extraction run on
    # the same tree.
    from sklearn.cluster._hdbscan._linkage import HIERARCHY_dtype
    from sklearn.cluster._hdbscan._tree import tree_to_labels as sk_tree_to_labels
    X = make_c3()
    n, k = len(X), 25
    runs = {}
    try:
        for name, method, prune in [("auto", "auto", "auto"), ("learned", "learned", "auto"),
                                    ("unpruned", "fused", "never")]:
            _lib.set_pruning(prune)
            t0 = time.perf_counter()
            runs[name] = hdbscan_tree(X, k, core_method=method)
            note(f"C3 {name}: {time.perf_counter() - t0:.2f} s")
    finally:
        _lib.set_pruning("auto")
    Z, core = runs["auto"]
    for name in ["learned", "unpruned"]:
        np.testing.assert_array_equal(runs[name][1], core, err_msg=name)
        np.testing.assert_array_equal(runs[name][0], Z, err_msg=name)
    check_tree_invariants(Z, n)
    rows = np.random.default_rng(31).choice(n, 1000, replace=False)
    ref = fp64_kth_distances(X, rows, k)
    err = np.abs(core[rows] - ref) / ref
    note(f"C3 sampled core distances vs FP64: max relative error {err.max():.3e}")
    assert err.max() < 1e-5
    t0 = time.perf_counter()
    labels, probs = tree_to_labels(Z, k)
    note(f"C3 extraction: {time.perf_counter() - t0:.2f} s, {labels.max() + 1} clusters, "
         f"{(labels < 0).sum()} noise points")
    h = np.empty(n - 1, dtype=HIERARCHY_dtype)
    for i, name in enumerate(["left_node", "right_node", "value", "cluster_size"]):
        h[name] = Z[:, i]
    ref_labels, ref_probs = sk_tree_to_labels(h, k)
    np.testing.assert_array_equal(labels, ref_labels)
    np.testing.assert_allclose(probs, ref_probs, rtol=0, atol=1e-12)
    assert labels.max() + 1 >= 20


def dendrogram_report(name, Z, ref):
    ours, theirs = tree_clusters(Z), tree_clusters(ref)
    differ = len(ours.keys() - theirs.keys())
    err = max_rel_height_error(Z, ref)
    note(f"{name}: {differ} of {len(theirs)} clusters differ by tie order; "
         f"max relative height error {err:.3e}")
    return differ, err, len(theirs)


def test_c2_ward_100k_rnn_equals_dense_and_fp64_prefix():
    # C2: n = 100,000, d = 32, N(0,1), seed 1. RNN Ward (FP64 centroids) against
    # the GPU dense Lance-Williams Ward on the same points, and RNN Ward on the
    # first 20,000 points against scipy's FP64 Ward.
    X = make_points(100_000, 32, 1)
    t0 = time.perf_counter()
    Zr = cuhc.linkage(X, method="ward")
    note(f"C2 rnn: {time.perf_counter() - t0:.2f} s")
    t0 = time.perf_counter()
    Zd = cuhc.linkage(X, method="ward", algorithm="dense")
    note(f"C2 dense: {time.perf_counter() - t0:.2f} s")
    check_tree_invariants(Zr, len(X))
    differ, err, total = dendrogram_report("C2 rnn vs GPU dense", Zr, Zd)
    assert differ <= 0.001 * total and err < 1e-5
    P = X[:20_000]
    differ, err, total = dendrogram_report(
        "C2 20k prefix vs scipy FP64", cuhc.linkage(P, method="ward"),
        scipy_linkage(P.astype(np.float64), "ward"))
    assert differ <= 0.001 * total and err < 1e-5


def cluster_members(Z, n, wanted):
    """{row: (points of side a, points of side b)} for the merges `wanted`, all
    of which make clusters of at most 64 points."""
    points = {}
    out = {}
    wanted = set(wanted.tolist())
    for i, (a, b) in enumerate(Z[:, :2].astype(np.int64).tolist()):
        if Z[i, 3] > 64:
            continue
        sides = [[a] if a < n else points[a], [b] if b < n else points[b]]
        points[n + i] = sides[0] + sides[1]
        if i in wanted:
            out[i] = sides
    return out


def test_c5_average_cosine_100k_tree_and_fp64_heights():
    # C5: n = 100,000, d = 64, N(0,1) rows L2-normalised in float32, seed 4,
    # average linkage on cosine distances (20 GB condensed matrix). The tree is
    # valid with non-decreasing heights; 2,000 sampled merges of clusters of
    # at most 64 points have the FP64 mean pairwise cosine distance of their
    # two sides as height; the first 20,000 points match scipy in FP64.
    X = make_points(100_000, 64, 4)
    X /= np.linalg.norm(X, axis=1, keepdims=True)
    n = len(X)
    t0 = time.perf_counter()
    Z = cuhc.linkage(X, method="average", metric="cosine")
    note(f"C5: {time.perf_counter() - t0:.2f} s")
    check_tree_invariants(Z, n)
    small = np.flatnonzero(Z[:, 3] <= 64)
    wanted = np.random.default_rng(32).choice(small, min(2000, len(small)), replace=False)
    X64 = X.astype(np.float64)
    norms = np.linalg.norm(X64, axis=1)
    errs = []
    for i, (pa, pb) in cluster_members(Z, n, wanted).items():
        A, B = X64[pa], X64[pb]
        cos = 1.0 - (A @ B.T) / np.outer(norms[pa], norms[pb])
        ref = cos.mean()
        errs.append(abs(Z[i, 2] - ref) / ref)
    note(f"C5 sampled small-cluster heights vs FP64 means: max relative error {max(errs):.3e}")
    assert max(errs) < 1e-5
    P = X[:20_000]
    differ, err, total = dendrogram_report(
        "C5 20k prefix vs scipy FP64", cuhc.linkage(P, method="average", metric="cosine"),
        scipy_linkage(P.astype(np.float64), "average", metric="cosine"))
    assert differ <= 0.001 * total and err < 1e-5
