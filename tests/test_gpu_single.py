"""Single linkage through Boruvka on the GPU: plain FP32 tiles, the pruned walk,
and the FP16-MFMA certified-candidate form all give the oracle's tree bit for
bit (reference tests/test_mst.py, tests/test_spatial_pruning.py)."""

import numpy as np
import pytest
from scipy.cluster.hierarchy import linkage as scipy_linkage

import cuhc_amd as cuhc
import oracle
from _helpers import assert_valid_tree, blobs, max_rel_height_error, tree_clusters
from cuhc_amd import _lib

pytestmark = pytest.mark.gpu


def run(X, metric="euclidean", precision="fp32", pruning="never", p=2.0):
    _lib.set_precision(precision)
    _lib.set_pruning(pruning)
    try:
        return cuhc.linkage(X, method="single", metric=metric, p=p, algorithm="mst")
    finally:
        _lib.set_precision("auto")
        _lib.set_pruning("auto")


@pytest.mark.parametrize("metric", ["euclidean", "sqeuclidean", "manhattan", "chebyshev",
                                    "cosine"])
@pytest.mark.parametrize("n", [2, 3, 127, 128, 129, 300])
def test_plain_matches_oracle(metric, n):
    X = blobs(n_samples=n, seed=n)
    np.testing.assert_array_equal(run(X, metric), oracle.single_linkage(X, metric))


@pytest.mark.parametrize("n, d", [(2, 32), (129, 16), (300, 32), (1000, 20), (2500, 64),
                                  (3000, 32)])
@pytest.mark.parametrize("metric", ["euclidean", "sqeuclidean"])
def test_mixed_matches_oracle(metric, n, d):
    X = np.random.default_rng(n + d).standard_normal((n, d), dtype=np.float32)
    ours = run(X, metric, precision="mixed")
    np.testing.assert_array_equal(ours, oracle.single_linkage(X, metric))


def test_mixed_on_clustered_and_offset_data():
    # Far from the origin and strongly clustered: the FP16 bound must still
    # cover every candidate (centring and scaling are per dataset).
    X = blobs(n_samples=2000, n_features=24, centers=7, seed=5) + np.float32(1000.0)
    np.testing.assert_array_equal(run(X, precision="mixed"), oracle.single_linkage(X))
    _lib.mixed_stats()
    Y = np.repeat(np.random.default_rng(4).normal(size=(300, 16)), 3, axis=0).astype(np.float32)
    np.testing.assert_array_equal(run(Y, precision="mixed"), run(Y, precision="fp32"))


@pytest.mark.parametrize("d", [16, 18])
def test_mixed_overflowing_lists(d):
    # 150 copies of each of 200 points, shuffled: every row has 149 columns at
    # weight 0 spread over about as many half-blocks, more than a candidate
    # list holds, so its rows are rechecked by the chunked full scan (d = 18:
    # rows not a multiple of 4 floats, the scalar load paths).
    rng = np.random.default_rng(8)
    base = rng.normal(size=(200, d)).astype(np.float32)
    X = np.repeat(base, 150, axis=0)[rng.permutation(30_000)]
    print(f"[single] overflowing lists n=30000 d={d}", flush=True)
    _lib.mixed_stats()
    Z = run(X, precision="mixed")
    stats = _lib.mixed_stats()
    assert stats["overflow_rows"] > 0, stats
    np.testing.assert_array_equal(Z, run(X, precision="fp32"))


@pytest.mark.parametrize("pruning", ["never", "always"])
def test_majority_component_rounds(pruning):
    # One dense cloud holding most points and a few small far groups: once the
    # cloud holds more than half of the points it proposes no edge and joins
    # through the small groups' smallest edges (mixed and pruned walk passes).
    rng = np.random.default_rng(17)
    big = rng.standard_normal((5000, 32)).astype(np.float32)
    small = [rng.normal(loc=c, scale=0.3, size=(40, 32)) for c in (12.0, -15.0, 25.0)]
    line = np.zeros((30, 32), np.float32)
    line[:, 0] = 40.0 + np.arange(30, dtype=np.float32)  # a chain of equal gaps
    X = np.concatenate([big, *small, line]).astype(np.float32)
    X = X[rng.permutation(len(X))]
    expected = oracle.single_linkage(X)
    np.testing.assert_array_equal(run(X, precision="mixed"), expected)
    np.testing.assert_array_equal(run(X, pruning=pruning), expected)


def test_mixed_on_tied_grid():
    g = np.stack(np.meshgrid(np.arange(20), np.arange(20)), axis=-1).reshape(-1, 2)
    X = np.hstack([g, np.zeros((400, 14))]).astype(np.float32)
    Z = run(X, precision="mixed")
    np.testing.assert_array_equal(Z, oracle.single_linkage(X))
    np.testing.assert_array_equal(Z[:, 2], 1.0)


@pytest.mark.parametrize("n_features", [1, 2, 3, 8, 20])
@pytest.mark.parametrize("metric", ["euclidean", "sqeuclidean", "manhattan", "chebyshev"])
def test_pruned_equals_unpruned(metric, n_features):
    X = blobs(n_samples=3000, n_features=n_features, centers=30, seed=n_features)
    pruned = run(X, metric, pruning="always")
    full = run(X, metric, pruning="never")
    np.testing.assert_array_equal(pruned, full)
    np.testing.assert_array_equal(full, oracle.single_linkage(X, metric))


def test_pruning_skips_tiles_on_clustered_data():
    X = blobs(n_samples=20000, n_features=2, centers=30, seed=30)
    _lib.prune_stats()
    run(X, pruning="always")
    computed, total = _lib.prune_stats()
    assert total > 0 and computed < 0.2 * total


def test_grid_ties_and_sizes_off_tiles():
    g = np.stack(np.meshgrid(np.arange(20), np.arange(20)), axis=-1).reshape(-1, 2)
    X = g.astype(np.float32)
    Z = run(X)
    np.testing.assert_array_equal(Z, oracle.single_linkage(X))
    np.testing.assert_array_equal(Z[:, 2], 1.0)
    for n in [129, 255, 257, 1000]:
        Y = blobs(n_samples=n, n_features=3, seed=n)
        np.testing.assert_array_equal(run(Y, pruning="always"), oracle.single_linkage(Y))


def test_c1_recipe_against_oracle_and_scipy():
    # C1: n = 20,000, d = 32, N(0,1) float32, seed 0. Bit-identical to the oracle
    # in both precisions; the same dendrogram as scipy in float64 with the
    # largest relative height error reported.
    X = np.random.default_rng(0).standard_normal((20_000, 32), dtype=np.float32)
    ref = oracle.single_linkage(X)
    plain = run(X, precision="fp32")
    mixed = run(X, precision="mixed")
    np.testing.assert_array_equal(plain, ref)
    np.testing.assert_array_equal(mixed, ref)
    # Against scipy in float64: float32 rounding makes some heights tie (or
    # swap order) that float64 separates, which renumbers a few clusters; count
    # them and bound the height error on all clusters the trees share.
    sp = scipy_linkage(X.astype(np.float64), "single")
    ours, ref = tree_clusters(mixed), tree_clusters(sp)
    swapped = len(ours.keys() - ref.keys())
    err = max_rel_height_error(mixed, sp)
    print(f"C1 vs scipy float64: {swapped} of {len(ref)} clusters differ by tie order; "
          f"max relative height error {err:.3e}")
    assert swapped <= 0.001 * len(ref)
    assert err < 1e-5


def test_auto_and_estimator():
    X = blobs(n_samples=500, seed=14)
    np.testing.assert_array_equal(cuhc.linkage(X), run(X))
    labels = cuhc.AgglomerativeClustering(n_clusters=6, linkage="single").fit(X).labels_
    assert len(set(labels.tolist())) == 6
    assert_valid_tree(cuhc.linkage(X), 500)


def test_device_entry_point_matches_host_entry():
    X = np.random.default_rng(9).standard_normal((3000, 32), dtype=np.float32)
    dev = _lib.DeviceArray(X)
    tree = np.empty((2999, 4), dtype=np.float64)
    for mode in ["fp32", "mixed"]:
        _lib.set_precision(mode)
        try:
            _lib.check(_lib.lib.cuhc_linkage_device(dev.ptr, 3000, 32, 0, 2.0, 0, 1,
                                                    _lib.double_ptr(tree)))
        finally:
            _lib.set_precision("auto")
        np.testing.assert_array_equal(tree, oracle.single_linkage(X))
    H = np.ascontiguousarray(X[:, :8])
    hdev = _lib.DeviceArray(H)
    core = np.empty(3000, dtype=np.float32)
    _lib.check(_lib.lib.cuhc_hdbscan_tree_device(hdev.ptr, 3000, 8, 0, 2.0, 5, 1.0, 0,
                                                 _lib.double_ptr(tree), _lib.float_ptr(core)))
    Zo, co = oracle.hdbscan_tree(H, 5)
    np.testing.assert_array_equal(core, co)
    np.testing.assert_array_equal(tree, Zo)
    dev.free()
    hdev.free()


def kruskal_tree(n, edges):
    """Scipy-format tree of edges (w, u, v) taken in the given order (Kruskal)."""
    parent = np.arange(n)
    cid, size = np.arange(n), np.ones(n, np.int64)

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x
    rows = []
    for w, u, v in edges:
        ru, rv = find(u), find(v)
        if ru == rv:
            continue
        a, b = cid[ru], cid[rv]
        rows.append([min(a, b), max(a, b), w, size[ru] + size[rv]])
        parent[rv] = ru
        size[ru] += size[rv]
        cid[ru] = n + len(rows) - 1
    return np.array(rows, dtype=np.float64)


@pytest.mark.parametrize("pruning", ["never", "always"])
def test_large_grid_every_weight_tied(pruning):
    # 300 x 240 integer grid (72,000 points, above the size where the edge
    # order is sorted on the GPU): every MST edge weighs exactly 1, so the tree
    # is Kruskal over the unit edges in (u, v) order, which the keys (weight,
    # smaller endpoint) alone do not fix.
    W, H = 300, 240
    g = np.stack(np.meshgrid(np.arange(W), np.arange(H)), axis=-1).reshape(-1, 2)
    X = g.astype(np.float32)
    n = len(X)
    idx = np.arange(n).reshape(H, W)
    right = np.stack([idx[:, :-1].ravel(), idx[:, 1:].ravel()], 1)
    down = np.stack([idx[:-1, :].ravel(), idx[1:, :].ravel()], 1)
    unit = np.concatenate([right, down])
    unit = unit[np.lexsort((unit[:, 1], unit[:, 0]))]
    expected = kruskal_tree(n, [(1.0, int(u), int(v)) for u, v in unit])
    np.testing.assert_array_equal(run(X, pruning=pruning), expected)


@pytest.mark.parametrize("pruning", ["never", "always"])
def test_large_shuffled_line_with_tied_gaps(pruning):
    # 80,000 shuffled points on a line at integer positions with gaps in
    # [1, 50]: distances are exact, every non-neighbouring pair is strictly
    # farther than both gaps between them, so the MST is the chain of
    # neighbours; its order (gap, smaller index, larger index) has ~1,600 edges
    # per tied weight.
    rng = np.random.default_rng(17)
    n = 80_000
    pos = np.cumsum(rng.integers(1, 51, n)).astype(np.float32)
    perm = rng.permutation(n)
    X = np.empty((n, 1), np.float32)
    X[perm, 0] = pos  # point perm[i] sits at the i-th position
    u, v = perm[:-1], perm[1:]
    lo, hi = np.minimum(u, v), np.maximum(u, v)
    gap = np.diff(pos).astype(np.float64)
    order = np.lexsort((hi, lo, gap))
    expected = kruskal_tree(n, [(gap[i], int(lo[i]), int(hi[i])) for i in order])
    np.testing.assert_array_equal(run(X, pruning=pruning), expected)


@pytest.mark.parametrize("metric", ["euclidean", "chebyshev"])
def test_pruned_large_clustered_identical(metric):
    # 60,000 points in 4-d blobs of unequal spread: many pruned rounds with
    # component bounds; the unpruned pass is the reference.
    rng = np.random.default_rng(8)
    centres = rng.uniform(-30, 30, (25, 4))
    X = np.concatenate([c + rng.normal(scale=s, size=(2_400, 4))
                        for c, s in zip(centres, rng.uniform(0.2, 3.0, 25))]).astype(np.float32)
    np.testing.assert_array_equal(run(X, metric, pruning="always"),
                                  run(X, metric, pruning="never"))
