"""Pins the CPU oracle before it is trusted as the GPU parity checker.

Each oracle function is checked against something independent of it: the
scipy/sklearn golden vectors (tests/golden/golden.npz), the reference's own
known answers (reference python/cuhc/_linkage.py:106-111, tests/test_mst.py:65-80),
brute force on tiny inputs, and invariants. Runs on the CPU.
"""

import itertools
import pathlib

import numpy as np
import pytest

import oracle
from _helpers import assert_same_dendrogram, assert_valid_tree, blobs, max_rel_height_error

GOLDEN = np.load(pathlib.Path(__file__).resolve().parent / "golden" / "golden.npz")


def brute_force_mst(X, metric="euclidean", core=None, alpha=1.0):
    """Kruskal over all edges sorted by (w, u, v), in float32 via numpy per pair."""
    n = len(X)
    edges = []
    for u, v in itertools.combinations(range(n), 2):
        w = np.float32(oracle.pairwise(X[u:u + 1], X[v:v + 1], metric)[0, 0])
        if core is not None:
            w = np.float32(max(core[u], core[v], w if alpha == 1.0 else np.float32(w / alpha)))
        edges.append((float(w), u, v))
    edges.sort()
    parent = list(range(n))

    def find(x):
        while parent[x] != x:
            x = parent[x]
        return x
    out = []
    for w, u, v in edges:
        ru, rv = find(u), find(v)
        if ru != rv:
            parent[ru] = rv
            out.append((u, v, w))
    return sorted(out)


def test_known_answer_single_linkage():
    # reference python/cuhc/_linkage.py:106-111
    Z = oracle.single_linkage(np.array([[0.0], [1.0], [5.0], [7.0]]))
    np.testing.assert_array_equal(Z, [[0, 1, 1, 2], [2, 3, 2, 2], [4, 5, 4, 4]])


def test_pairwise_matches_sklearn_golden():
    P = GOLDEN["pair_X"]
    for metric in ["euclidean", "sqeuclidean", "manhattan", "chebyshev", "cosine"]:
        np.testing.assert_allclose(oracle.pairwise(P, metric=metric), GOLDEN[f"pair_{metric}"],
                                   rtol=1e-5, atol=1e-5, err_msg=metric)
    np.testing.assert_allclose(oracle.pairwise(P, metric="minkowski", p=3.0),
                               GOLDEN["pair_minkowski3"], rtol=1e-5, atol=1e-5)


def test_pairwise_is_the_fma_fold():
    # A dropped fma, a reordered fold or a missing sqrt changes the bits.
    rng = np.random.default_rng(3)
    a, b = rng.normal(size=(2, 37)).astype(np.float32)
    acc = np.float32(0.0)
    for k in range(37):
        diff = np.float32(a[k] - b[k])
        acc = np.float32(np.float64(diff) * np.float64(diff) + np.float64(acc))  # exact fma
    assert oracle.pairwise(a[None], b[None], "sqeuclidean")[0, 0] == acc
    assert oracle.pairwise(a[None], b[None], "euclidean")[0, 0] == np.sqrt(acc, dtype=np.float32)
    # cosine of a zero vector is 1 (reference src/distances.cuh:78-82)
    z = np.zeros((1, 3), np.float32)
    assert oracle.pairwise(z, np.ones((1, 3), np.float32), "cosine")[0, 0] == 1.0


@pytest.mark.parametrize("method, metric", [("single", "euclidean"), ("single", "cosine"),
                                            ("complete", "manhattan"), ("average", "cosine"),
                                            ("average", "euclidean"), ("ward", "euclidean")])
def test_trees_match_scipy_golden(method, metric):
    X = GOLDEN["link_X"]
    ref = GOLDEN[f"link_{method}_{metric}"]
    if method == "single":
        Z = oracle.single_linkage(X, metric)
    elif method == "ward":
        Z = oracle.dense_on_matrix(oracle.pairwise(X), "ward")
    else:
        Z = oracle.dense_linkage(X, method, metric)
    np.testing.assert_array_equal(np.sort(Z[:, :2], 1), np.sort(ref[:, :2], 1))
    np.testing.assert_allclose(Z[:, 2], ref[:, 2], rtol=1e-4, atol=1e-5)
    assert_valid_tree(Z, len(X))


@pytest.mark.parametrize("n", [2, 3, 5, 17, 40])
@pytest.mark.parametrize("metric", ["euclidean", "manhattan", "cosine"])
def test_mst_equals_brute_force(n, metric):
    X = np.random.default_rng(n).normal(size=(n, 3)).astype(np.float32)
    eu, ev, ew = oracle.mst(X, metric)
    ours = sorted(zip(eu.tolist(), ev.tolist(), ew.tolist()))
    assert ours == brute_force_mst(X, metric)


def test_mst_ties_follow_the_total_order():
    # Integer grid: every nearest distance ties at 1.0; the total order
    # (w, u, v) fixes one tree, which brute-force Kruskal also finds.
    g = np.stack(np.meshgrid(np.arange(5), np.arange(4)), axis=-1).reshape(-1, 2)
    X = g.astype(np.float32)
    eu, ev, ew = oracle.mst(X)
    assert sorted(zip(eu.tolist(), ev.tolist(), ew.tolist())) == brute_force_mst(X)


def test_grid_heights_all_one():
    # reference tests/test_mst.py:65-80
    g = np.stack(np.meshgrid(np.arange(20), np.arange(20)), axis=-1).reshape(-1, 2)
    Z = oracle.single_linkage(g.astype(np.float32))
    np.testing.assert_array_equal(Z[:, 2], 1.0)
    assert_valid_tree(Z, 400)


def test_mreach_mst_equals_brute_force():
    X = blobs(n_samples=30, n_features=3, seed=2)
    core = oracle.core_distances(X, 4)
    for alpha in [1.0, 0.7]:
        eu, ev, ew = oracle.mst(X, core=core, alpha=alpha)
        assert sorted(zip(eu.tolist(), ev.tolist(), ew.tolist())) == \
            brute_force_mst(X, core=core, alpha=alpha)


@pytest.mark.parametrize("k", [1, 2, 7, 40])
def test_core_distances_match_golden(k):
    X = GOLDEN["core_X"]
    core = oracle.core_distances(X, k)
    np.testing.assert_allclose(core, GOLDEN[f"core_k{k}"], rtol=1e-5, atol=1e-6)
    # bit-exact against np.partition of the oracle's own FP32 matrix
    D = oracle.pairwise(X)
    np.fill_diagonal(D, 0.0)
    np.testing.assert_array_equal(core, np.sort(np.maximum(D, 0), axis=1)[:, k - 1])


def test_core_distances_with_duplicates():
    X = np.repeat(np.random.default_rng(3).normal(size=(50, 3)), 4, axis=0).astype(np.float32)
    np.testing.assert_array_equal(oracle.core_distances(X, 4), 0.0)
    assert np.all(oracle.core_distances(X, 5) > 0)


def test_hdbscan_tree_min_samples_one_is_single_linkage():
    X = GOLDEN["hdb_X"]
    Z, core = oracle.hdbscan_tree(X, 1)
    np.testing.assert_array_equal(core, 0.0)
    np.testing.assert_array_equal(Z, oracle.single_linkage(X))


def test_dense_on_matrix_equals_naive():
    # O(n^3) restatement: merge the smallest (d, i, j) with i < j each step.
    rng = np.random.default_rng(11)
    X = rng.normal(size=(25, 4)).astype(np.float32)
    for method in ["single", "complete", "average", "ward"]:
        D = oracle.pairwise(X).astype(np.float64)
        D = D.astype(np.float32)
        n = len(X)
        size, ids, active = [1] * n, list(range(n)), [True] * n
        rows = []
        for step in range(n - 1):
            best = None
            for i in range(n):
                for j in range(i + 1, n):
                    if active[i] and active[j] and (best is None or D[i, j] < best[0]):
                        best = (D[i, j], i, j)
            d, a, b = best
            rows.append([min(ids[a], ids[b]), max(ids[a], ids[b]), float(d), size[a] + size[b]])
            active[b] = False
            for k in range(n):
                if active[k] and k != a:
                    dak, dbk = float(D[a, k]), float(D[b, k])
                    na, nb, nk = size[a], size[b], size[k]
                    if method == "single":
                        v = min(dak, dbk)
                    elif method == "complete":
                        v = max(dak, dbk)
                    elif method == "average":
                        v = (na * dak + nb * dbk) / (na + nb)
                    else:
                        v = np.sqrt(max(((na + nk) * dak * dak + (nb + nk) * dbk * dbk
                                         - nk * float(d) * float(d)) / (na + nb + nk), 0.0))
                    D[a, k] = D[k, a] = np.float32(v)
            size[a] += size[b]
            ids[a] = n + step
        np.testing.assert_array_equal(oracle.dense_on_matrix(oracle.pairwise(X), method),
                                      np.array(rows), err_msg=method)


def test_oracle_single_linkage_c1_scale_against_scipy():
    # 3,000 points of the C1 recipe (N(0,1), d=32): same children as scipy in
    # float64, heights within float32 rounding.
    from scipy.cluster.hierarchy import linkage
    X = np.random.default_rng(0).standard_normal((3000, 32), dtype=np.float32)
    Z = oracle.single_linkage(X)
    ref = linkage(X.astype(np.float64), "single")
    assert_same_dendrogram(Z, ref, rtol=1e-6, atol=0)
    assert max_rel_height_error(Z, ref) < 1e-6


def test_mlp_gradient_matches_finite_differences():
    # Central differences of the oracle's own FP64 loss pin its chain rule: a
    # dropped tanh' factor or a transposed W2 fails by orders of magnitude.
    rng = np.random.default_rng(0)
    p = rng.normal(scale=0.3, size=oracle.MLP_PARAMS).astype(np.float32)
    X = rng.normal(size=(17, 4)).astype(np.float32)
    t = rng.normal(size=17).astype(np.float32)
    g, loss = oracle.mlp_gradient(p, X, t)
    y = oracle.mlp_forward(p, X)
    assert loss == pytest.approx(0.5 * np.mean((y - t.astype(np.float64)) ** 2), rel=1e-12)
    # p is float32, so perturb in float32 steps large enough to be exact.
    idx = rng.choice(oracle.MLP_PARAMS, size=60, replace=False)
    for q in idx:
        h = np.float32(1e-2) * max(1.0, abs(float(p[q])))
        up, dn = p.copy(), p.copy()
        up[q] += h
        dn[q] -= h
        fd = (oracle.mlp_gradient(up, X, t)[1] - oracle.mlp_gradient(dn, X, t)[1]) / \
            (float(up[q]) - float(dn[q]))
        assert g[q] == pytest.approx(fd, rel=2e-3, abs=1e-6), q
