"""CPU tests of the C ABI library and the host-side code (no GPU work).

libcuhc.so loads and exports every symbol include/cuhc.h declares; the
host-only entry points (condense, prune mode, error strings) and the Python
extraction reproduce sklearn's golden outputs; argument checks raise the
reference's errors before any GPU call.
"""

import ctypes
import pathlib
import re

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
GOLDEN = np.load(ROOT / "tests" / "golden" / "golden.npz")

import cuhc_amd as cuhc  # noqa: E402
from cuhc_amd import _hdbscan, _lib  # noqa: E402


def declared_symbols():
    text = (ROOT / "include" / "cuhc.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cuhc_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared_symbols()
    assert len(names) >= 8
    lib = ctypes.CDLL(str(ROOT / "cuhc_amd" / "libcuhc.so"))
    for name in names:
        assert hasattr(lib, name), name
    # the eight reference entry points are among them
    for name in ["cuhc_pairwise_distances", "cuhc_linkage", "cuhc_ward_rnn_with_options",
                 "cuhc_hdbscan_tree", "cuhc_condense_tree", "cuhc_set_pruning",
                 "cuhc_prune_stats", "cuhc_error_string"]:
        assert name in names


def test_error_strings_and_codes():
    assert _lib.lib.cuhc_error_string(-1).decode() == "invalid metric id"
    assert _lib.lib.cuhc_error_string(-2).decode() == "array dimensions must be positive"
    assert _lib.lib.cuhc_error_string(-3).decode() == "invalid linkage id"
    assert _lib.lib.cuhc_error_string(-4).decode() == "invalid algorithm id for this linkage"
    assert "NaN" in _lib.lib.cuhc_error_string(-5).decode()
    assert "rescale" in _lib.lib.cuhc_error_string(-6).decode()
    with pytest.raises(RuntimeError, match=r"cuhc: invalid metric id \(code -1\)"):
        _lib.check(-1)


def test_argument_codes_before_any_gpu_work():
    f = np.zeros((4, 2), np.float32)
    tree = np.zeros((3, 4))
    L = _lib.lib
    assert L.cuhc_linkage(_lib.float_ptr(f), 0, 2, 0, 2.0, 0, 1, _lib.double_ptr(tree)) == -2
    assert L.cuhc_linkage(_lib.float_ptr(f), 4, 2, 0, 2.0, 7, 0, _lib.double_ptr(tree)) == -3
    assert L.cuhc_linkage(_lib.float_ptr(f), 4, 2, 0, 2.0, 1, 1, _lib.double_ptr(tree)) == -4
    assert L.cuhc_linkage(_lib.float_ptr(f), 4, 2, 5, 2.0, 3, 3, _lib.double_ptr(tree)) == -4
    assert L.cuhc_linkage(_lib.float_ptr(f), 4, 2, 9, 2.0, 0, 0, _lib.double_ptr(tree)) == -1
    core = np.zeros(4, np.float32)
    assert L.cuhc_hdbscan_tree(_lib.float_ptr(f), 4, 2, 0, 2.0, 5, 1.0, 0,
                               _lib.double_ptr(tree), _lib.float_ptr(core)) == -2
    assert L.cuhc_hdbscan_tree(_lib.float_ptr(f), 4, 2, 0, 2.0, 2, 1.0, 4,
                               _lib.double_ptr(tree), _lib.float_ptr(core)) == -4
    assert L.cuhc_set_pruning(3) == -4
    assert L.cuhc_set_precision(3) == -4
    assert L.cuhc_set_pruning(0) == 0
    assert _lib.prune_stats() == (0, 0)
    with pytest.raises(KeyError):
        _lib.set_pruning("sometimes")


@pytest.mark.parametrize("mcs", [5, 25])
def test_condense_matches_sklearn_golden(mcs):
    ref = GOLDEN[f"condensed_{mcs}"]
    for parents, children, lambdas, sizes in [_hdbscan.condense_tree(GOLDEN["extract_Z"], mcs),
                                              _hdbscan.condense_tree_reference(
                                                  GOLDEN["extract_Z"], mcs)]:
        np.testing.assert_array_equal(parents, ref[:, 0])
        np.testing.assert_array_equal(children, ref[:, 1])
        np.testing.assert_array_equal(lambdas, ref[:, 2])
        np.testing.assert_array_equal(sizes, ref[:, 3])


def test_condense_capacity_protocol():
    Z = np.ascontiguousarray(GOLDEN["extract_Z"])
    n = len(Z) + 1
    buf = [np.empty(4, np.int64), np.empty(4, np.int64), np.empty(4), np.empty(4, np.int64)]
    rows = _lib.lib.cuhc_condense_tree(_lib.double_ptr(Z), n, 5, 4, _lib.int64_ptr(buf[0]),
                                       _lib.int64_ptr(buf[1]), _lib.double_ptr(buf[2]),
                                       _lib.int64_ptr(buf[3]))
    assert rows == -len(GOLDEN["condensed_5"])


@pytest.mark.parametrize("case, params", list(enumerate([
    dict(min_cluster_size=10), dict(min_cluster_size=10, cluster_selection_method="leaf"),
    dict(min_cluster_size=25, allow_single_cluster=True),
    dict(min_cluster_size=10, max_cluster_size=60),
    dict(min_cluster_size=10, cluster_selection_epsilon=0.5)])))
def test_extraction_matches_sklearn_golden(case, params):
    labels, probs = _hdbscan.tree_to_labels(GOLDEN["extract_Z"], **params)
    np.testing.assert_array_equal(labels, GOLDEN[f"extract_{case}_labels"])
    np.testing.assert_allclose(probs, GOLDEN[f"extract_{case}_probs"], rtol=0, atol=1e-12)


def small_condensed_tree():
    """Reference tests/test_hdbscan.py:344-357: 10 points, root 10 splits at
    distance 3.0 into 11 (points 0-5) and 12 (6-9); 11 splits at 0.4 into 13
    (0-2) and 14 (3-5); points fall out of their leaf at 0.2."""
    rows = [(10, 11, 1 / 3.0, 6), (10, 12, 1 / 3.0, 4), (11, 13, 1 / 0.4, 3), (11, 14, 1 / 0.4, 3)]
    rows += [(13, p, 5.0, 1) for p in range(0, 3)]
    rows += [(14, p, 5.0, 1) for p in range(3, 6)]
    rows += [(12, p, 5.0, 1) for p in range(6, 10)]
    parents, children, lambdas, sizes = map(np.array, zip(*rows))
    return parents, children, lambdas.astype(np.float64), sizes


@pytest.mark.parametrize("epsilon, allow_single, expected", [
    (0.0, False, {12, 13, 14}), (0.3, False, {12, 13, 14}), (0.5, False, {11, 12}),
    (5.0, False, {11, 12}), (5.0, True, {10})])
def test_leaf_selection_with_epsilon_known_answers(epsilon, allow_single, expected):
    # Expected sets are the reference's hand-worked answers (tests/test_hdbscan.py:365-371).
    condensed = small_condensed_tree()
    clusters = _hdbscan.select_clusters(condensed, _hdbscan.compute_stability(condensed), "leaf",
                                        allow_single, epsilon)
    assert clusters == expected
    labels = _hdbscan.label_points(condensed, clusters, allow_single, epsilon)
    assert (labels >= 0).all()


def test_fcluster_and_cut_tree_match_scipy():
    from scipy.cluster.hierarchy import fcluster as scipy_fcluster
    from _helpers import assert_same_partition
    Z = GOLDEN["link_average_euclidean"]
    for t in [1, 2, 7, 50, 400]:
        assert_same_partition(cuhc.fcluster(Z, t), scipy_fcluster(Z, t, "maxclust"))
    for q in [0.0, 0.5, 0.9, 1.0]:
        t = float(np.quantile(Z[:, 2], q))
        assert_same_partition(cuhc.fcluster(Z, t, "distance"), scipy_fcluster(Z, t, "distance"))
    # reference known answer: _linkage.py:201-204
    Zk = np.array([[0, 1, 1, 2], [2, 3, 2, 2], [4, 5, 4, 4]], dtype=np.float64)
    np.testing.assert_array_equal(cuhc.fcluster(Zk, 2), [1, 1, 2, 2])


@pytest.mark.parametrize("call, message", [
    (lambda: cuhc.pairwise_distances(np.zeros((3, 2)), metric="nope"), "unknown metric"),
    (lambda: cuhc.pairwise_distances(np.zeros((3, 2)), np.zeros((3, 4))), "same number of features"),
    (lambda: cuhc.pairwise_distances(np.zeros(3)), "must be 2D"),
    (lambda: cuhc.pairwise_distances(np.zeros((0, 3))), "non-empty"),
    (lambda: cuhc.pairwise_distances(np.zeros((3, 2)), metric="minkowski", p=0), "positive"),
    (lambda: cuhc.pairwise_distances(np.array([[0.0, np.nan]])), "NaN or infinity"),
    (lambda: cuhc.linkage(np.zeros((5, 2)), method="median-of-means"), "method"),
    (lambda: cuhc.linkage(np.zeros((5, 2)), method="ward", metric="cosine"), "euclidean"),
    (lambda: cuhc.linkage(np.zeros((5, 2)), method="complete", algorithm="mst"), "supports linkage"),
    (lambda: cuhc.linkage(np.zeros((5, 2)), algorithm="fastest"), "unknown algorithm"),
    (lambda: cuhc.fcluster(np.zeros((3, 4)), 0), "t >= 1"),
    (lambda: cuhc.fcluster(np.zeros((3, 4)), 3, criterion="inconsistent"), "criterion"),
    (lambda: cuhc.HDBSCAN(min_cluster_size=1).fit(np.zeros((5, 2))), "min_cluster_size"),
    (lambda: cuhc.HDBSCAN(min_samples=0).fit(np.zeros((5, 2))), "min_samples"),
    (lambda: cuhc.HDBSCAN(min_samples=1000).fit(np.zeros((5, 2))), "at most the number"),
    (lambda: cuhc.HDBSCAN(alpha=0.0).fit(np.zeros((5, 2))), "alpha"),
    (lambda: cuhc.HDBSCAN(cluster_selection_epsilon=-1.0).fit(np.zeros((5, 2))), "epsilon"),
    (lambda: cuhc.HDBSCAN(cluster_selection_method="biggest").fit(np.zeros((5, 2))),
     "cluster_selection_method"),
    (lambda: cuhc.HDBSCAN(metric="hamming").fit(np.zeros((5, 2))), "unknown metric"),
    (lambda: cuhc.AgglomerativeClustering(n_clusters=3, distance_threshold=1.0).fit(
        np.zeros((5, 2))), "n_clusters"),
    (lambda: cuhc.AgglomerativeClustering(linkage="ward", metric="cosine").fit(np.zeros((5, 2))),
     "euclidean"),
])
def test_rejects_bad_input_before_the_gpu(call, message):
    with pytest.raises(ValueError, match=message):
        call()


def test_rejects_float32_overflow():
    X = (1.01 ** np.arange(200, dtype=np.float64) * 1e20)[:, None].astype(np.float32)
    for call in [lambda: cuhc.linkage(X), lambda: cuhc.pairwise_distances(X),
                 lambda: cuhc.HDBSCAN().fit(X), lambda: cuhc.linkage(X, method="ward")]:
        with pytest.raises(ValueError, match="too large for float32"):
            call()
    # Ward's size factor (up to n) is included in the bound.
    Y = np.array([[0.0], [1e18]] * 500, dtype=np.float32)
    with pytest.raises(ValueError, match="too large for float32"):
        cuhc.linkage(Y, method="ward")


def test_estimators_are_sklearn_estimators():
    from sklearn.base import ClusterMixin, clone
    from sklearn.cluster import HDBSCAN as SkHDBSCAN
    from sklearn.cluster import AgglomerativeClustering as SkAgg
    for ours, ref, names in [
        (cuhc.HDBSCAN(), SkHDBSCAN(), ["min_cluster_size", "min_samples",
                                       "cluster_selection_epsilon", "max_cluster_size", "metric",
                                       "alpha", "cluster_selection_method",
                                       "allow_single_cluster"]),
        (cuhc.AgglomerativeClustering(), SkAgg(), ["n_clusters", "metric", "linkage",
                                                   "distance_threshold", "compute_distances"])]:
        assert isinstance(ours, ClusterMixin)
        assert clone(ours).get_params() == ours.get_params()
        for name in names:
            assert getattr(ours, name) == getattr(ref, name), name


def _sklearn_tree_to_labels(Z, **params):
    from sklearn.cluster._hdbscan._linkage import HIERARCHY_dtype
    from sklearn.cluster._hdbscan._tree import tree_to_labels as sk_tree_to_labels
    hierarchy = np.empty(len(Z), dtype=HIERARCHY_dtype)
    hierarchy["left_node"] = Z[:, 0]
    hierarchy["right_node"] = Z[:, 1]
    hierarchy["value"] = Z[:, 2]
    hierarchy["cluster_size"] = Z[:, 3]
    import warnings
    with warnings.catch_warnings():  # scikit-learn's own epsilon search converts 1-element arrays
        warnings.simplefilter("ignore", DeprecationWarning)
        return sk_tree_to_labels(hierarchy, **params)


def _blob_tree(seed, n):
    # Single-linkage trees of blobs plus noise with ties (rounded coordinates).
    from scipy.cluster.hierarchy import linkage as scipy_linkage
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-8, 8, (rng.integers(2, 7), 2))
    X = np.vstack([c + rng.normal(0, rng.uniform(0.2, 1.2), (n // len(centres), 2))
                   for c in centres] + [rng.uniform(-10, 10, (n // 10, 2))])
    return scipy_linkage(np.round(X, 1), "single")


@pytest.mark.parametrize("seed", range(6))
def test_extraction_matches_sklearn_on_random_trees(seed):
    # The C++ extraction against scikit-learn's own tree_to_labels, every
    # selection option, on trees with tied heights.
    Z = _blob_tree(seed, 400 + 150 * seed)
    options = [dict(min_cluster_size=m, cluster_selection_method=meth, allow_single_cluster=a,
                    cluster_selection_epsilon=e, max_cluster_size=top)
               for m in (3, 12) for meth in ("eom", "leaf") for a in (False, True)
               for e in (0.0, 0.35) for top in (None, 80)]
    for params in options:
        ours_l, ours_p = _hdbscan.tree_to_labels(Z, **params)
        ref_l, ref_p = _sklearn_tree_to_labels(Z, **params)
        np.testing.assert_array_equal(ours_l, ref_l, err_msg=str(params))
        np.testing.assert_allclose(ours_p, ref_p, rtol=0, atol=1e-12, err_msg=str(params))


@pytest.mark.parametrize("seed", range(4))
def test_condensed_tree_two_methods_agree_with_sklearn(seed):
    from sklearn.cluster._hdbscan._linkage import HIERARCHY_dtype
    from sklearn.cluster._hdbscan._tree import _condense_tree
    Z = _blob_tree(10 + seed, 700)
    hierarchy = np.empty(len(Z), dtype=HIERARCHY_dtype)
    for i, name in enumerate(["left_node", "right_node", "value", "cluster_size"]):
        hierarchy[name] = Z[:, i]
    for mcs in (2, 5, 40):
        ref = _condense_tree(hierarchy, mcs)
        for ours in (_hdbscan.condense_tree(Z, mcs), _hdbscan.condense_tree_reference(Z, mcs)):
            for col, name in zip(ours, ["parent", "child", "value", "cluster_size"]):
                np.testing.assert_array_equal(col, ref[name])
