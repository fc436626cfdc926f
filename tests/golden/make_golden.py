"""Writes tests/golden/golden.npz: small seeded inputs with scipy/sklearn outputs.

These third-party results are the golden vectors that pin the oracle (and, on
the GPU, the library): scipy's linkage on float64 copies of float32 inputs,
sklearn's pairwise distances, HDBSCAN and its private tree extraction. The
reference repository itself is not imported or run.

    python tests/golden/make_golden.py
"""

import pathlib
import sys

import numpy as np
from scipy.cluster.hierarchy import linkage
from scipy.spatial.distance import pdist
from sklearn.cluster import HDBSCAN
from sklearn.cluster._hdbscan._linkage import HIERARCHY_dtype
from sklearn.cluster._hdbscan._tree import _condense_tree, tree_to_labels
from sklearn.metrics import pairwise_distances

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
from _helpers import blobs  # noqa: E402


def structured(Z):
    tree = np.empty(len(Z), dtype=HIERARCHY_dtype)
    tree["left_node"], tree["right_node"] = Z[:, 0], Z[:, 1]
    tree["value"], tree["cluster_size"] = Z[:, 2], Z[:, 3]
    return tree


def main():
    out = {}
    # Linkage trees (scipy, float64) on tie-free blobs.
    X = blobs(n_samples=400, n_features=8, seed=400)
    out["link_X"] = X
    X64 = X.astype(np.float64)
    for method, metric in [("single", "euclidean"), ("single", "cosine"),
                           ("complete", "manhattan"), ("average", "cosine"),
                           ("average", "euclidean"), ("ward", "euclidean")]:
        if method == "ward":
            Z = linkage(X64, "ward")
        else:
            Z = linkage(pdist(X64, metric={"manhattan": "cityblock"}.get(metric, metric)), method)
        out[f"link_{method}_{metric}"] = Z
    # Pairwise distances (sklearn, float64).
    P = np.random.default_rng(7).normal(size=(129, 9)).astype(np.float32)
    out["pair_X"] = P
    for metric in ["euclidean", "sqeuclidean", "manhattan", "chebyshev", "cosine"]:
        out[f"pair_{metric}"] = pairwise_distances(P.astype(np.float64), metric=metric)
    out["pair_minkowski3"] = pairwise_distances(P.astype(np.float64), metric="minkowski", p=3.0)
    # Core distances: k-th smallest of the float64 matrix with diagonal 0.
    C = blobs(n_samples=300, n_features=5, seed=5)
    out["core_X"] = C
    D = pairwise_distances(C.astype(np.float64))
    np.fill_diagonal(D, 0.0)
    for k in [1, 2, 7, 40]:
        out[f"core_k{k}"] = np.sort(D, axis=1)[:, k - 1]
    # HDBSCAN end to end (min_samples=1: mutual reachability = distance, no ties).
    H = blobs(n_samples=600, n_features=4, centers=6, noise_fraction=0.1, seed=20)
    out["hdb_X"] = H
    model = HDBSCAN(min_cluster_size=20, min_samples=1, copy=True).fit(H.astype(np.float64))
    out["hdb_labels"] = model.labels_
    out["hdb_probabilities"] = model.probabilities_
    # Condensed tree and extraction on a fixed single-linkage tree.
    Zs = linkage(H.astype(np.float64), "single")
    out["extract_Z"] = Zs
    for mcs in [5, 25]:
        ct = _condense_tree(structured(Zs), mcs)
        out[f"condensed_{mcs}"] = np.column_stack([ct["parent"], ct["child"], ct["value"],
                                                   ct["cluster_size"]])
    cases = [dict(min_cluster_size=10), dict(min_cluster_size=10, cluster_selection_method="leaf"),
             dict(min_cluster_size=25, allow_single_cluster=True),
             dict(min_cluster_size=10, max_cluster_size=60),
             dict(min_cluster_size=10, cluster_selection_epsilon=0.5)]
    for i, params in enumerate(cases):
        labels, probs = tree_to_labels(structured(Zs), **params)
        out[f"extract_{i}_labels"] = labels
        out[f"extract_{i}_probs"] = probs
    np.savez_compressed(HERE / "golden.npz", **out)
    print(f"wrote {HERE / 'golden.npz'} ({len(out)} arrays)")


if __name__ == "__main__":
    main()
