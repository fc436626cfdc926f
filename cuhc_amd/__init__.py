"""Exact hierarchical clustering and HDBSCAN on one AMD Instinct MI355X (gfx950).

A drop-in for the reference `cuhc` package (`import cuhc_amd as cuhc`): the same
`AgglomerativeClustering`, `HDBSCAN`, `linkage`, `fcluster` and
`pairwise_distances`, backed by hand-written HIP kernels in libcuhc.so.
"""

from ._distances import METRIC_IDS, pairwise_distances
from ._estimators import AgglomerativeClustering
from ._hdbscan import HDBSCAN
from ._linkage import LINKAGE_IDS, fcluster, linkage

__all__ = [
    "AgglomerativeClustering",
    "HDBSCAN",
    "LINKAGE_IDS",
    "METRIC_IDS",
    "fcluster",
    "linkage",
    "pairwise_distances",
]
