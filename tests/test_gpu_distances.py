"""GPU pairwise distances: bit-identical to the oracle's FP32 folds, and within
float32 tolerance of sklearn's float64 (reference tests/test_pairwise_distances.py)."""

import numpy as np
import pytest
from sklearn.metrics import pairwise_distances as sk_pairwise

import cuhc_amd as cuhc
import oracle

pytestmark = pytest.mark.gpu

METRICS = [("euclidean", {}), ("sqeuclidean", {}), ("manhattan", {}), ("chebyshev", {}),
           ("minkowski", {"p": 3.0}), ("minkowski", {"p": 1.5}), ("cosine", {})]
# Tile edges: 128 x 128 (64 x 64 for cosine) tiles, 8- and 32-feature LDS fills.
SHAPES = [(1, 1), (7, 3), (64, 8), (129, 9), (300, 64), (257, 100)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"{s[0]}x{s[1]}")
@pytest.mark.parametrize("metric, kw", METRICS, ids=lambda m: str(m))
def test_matches_oracle_and_sklearn(metric, kw, shape):
    X = np.random.default_rng(0).normal(size=shape).astype(np.float32)
    ours = cuhc.pairwise_distances(X, metric=metric, **kw)
    assert ours.dtype == np.float32 and ours.shape == (shape[0], shape[0])
    ref = oracle.pairwise(X, metric=metric, p=kw.get("p", 2.0))
    if metric == "minkowski":  # powf differs between libraries
        np.testing.assert_allclose(ours, ref, rtol=2e-6, atol=1e-6)
    else:
        np.testing.assert_array_equal(ours, ref)
    np.testing.assert_allclose(ours, sk_pairwise(X.astype(np.float64), metric=metric, **kw),
                               rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("metric, kw", METRICS, ids=lambda m: str(m))
def test_two_inputs(metric, kw):
    rng = np.random.default_rng(1)
    X = rng.normal(size=(150, 20)).astype(np.float32)
    Y = rng.normal(loc=0.5, size=(333, 20)).astype(np.float32)
    ours = cuhc.pairwise_distances(X, Y, metric=metric, **kw)
    assert ours.shape == (150, 333)
    np.testing.assert_allclose(ours, sk_pairwise(X.astype(np.float64), Y.astype(np.float64),
                                                 metric=metric, **kw), rtol=1e-4, atol=1e-4)


def test_cosine_zero_vector_and_large_values():
    X = np.array([[0.0, 0.0, 0.0], [1.0, 2.0, 3.0]], dtype=np.float32)
    Y = np.array([[0.0, 0.0, 0.0], [-1.0, 0.5, 2.0]], dtype=np.float32)
    np.testing.assert_allclose(cuhc.pairwise_distances(X, Y, metric="cosine"),
                               sk_pairwise(X.astype(np.float64), Y.astype(np.float64),
                                           metric="cosine"), atol=1e-4)
    Z = np.random.default_rng(3).normal(scale=1e4, size=(100, 32)).astype(np.float32)
    np.testing.assert_array_equal(cuhc.pairwise_distances(Z), oracle.pairwise(Z))


def test_symmetric_with_zero_diagonal_and_aliases():
    X = np.random.default_rng(2).normal(size=(200, 17)).astype(np.float32)
    D = cuhc.pairwise_distances(X)
    np.testing.assert_array_equal(D, D.T)
    np.testing.assert_array_equal(np.diag(D), 0.0)
    np.testing.assert_array_equal(cuhc.pairwise_distances(X, metric="l2"), D)
    np.testing.assert_array_equal(cuhc.pairwise_distances(X, metric="l1"),
                                  cuhc.pairwise_distances(X, metric="manhattan"))
    np.testing.assert_array_equal(cuhc.pairwise_distances(X.tolist()), D)
    np.testing.assert_array_equal(cuhc.pairwise_distances(np.asfortranarray(X)), D)
