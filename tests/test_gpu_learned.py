"""The learned-radius core distances and the radius network's own kernels.

The network's forward pass and gradients (mlp.hip) are checked against the FP64
CPU oracle (whose gradient is pinned by finite differences in test_oracle.py);
the first Adam step against its closed form (every parameter moves by lr in the
direction opposite its gradient sign); the learned core distances bit for bit
against the oracle and the exact methods; and the trained network against the
untrained one on the same data (fewer exact fallbacks and gathered candidates).
"""

import numpy as np
import pytest

import oracle
from _helpers import blobs
from cuhc_amd import _lib
from cuhc_amd._hdbscan import hdbscan_tree

pytestmark = pytest.mark.gpu


def random_net(seed, scale=0.4):
    rng = np.random.default_rng(seed)
    return rng.normal(scale=scale, size=_lib.MLP_PARAMS).astype(np.float32)


@pytest.mark.parametrize("m", [1, 5, 257, 4096])
def test_forward_matches_oracle(m):
    p = random_net(m)
    X = np.random.default_rng(m + 1).normal(size=(m, 4)).astype(np.float32)
    y = _lib.mlp_forward(p, X)
    np.testing.assert_allclose(y, oracle.mlp_forward(p, X), rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("m", [1, 37, 4096])
def test_gradient_matches_oracle(m):
    # FP32 activations against the FP64 oracle: relative to the gradient's
    # scale, within 1e-4.
    p = random_net(100 + m)
    rng = np.random.default_rng(m)
    X = rng.normal(size=(m, 4)).astype(np.float32)
    t = rng.normal(size=m).astype(np.float32)
    g, loss = _lib.mlp_gradient(p, X, t)
    go, lo = oracle.mlp_gradient(p, X, t)
    assert loss == pytest.approx(lo, rel=1e-5)
    scale = np.abs(go).max()
    np.testing.assert_allclose(g, go, rtol=1e-3, atol=1e-4 * scale)
    # every layer's block of parameters is exercised (none identically zero)
    bounds = np.cumsum([0, 128, 32, 1024, 32, 32, 1])
    for a, b in zip(bounds[:-1], bounds[1:]):
        assert np.abs(go[a:b]).max() > 0


def test_first_adam_step_is_signed_lr():
    # Adam's bias-corrected first step is lr * g / (|g| + eps): lr times the
    # sign of the gradient for every parameter whose gradient is not tiny.
    p = random_net(7)
    rng = np.random.default_rng(8)
    X = rng.normal(size=(500, 4)).astype(np.float32)
    t = rng.normal(size=500).astype(np.float32)
    go, _ = oracle.mlp_gradient(p, X, t)
    lr = 0.01
    p1 = _lib.mlp_train(p, X, t, steps=1, lr=lr)
    big = np.abs(go) > 1e-3 * np.abs(go).max()
    step = (p1 - p).astype(np.float64)
    np.testing.assert_allclose(step[big], -lr * np.sign(go[big]), rtol=1e-3, atol=1e-6)
    assert np.all(np.abs(step) <= lr * 1.001)


def test_training_fits_a_smooth_target():
    rng = np.random.default_rng(3)
    X = rng.normal(size=(2048, 4)).astype(np.float32)
    t = (np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2] - 0.3 * X[:, 3]).astype(np.float32)
    p0 = random_net(4, scale=0.2)
    _, loss0 = oracle.mlp_gradient(p0, X, t)
    p = _lib.mlp_train(p0, X, t, steps=400, lr=0.01)
    _, loss = oracle.mlp_gradient(p, X, t)
    assert loss < 0.1 * loss0, (loss0, loss)


def test_mlp_argument_errors():
    with pytest.raises(ValueError):
        _lib.mlp_forward(np.zeros(3, np.float32), np.zeros((2, 4), np.float32))
    p = random_net(1)
    X = np.zeros((0, 4), np.float32)
    with pytest.raises(RuntimeError, match="positive"):
        _lib.mlp_forward(p, X)
    with pytest.raises(RuntimeError, match="positive"):
        _lib.set_learned_training(-1)


def varied_data():
    rng = np.random.default_rng(5)
    yield "blobs_d2", blobs(n_samples=6000, n_features=2, centers=12, seed=1)
    yield "blobs_d8_noise", blobs(n_samples=8000, n_features=8, centers=20, noise_fraction=0.2,
                                  seed=2)
    yield "uniform_d3", rng.uniform(-1, 1, (5000, 3)).astype(np.float32)
    yield "gauss_d24", rng.standard_normal((4000, 24)).astype(np.float32)
    spread = np.concatenate([rng.normal(scale=s, size=(1500, 4)) + 20 * i
                             for i, s in enumerate([0.01, 0.1, 1.0, 10.0])])
    yield "mixed_scales", spread.astype(np.float32)
    yield "duplicates", np.repeat(rng.normal(size=(700, 3)), 6, axis=0).astype(np.float32)
    yield "small_n", rng.normal(size=(130, 5)).astype(np.float32)


DATA = dict(varied_data())


@pytest.mark.parametrize("name", list(DATA))
@pytest.mark.parametrize("k", [1, 5, 40, 64, 100])
def test_learned_core_equals_oracle(name, k):
    X = DATA[name]
    if k > len(X):
        pytest.skip("k > n")
    _lib.learned_stats()
    Z, core = hdbscan_tree(X, k, core_method="learned")
    st = _lib.learned_stats()
    np.testing.assert_array_equal(core, oracle.core_distances(X, k))
    assert st["proposals"] == len(X)
    assert st["hits"] + st["fallbacks"] == len(X)
    Zr, _ = hdbscan_tree(X, k, core_method="radix")
    np.testing.assert_array_equal(Z, Zr)


@pytest.mark.parametrize("metric", ["manhattan", "chebyshev", "minkowski", "sqeuclidean",
                                    "cosine"])
def test_learned_core_other_metrics(metric):
    X = DATA["blobs_d8_noise"][:5000]
    kw = {"p": 3.0} if metric == "minkowski" else {}
    _, core = hdbscan_tree(X, 10, metric=metric, core_method="learned", **kw)
    _, radix = hdbscan_tree(X, 10, metric=metric, core_method="radix", **kw)
    np.testing.assert_array_equal(core, radix)
    expected = oracle.core_distances(X, 10, metric, **kw)
    if metric == "minkowski":  # the device powf is not the host libm's
        np.testing.assert_allclose(core, expected, rtol=1e-6)
    else:
        np.testing.assert_array_equal(core, expected)


@pytest.mark.parametrize("prune", ["never", "always"])
def test_learned_core_any_order(prune):
    X = DATA["blobs_d8_noise"]
    _lib.set_pruning(prune)
    try:
        _, core = hdbscan_tree(X, 7, core_method="learned")
    finally:
        _lib.set_pruning("auto")
    np.testing.assert_array_equal(core, oracle.core_distances(X, 7))


def test_trained_network_gains_over_untrained():
    # 60,000 points whose density varies by three orders of magnitude between
    # clusters: the trained radii hold for far more rows than the untrained
    # network's (whose residual margin must cover its own scatter), and gather
    # far fewer candidates. The core distances are identical either way.
    rng = np.random.default_rng(11)
    parts = [rng.normal(scale=s, size=(10_000, 4)) + rng.uniform(-50, 50, 4)
             for s in [0.01, 0.05, 0.3, 1.0, 3.0]]
    parts.append(rng.uniform(-60, 60, (10_000, 4)))
    X = np.concatenate(parts).astype(np.float32)
    k = 16
    stats = {}
    cores = {}
    try:
        for steps in [300, 0]:
            _lib.set_learned_training(steps)
            _lib.learned_stats()
            _, cores[steps] = hdbscan_tree(X, k, core_method="learned")
            stats[steps] = _lib.learned_stats()
    finally:
        _lib.set_learned_training(300)
    np.testing.assert_array_equal(cores[300], cores[0])
    _, exact = hdbscan_tree(X, k, core_method="radix")
    np.testing.assert_array_equal(cores[300], exact)
    trained, untrained = stats[300], stats[0]
    print("trained", trained, "untrained", untrained)
    assert trained["hits"] >= 0.95 * len(X)
    assert trained["fallbacks"] < 0.5 * untrained["fallbacks"]
    assert trained["candidates"] < untrained["candidates"]


@pytest.mark.parametrize("name", ["blobs_d8_noise", "blobs_d2", "duplicates", "mixed_scales"])
@pytest.mark.parametrize("k", [5, 25])
def test_learned_lists_settle_boruvka_rows(name, k):
    # With pruned passes the learned method keeps its candidate lists, and a
    # Boruvka row whose best listed partner outside its component lies within
    # the radius is answered from the list and closed to the walk; the other
    # rows are searched. The tree must equal the oracle's bit for bit, with
    # both kinds of rows present.
    X = DATA[name]
    _lib.set_pruning("always")
    try:
        _lib.learned_stats()
        Z, core = hdbscan_tree(X, k, core_method="learned")
        st = _lib.learned_stats()
    finally:
        _lib.set_pruning("auto")
    Zo, co = oracle.hdbscan_tree(X, k)
    np.testing.assert_array_equal(core, co)
    np.testing.assert_array_equal(Z, Zo)
    print(name, k, st)
    assert st["settled_rows"] > 0
    assert st["settled_rows"] < st["offered_rows"]


@pytest.mark.parametrize("alpha", [0.7, 1.6])
def test_learned_lists_off_for_scaled_reachability(alpha):
    # alpha != 1: the weight max(core_i, core_j, d / alpha) can fall below the
    # distance, so the lists do not certify rows; the tree is still exact.
    X = DATA["blobs_d8_noise"]
    _lib.set_pruning("always")
    try:
        _lib.learned_stats()
        Z, core = hdbscan_tree(X, 10, alpha=alpha, core_method="learned")
        st = _lib.learned_stats()
    finally:
        _lib.set_pruning("auto")
    Zo, co = oracle.hdbscan_tree(X, 10, alpha=alpha)
    np.testing.assert_array_equal(core, co)
    np.testing.assert_array_equal(Z, Zo)
    assert st["settled_rows"] == 0 and st["offered_rows"] == 0


def test_auto_uses_the_learned_radius_above_k64():
    # For k > 64 on large inputs auto hands the core distances to the learned
    # radius (the alternative is the full-row radix select); the result is the
    # radix select's bit for bit, and it gathers far fewer candidates than the
    # n^2 distances the radix select writes and reads.
    rng = np.random.default_rng(21)
    parts = [rng.normal(scale=s, size=(30_000, 5)) + rng.uniform(-40, 40, 5)
             for s in [0.2, 0.5, 1.0, 2.0]]
    parts.append(rng.uniform(-45, 45, (20_000, 5)))
    X = np.concatenate(parts).astype(np.float32)
    n, k = len(X), 100
    _lib.learned_stats()
    _, core_auto = hdbscan_tree(X, k)
    st = _lib.learned_stats()
    _, core_radix = hdbscan_tree(X, k, core_method="radix")
    np.testing.assert_array_equal(core_auto, core_radix)
    print(st)
    assert st["proposals"] == n
    assert st["candidates"] < 0.05 * n * n
