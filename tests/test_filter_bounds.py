"""The argmin pair filter's bounds, checked against correctly rounded FP32.

The walk, fused-core and radius kernels skip a pair when its raw accumulator
exceeds accBound(need) (cuhc_amd/csrc/common.hpp). That is exact only if every
accumulator a with finish(a) <= need lies within the bound, i.e. the first
float above the bound already finishes above need. These tests evaluate the
bound formulas in float32 (left to right, no fused operations, as the library
is built with -ffp-contract=off) and check that property with numpy's
correctly rounded float32 sqrt and division over random and edge values,
denormals and overflow included. Runs on the CPU.
"""

import numpy as np

F32 = np.float32
FLT_MIN = F32(np.finfo(np.float32).tiny)


def euclidean_acc_bound(x):
    with np.errstate(over="ignore"):
        return (x * x) * F32(1.0 + 2.0 ** -20) + FLT_MIN


def mreach_scaled_dist_bound(w, alpha):
    with np.errstate(over="ignore"):
        return (w * alpha) * F32(1.0 + 2.0 ** -21) + FLT_MIN


def sample_floats(rng, count):
    """Non-negative float32 values spread over every binade, plus edge cases."""
    bits = rng.integers(0, 0x7F800000, size=count, dtype=np.uint32)
    special = np.array([0.0, FLT_MIN, 1e-45, 1e-40, 1.0, 2.0, 3.0, 1e19, 1.8e19, 1.9e19,
                        3.4e38, np.finfo(np.float32).max], dtype=np.float32)
    return np.concatenate([bits.view(np.float32), special])


def test_euclidean_bound_covers_every_accumulator_within_need():
    rng = np.random.default_rng(11)
    need = sample_floats(rng, 400_000)
    bound = euclidean_acc_bound(need)
    assert bound.dtype == np.float32
    assert np.all(bound >= 0)
    above = np.nextafter(bound, F32(np.inf))
    finite = np.isfinite(above)
    # The first float above the bound finishes (sqrt) above the need ...
    assert np.all(np.sqrt(above[finite]) > need[finite])
    # ... and the bound is tight enough to filter: within 2^-18 of need^2
    # (away from the FLT_MIN term and from overflow).
    normal = (need > F32(1e-15)) & (need < F32(1e18))
    ratio = bound[normal].astype(np.float64) / (need[normal].astype(np.float64) ** 2)
    assert np.all(ratio < 1.0 + 2.0 ** -18)


def test_euclidean_bound_brute_force_small_binades():
    # Every float accumulator in a few binades: sqrtf(a) <= x implies a <= bound(x).
    for lo in (F32(1e-44), F32(1e-38), F32(0.5), F32(7.0)):
        a = np.arange(lo.view(np.uint32), lo.view(np.uint32) + 200_000,
                      dtype=np.uint32).view(np.float32)
        x = np.sqrt(a)  # the need equal to each accumulator's own finish
        assert np.all(a <= euclidean_acc_bound(x))


def test_scaled_mreach_bound_covers_every_distance_within_weight():
    rng = np.random.default_rng(12)
    w = sample_floats(rng, 200_000)
    for alpha in (F32(0.3), F32(0.5), F32(1.7), F32(2.0), F32(1e-3), F32(1e3)):
        bound = mreach_scaled_dist_bound(w, alpha)
        above = np.nextafter(bound, F32(np.inf))
        finite = np.isfinite(above)
        with np.errstate(over="ignore"):
            assert np.all(above[finite] / alpha > w[finite])
