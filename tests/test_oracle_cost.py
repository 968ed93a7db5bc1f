"""Pins the oracle's configuration-pair costs and distances (or_config_cost) to facts it does not
restate itself.

* The reference's own tests of these functions (tests/test_config.py): zero distance, equal points,
  translation invariance, one agent (max with w = 0 equals sum), and the two manual comparisons
  that demand equality with np.linalg.norm / np.max per robot summed in Python (:110-126,
  :129-145).
* Exact arithmetic: each fused step (sum_squared += d * d, max + w * sum) equals the correctly
  rounded value of the exact rational result (fractions.Fraction), the square root is Python's
  correctly rounded math.sqrt, and the other steps round separately.  A dropped fusion, a
  reordered sum or a wrong slice fails these.
* NaN rules: np.max propagates a NaN (metric "max" of batch_config_cost); compute_abs_max_reduction
  keeps a NaN only in the first coordinate (abs_val > max_val is false for NaN).
"""
import math
from fractions import Fraction

import numpy as np
import pytest

from helpers import O

DIMS = [[2, 2], [7, 7], [3, 3, 3], [2, 5], [14]]
NUM_POINTS = [1, 10, 100, 1000, 5000]


def _slices(dims):
    return [(sum(dims[:i]), sum(dims[: i + 1])) for i in range(len(dims))]


def _fused(a, b, c):
    """fma(a, b, c) from exact rationals, rounded once (Fraction -> float is correctly rounded)."""
    return float(Fraction(a) * Fraction(b) + Fraction(c))


# ---------------------------------------------------------------------------------------------
# tests/test_config.py, restated against the oracle
@pytest.mark.parametrize("reduction", ["max", "sum"])
@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("num_points", NUM_POINTS)
def test_cost_zero_and_equal_points(reduction, dims, num_points):
    rng = np.random.default_rng(num_points)
    D = sum(dims)
    assert np.all(O.config_cost(np.zeros(D), np.zeros((num_points, D)), _slices(dims), "max", reduction) == 0)
    pt = rng.random(D)
    assert np.all(O.config_cost(pt, np.tile(pt, (num_points, 1)), _slices(dims), "max", reduction) == 0)


@pytest.mark.parametrize("reduction", ["max", "sum"])
@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("num_points", NUM_POINTS)
def test_cost_translation(reduction, dims, num_points):
    rng = np.random.default_rng(7 + num_points)
    D = sum(dims)
    pt, pts = rng.random(D), rng.random((num_points, D))
    a = O.config_cost(pt, pts, _slices(dims), "max", reduction)
    b = O.config_cost(pt + 1, pts + 1, _slices(dims), "max", reduction)
    assert np.allclose(a, b)


@pytest.mark.parametrize("dims", [[2], [6], [14]])
@pytest.mark.parametrize("num_points", NUM_POINTS)
def test_cost_single_agent(dims, num_points):
    rng = np.random.default_rng(11)
    D = sum(dims)
    pt, pts = rng.random(D), rng.random((num_points, D))
    m = O.config_cost(pt, pts, _slices(dims), "max", "max", w=0.0)
    s = O.config_cost(pt, pts, _slices(dims), "max", "sum")
    assert np.array_equal(m, s)


@pytest.mark.parametrize("dims", DIMS)
def test_cost_manual_comparisons(dims):
    """test_config.py:110-145: one point; the per-robot np.linalg.norm (euclidean) or
    np.max(np.abs(.)) (max), summed from 0 in Python, equals the function's value exactly."""
    rng = np.random.default_rng(3)
    D = sum(dims)
    for trial in range(200):
        pt, pts = rng.random(D), rng.random((1, D))
        for metric, per_robot in (("euclidean", np.linalg.norm), ("max", lambda v: np.max(np.abs(v)))):
            got = O.config_cost(pt, pts, _slices(dims), metric, "sum")[0]
            manual = 0
            for s, e in _slices(dims):
                manual += per_robot(pts[0][s:e] - pt[s:e])
            assert got == manual, (metric, trial)


@pytest.mark.parametrize("metric", ["absmax", "euclidean"])
@pytest.mark.parametrize("reduction", ["max_only", "sum"])
@pytest.mark.parametrize("dims", DIMS)
def test_dist_zero_equal_translation(metric, reduction, dims):
    """batch_config_dist's forms (test_config.py:148-185): "max" = absmax over all coordinates,
    "euclidean" = one slice, "sum_euclidean" / "max_euclidean" = per-robot slices."""
    rng = np.random.default_rng(5)
    D = sum(dims)
    sl = [(0, D)] if metric == "absmax" else _slices(dims)
    assert np.all(O.config_cost(np.zeros(D), np.zeros((50, D)), sl, metric, reduction) == 0)
    pt, pts = rng.random(D), rng.random((100, D))
    assert np.all(O.config_cost(pt, np.tile(pt, (50, 1)), sl, metric, reduction) == 0)
    assert np.allclose(O.config_cost(pt, pts, sl, metric, reduction),
                       O.config_cost(pt + 1, pts + 1, sl, metric, reduction))


# ---------------------------------------------------------------------------------------------
# exact arithmetic
def _exact_row(p, q, slices, metric, reduction, w):
    d = [float(a) - float(b) for a, b in zip(p, q)]
    vals = []
    for s, e in slices:
        if metric == "euclidean":
            acc = 0.0
            for k in range(s, e):
                acc = _fused(d[k], d[k], acc)
            vals.append(math.sqrt(acc))
        else:
            vals.append(max(abs(x) for x in d[s:e]))
    if reduction == "sum":
        t = 0.0
        for v in vals:
            t = t + v
        return t
    mx, sm = vals[0], vals[0]
    for v in vals[1:]:
        mx = v if v > mx else mx
        sm = sm + v
    return mx if reduction == "max_only" else _fused(w, sm, mx)


@pytest.mark.parametrize("metric", ["max", "euclidean"])
@pytest.mark.parametrize("reduction", ["max", "sum", "max_only"])
@pytest.mark.parametrize("dims", [[7, 7, 7, 7], [3, 3, 3], [14], [1, 20, 2]])
def test_cost_exact_arithmetic(metric, reduction, dims):
    rng = np.random.default_rng(hash((metric, reduction, len(dims))) & 0xFFFF)
    D = sum(dims)
    # spread magnitudes so that fused and unfused steps differ in many rows
    pts = rng.normal(size=(300, D)) * np.exp2(rng.integers(-20, 20, size=(300, D)))
    pt = rng.normal(size=D)
    w = 0.01
    got = O.config_cost(pt, pts, _slices(dims), metric, reduction, w=w)
    exp = np.array([_exact_row(pt, r, _slices(dims), metric, reduction, w) for r in pts])
    assert np.array_equal(got, exp)


def test_fusion_is_observable():
    """The exact-arithmetic check above can tell fused from unfused steps: on these inputs the two
    differ, so a dropped fma would fail it."""
    rng = np.random.default_rng(0)
    differ = 0
    for row in rng.normal(size=(2000, 7)):
        a = b = 0.0
        for x in row:
            a = _fused(x, x, a)
            b = b + x * x
        differ += a != b
    assert differ > 100
    differ = sum(_fused(0.01, x, y) != y + 0.01 * x for x, y in rng.random((2000, 2)) * 10)
    assert differ > 10


def test_nan_rules():
    nan = float("nan")
    q = np.zeros((4, 3))
    p = np.array([[nan, 1.0, 2.0], [1.0, nan, 2.0], [1.0, 2.0, nan], [3.0, 2.0, 1.0]])
    sl = [(0, 3)]
    # np.max(np.abs(.)) propagates the NaN wherever it is
    got = O.config_cost(p, q, sl, "max", "max_only")
    assert np.isnan(got[:3]).all() and got[3] == 3.0
    assert np.array_equal(np.isnan(got), np.isnan(np.max(np.abs(p - q), axis=1)))
    # compute_abs_max_reduction: a NaN sticks only from the first coordinate
    got = O.config_cost(p, q, sl, "absmax", "max_only")
    assert np.isnan(got[0]) and got[1] == 2.0 and got[2] == 2.0 and got[3] == 3.0


def test_bad_arguments():
    with pytest.raises(ValueError):
        O.config_cost(np.zeros(4), np.zeros((2, 4)), [(0, 5)])
    with pytest.raises(ValueError):
        O.config_cost(np.zeros(4), np.zeros((2, 4)), [(2, 2)], "max")
    # an empty slice is a zero norm
    assert np.all(O.config_cost(np.ones(4), np.zeros((2, 4)), [(2, 2)], "euclidean", "sum") == 0)
