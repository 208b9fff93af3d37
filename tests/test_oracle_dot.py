"""The oracle's scalar product (the restated BiCGStab's dot) is the exact sum
of the products rounded once to the nearest double.  Pinned against Python's
math.fsum, which returns the correctly rounded sum of its arguments, and
against exact rational arithmetic where fsum overflows; the products are
formed in Python with the same IEEE double multiply the restatement uses."""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle as orc


def _exact(a):
    """Correctly rounded sum of the doubles in a (ties to even), +0 for a zero sum."""
    vals = [float(v) for v in a]
    if any(math.isnan(v) for v in vals):
        return float("nan")
    pinf = any(v == math.inf for v in vals)
    ninf = any(v == -math.inf for v in vals)
    if pinf and ninf:
        return float("nan")
    if pinf or ninf:
        return math.inf if pinf else -math.inf
    try:
        return math.fsum(vals) + 0.0
    except OverflowError:
        F = sum(Fraction(v) for v in vals)
        limit = Fraction(2 ** 1024 - 2 ** 970)
        if abs(F) >= limit:
            return math.inf if F > 0 else -math.inf
        return float(F)


def _same(a, b):
    if math.isnan(a) or math.isnan(b):
        return math.isnan(a) and math.isnan(b)
    return a == b and math.copysign(1.0, a) == math.copysign(1.0, b)


def _check(x, y):
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    with np.errstate(over="ignore", invalid="ignore", under="ignore"):
        prods = x * y
    got = orc.dot(x, y)
    want = _exact(prods)
    assert _same(got, want), (got, want)


def _cases(rng, count):
    for t in range(count):
        n = int(rng.integers(1, 300))
        kind = t % 7
        if kind == 0:
            x, y = rng.standard_normal(n), rng.standard_normal(n)
        elif kind == 1:      # products over the whole exponent range
            x = rng.standard_normal(n) * 2.0 ** rng.integers(-600, 500, n)
            y = rng.standard_normal(n) * 2.0 ** rng.integers(-470, 500, n)
        elif kind == 2:      # cancellation down to a tiny remainder
            b = rng.standard_normal(n) * 10.0 ** rng.integers(-30, 30, n)
            x = np.concatenate([b, -b, rng.standard_normal(3) * 1e-300])
            y = np.ones_like(x)
            p = rng.permutation(x.size)
            x, y = x[p], y[p]
        elif kind == 3:      # subnormal products
            x = rng.integers(-2 ** 40, 2 ** 40, n) * 2.0 ** -1074
            y = rng.integers(1, 4096, n).astype(np.float64)
        elif kind == 4:      # near overflow
            x = rng.standard_normal(n) * 1e300
            y = rng.standard_normal(n) * 1e7
        elif kind == 5:      # mixed scales of a black-oil residual (p, S, Rs equations)
            x = rng.standard_normal(3 * n) * np.tile([1e5, 1e-3, 1e2], n)
            y = rng.standard_normal(3 * n) * np.tile([1e-4, 1e3, 1e-9], n)
        else:                # large values cancelling above a small tail
            v = rng.standard_normal(n)
            x = np.concatenate([v * 2.0 ** 60, -v * 2.0 ** 60, v * 2.0 ** -60])
            y = np.ones_like(x)
        yield x, y


def test_dot_is_the_correctly_rounded_sum():
    rng = np.random.default_rng(2024)
    for x, y in _cases(rng, 1400):
        _check(x, y)


@pytest.mark.parametrize("vals", [
    [1.0, 2.0 ** -53],                         # tie: stays at the even neighbour 1.0
    [1.0 + 2.0 ** -52, 2.0 ** -53],            # tie: rounds up to the even neighbour
    [1.0, 2.0 ** -53, 2.0 ** -200],            # just above the tie: up
    [1.0, 2.0 ** -53, -(2.0 ** -300)],         # just below the tie: down
    [-1.0, -(2.0 ** -53)],                     # negative tie
    [0.0, -0.0],                               # zero sum is +0
    [-0.0],
    [1.0, -1.0],
    [2.0 ** -1074, -(2.0 ** -1074), 2.0 ** -1074],
    [2.0 ** -1022, -(2.0 ** -1074)],           # normal minus a subnormal unit
    [1.7976931348623157e308, 1e292],           # rounds to infinity
    [2.0 ** 1023, 2.0 ** 1023, -(2.0 ** 1023)],  # intermediate overflow, finite sum
    [1.0, math.inf],
    [-math.inf, 3.0],
    [math.inf, -math.inf],
    [1.0, math.nan],
])
def test_dot_special_cases(vals):
    _check(vals, np.ones(len(vals)))


def test_dot_is_order_independent_where_the_sequential_sum_is_not():
    """A planted check that the pins discriminate: the sequential loop of the
    reference gives two different results for two orders of the same
    products, the restated dot gives the correctly rounded sum for both."""
    x = np.array([1.0, 1e-16, 1e-16, -1.0, 1e-32])
    seq = lambda v: sum(v.tolist(), 0.0)   # noqa: E731
    assert seq(x) != seq(x[::-1])
    assert orc.dot(x, np.ones(5)) == orc.dot(x[::-1], np.ones(5)) == math.fsum(x.tolist())


def test_dot_of_long_vectors():
    rng = np.random.default_rng(7)
    x = rng.standard_normal(300_000) * np.tile([1e5, 1e-3, 1e2], 100_000)
    y = rng.standard_normal(300_000)
    assert orc.dot(x, y) == math.fsum((x * y).tolist())
