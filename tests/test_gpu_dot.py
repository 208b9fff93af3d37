"""The device scalar product (BiCGStab's dot) against the oracle's and
against math.fsum: the exact sum of the products rounded once, bit for bit,
whatever the length, the magnitudes or the cancellation."""
import math

import numpy as np
import pytest

import omodels_amd as om
import oracle as orc
from test_oracle_dot import _cases, _exact, _same

pytestmark = pytest.mark.gpu


def _check(x, y):
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    got = om.scalar_product(x, y)
    want = orc.dot(x, y)
    with np.errstate(over="ignore", invalid="ignore", under="ignore"):
        ref = _exact(x * y)
    assert _same(want, ref), (want, ref)
    assert _same(got, want), (got, want)


def test_device_dot_random_cases():
    rng = np.random.default_rng(99)
    for x, y in _cases(rng, 210):
        _check(x, y)


@pytest.mark.parametrize("vals", [
    [1.0, 2.0 ** -53], [1.0 + 2.0 ** -52, 2.0 ** -53], [1.0, 2.0 ** -53, 2.0 ** -200],
    [1.0, 2.0 ** -53, -(2.0 ** -300)], [-1.0, -(2.0 ** -53)], [0.0, -0.0], [-0.0], [1.0, -1.0],
    [2.0 ** -1074, -(2.0 ** -1074), 2.0 ** -1074], [2.0 ** -1022, -(2.0 ** -1074)],
    [1.7976931348623157e308, 1e292], [2.0 ** 1023, 2.0 ** 1023, -(2.0 ** 1023)],
    [1.0, math.inf], [-math.inf, 3.0], [math.inf, -math.inf], [1.0, math.nan], [],
])
def test_device_dot_special_cases(vals):
    _check(vals, np.ones(len(vals)))


@pytest.mark.parametrize("m", [1, 63, 64, 65, 511, 512, 513, 2048 * 8 + 7, 1_500_000, 6_000_000])
def test_device_dot_lengths(m):
    """Tails of every launch shape, up to the 2M-cell workload's 6M entries:
    mixed scales of the three black-oil equations and a cancelling pair."""
    rng = np.random.default_rng(m)
    x = rng.standard_normal(m) * np.resize([1e5, 1e-3, 1e2], m)
    y = rng.standard_normal(m) * np.resize([1e-4, 1e3, 1e-9], m)
    if m > 2:
        x[0], x[-1], y[0], y[-1] = 1e200, -1e200, 1.0, 1.0
    got = om.scalar_product(x, y)
    assert got == orc.dot(x, y) == math.fsum((x * y).tolist())


def test_device_dot_is_order_independent():
    rng = np.random.default_rng(5)
    x = rng.standard_normal(300_001) * 10.0 ** rng.integers(-20, 20, 300_001)
    y = rng.standard_normal(300_001)
    p = rng.permutation(x.size)
    assert om.scalar_product(x, y) == om.scalar_product(x[p], y[p]) == math.fsum((x * y).tolist())
