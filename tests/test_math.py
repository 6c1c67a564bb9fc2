"""Accuracy of the shared elementary functions (include/gearhip_math.h),
which the expression-program kernel and the oracle both use: errors in
units in the last place against 80-bit mpmath values on seeded sweeps."""
import math

import mpmath
import numpy as np
import pytest

from oracle import oracle as O

mpmath.mp.prec = 80


def _ulp_err(name, xs, ref):
    ys = O.math_fn(name, xs)
    worst = 0.0
    for x, y in zip(xs, ys):
        t = ref(mpmath.mpf(float(x)))
        tf = float(t)
        if math.isinf(tf) or tf == 0.0:
            assert y == tf or abs(y) < 1e-300, (name, x, y, tf)
            continue
        ulp = math.ulp(tf)
        worst = max(worst, float(abs(mpmath.mpf(float(y)) - t) / ulp))
    return worst


def _pts(rng, lo, hi, n, log=False):
    if log:
        return np.sign(rng.standard_normal(n)) * 10.0 ** rng.uniform(lo, hi, n)
    return rng.uniform(lo, hi, n)


@pytest.mark.parametrize("name,ref,bound,ranges", [
    ("exp", mpmath.exp, 2.0, [(-745, 709, False), (-1e-3, 1e-3, False), (-1, 1, False)]),
    ("expm1", mpmath.expm1, 3.0, [(-50, 50, False), (-0.4, 0.4, False), (-12, -1, True)]),
    ("sin", mpmath.sin, 2.0, [(-10, 10, False), (-1e4, 1e4, False), (-12, -1, True)]),
    ("cos", mpmath.cos, 2.0, [(-10, 10, False), (-1e4, 1e4, False), (-12, -1, True)]),
    ("asin", mpmath.asin, 2.5, [(-1, 1, False), (-0.5, 0.5, False), (0.49, 0.51, False),
                                (0.999, 1.0, False), (-12, -1, True)]),
])
def test_ulp_error(name, ref, bound, ranges):
    rng = np.random.default_rng(len(name))
    for lo, hi, lg in ranges:
        xs = _pts(rng, lo, hi, 1500, lg)
        assert _ulp_err(name, xs, ref) <= bound, (name, lo, hi)


def test_log_ulp_error():
    rng = np.random.default_rng(7)
    xs = np.concatenate([10.0 ** rng.uniform(-300, 300, 1500), 1 + rng.uniform(-1e-3, 1e-3, 800),
                         rng.uniform(0.5, 2.0, 800), [5e-324, 2.2250738585072014e-308, 1.0]])
    assert _ulp_err("log", xs, mpmath.log) <= 2.5


def test_special_values():
    assert np.isnan(O.math_fn("log", [-1.0]))[0]
    assert O.math_fn("log", [0.0])[0] == -math.inf
    assert O.math_fn("exp", [800.0])[0] == math.inf
    assert O.math_fn("exp", [-800.0])[0] == 0.0
    assert O.math_fn("expm1", [-50.0])[0] == -1.0
    assert np.isnan(O.math_fn("sin", [2.0 ** 51]))[0]
    assert O.math_fn("sin", [0.0])[0] == 0.0 and O.math_fn("cos", [0.0])[0] == 1.0
    assert O.math_fn("asin", [1.0])[0] == math.pi / 2 and O.math_fn("asin", [-1.0])[0] == -math.pi / 2
    assert np.isnan(O.math_fn("asin", [1.0 + 2 ** -52]))[0]


@pytest.mark.parametrize("p", [2, 3, 4, 5, 6])
def test_integer_powers(p):
    rng = np.random.default_rng(p)
    xs = np.concatenate([rng.uniform(-10, 10, 2000), rng.uniform(-1e-3, 1e-3, 200)])
    assert _ulp_err(f"pow{p}", xs, lambda x: x ** p) <= 3.5
