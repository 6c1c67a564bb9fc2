"""Child key (oracle.plane_minimum_down / child_keys) against the exact
rational minimum of the plane over the box, the reference's own check
(CV/tests/test_child_key.py:78-111: random fuzz, never above the exact
minimum)."""
from fractions import Fraction

import numpy as np

from oracle import oracle as O


def _exact_min(a, b, xl, xu):
    return sum(min(Fraction(float(ai)) * Fraction(float(l)), Fraction(float(ai)) * Fraction(float(u)))
               for ai, l, u in zip(a, xl, xu)) + Fraction(float(b))


def test_never_above_exact_minimum_and_tight():
    rng = np.random.default_rng(0)
    for trial in range(2000):
        n = int(rng.integers(1, 40))
        scale = 10.0 ** rng.uniform(-3, 3)
        a = rng.standard_normal(n) * scale
        b = float(rng.standard_normal() * scale)
        xl = rng.uniform(-5, 5, n)
        xu = xl + rng.uniform(0, 3, n)
        pm = O.plane_minimum_down(a, b, xl, xu)
        ex = _exact_min(a, b, xl, xu)
        assert Fraction(float(pm)) <= ex, trial
        mag = np.abs(a) @ np.maximum(np.abs(xl), np.abs(xu)) + abs(b)
        assert float(ex) - pm <= 1e-13 * (1 + mag) * (n + 5), trial


def test_child_keys_layout():
    rng = np.random.default_rng(1)
    K, S, n = 5, 3, 4
    plb = rng.standard_normal((K, S))
    lA = rng.standard_normal((K, S, n))
    lbias = rng.standard_normal((K, S))
    L, U = -np.ones((K, n)), np.ones((K, n))
    idx = O.branch_widest(L, U)
    cL, cU = O.split(L, U, idx)
    keys = O.child_keys(plb, lA, lbias, cL, cU)
    assert keys.shape == (2 * K, S)
    assert np.array_equal(keys[:, 1:], np.vstack([plb, plb])[:, 1:])
    assert np.all(keys[:, 0] >= np.vstack([plb, plb])[:, 0])
    for k in range(2 * K):
        p = k % K
        pm = O.plane_minimum_down(lA[p, 0], lbias[p, 0], cL[k], cU[k])
        assert keys[k, 0] == max(plb[p, 0], pm)
