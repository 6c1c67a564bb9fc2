"""The LP-certificate restatement (oracle.lp_dual_value) against exact
rational arithmetic and against LP optima from scipy's HiGHS."""
from fractions import Fraction

import numpy as np
import pytest
from scipy.optimize import linprog

from oracle import oracle as O


def _exact_D(c, c0, G, h, l, u, y):
    F = Fraction
    y = [max(F(float(v)), F(0)) for v in y]
    r = [F(float(c[j])) + sum(F(float(G[i, j])) * y[i] for i in range(G.shape[0]))
         for j in range(G.shape[1])]
    t = [min(r[j] * F(float(l[j])), r[j] * F(float(u[j]))) for j in range(G.shape[1])]
    return F(float(c0)) - sum(F(float(h[i])) * y[i] for i in range(G.shape[0])) + sum(t)


def _lp(seed, m, n):
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((m, n))
    x0 = rng.uniform(-0.5, 0.5, n)
    h = G @ x0 + rng.uniform(0.1, 1.0, m)          # x0 is strictly feasible
    c = rng.standard_normal(n)
    c0 = float(rng.standard_normal())
    l, u = -np.ones(n), np.ones(n)
    return c, c0, G, h, l, u


@pytest.mark.parametrize("seed", range(6))
def test_rigorous_value_below_exact_and_tight(seed):
    c, c0, G, h, l, u = _lp(seed, 6, 9)
    rng = np.random.default_rng(100 + seed)
    for y in [np.zeros(6), rng.exponential(1.0, 6), rng.exponential(1e3, 6), np.r_[0, 0, 5.0, 0, 1, 0]]:
        lb = O.lp_dual_value(c, c0, G, h, l, u, y)
        ex = _exact_D(c, c0, G, h, l, u, y)
        assert Fraction(lb) <= ex
        assert float(ex) - lb <= 1e-12 * (1 + abs(float(ex)) + np.abs(y).sum() * 10)


@pytest.mark.parametrize("seed", range(4))
def test_weak_duality_and_lp_optimum(seed):
    """The optimal multipliers HiGHS returns give a certified value within a
    hair of the LP optimum, never above it."""
    c, c0, G, h, l, u = _lp(seed, 5, 8)
    res = linprog(c, A_ub=G, b_ub=h, bounds=list(zip(l, u)), method="highs")
    assert res.status == 0
    opt = res.fun + c0
    y = np.maximum(-res.ineqlin.marginals, 0.0)
    lb = O.lp_dual_value(c, c0, G, h, l, u, y)
    assert lb <= opt + 1e-9 * (1 + abs(opt))
    assert opt - lb <= 1e-7 * (1 + abs(opt))
    # any other y >= 0 certifies less
    rng = np.random.default_rng(seed)
    for _ in range(20):
        assert O.lp_dual_value(c, c0, G, h, l, u, rng.exponential(1.0, 5)) <= opt + 1e-9 * (1 + abs(opt))


def test_plane_only_value_is_box_minimum():
    """y = 0: the box minimum of the objective plane, rounded down."""
    c, c0, G, h, l, u = _lp(7, 3, 4)
    lb = O.lp_dual_value(c, c0, G, h, l, u, np.zeros(3))
    exact = c0 + np.minimum(c * l, c * u).sum()
    assert lb <= exact and exact - lb < 1e-13 * (1 + abs(exact))
