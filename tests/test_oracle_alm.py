"""The ALM schedule replay (oracle.alm_schedule_check) pinned by hand-made
traces: the reference's rules worked through by hand (complete_verifier/
nlp_solver/alm.py:1851-1914), and planted mistakes it must report."""
import math

import numpy as np

from gear_amd.alm import ALMConfig
from oracle import oracle as O

INF = math.inf


def _rec(rho, eps, eta, viol, phi, steps, flag, gnorm, mu, g):
    return [rho, eps, eta, viol, phi, steps, flag, gnorm, *mu, *g]


def _cfg(**kw):
    base = dict(outer_max_iter=4, inner_steps=200, obj_patience=-1, first_feas_patience=-1)
    base.update(kw)
    return ALMConfig(**base)


def test_bab_mode_by_hand():
    """rho0 10, kappa 5, alpha 1/2, beta 0.8: eps0 = 10^-1/2, eta0 = 10^-0.8.
    k0: violation 0.5 > eps and > 1e-3: rho 50, eps 50^-1/2, mu unchanged.
    k1: violation 0.1 <= eps (0.1414): mu = max(0 + 50 g, 0), eps /= 50.
    k2: violation 5e-4 > eps (2.8e-3)? no -- 5e-4 <= 2.83e-3: mu moves again,
    eps /= 50 -> 5.66e-5.  k3: violation 2e-4 > eps, <= 1e-3: rho stays 50,
    eps, eta reset to 50^-alpha, 50^-beta."""
    cfg = _cfg()
    e0, h0 = 10 ** -0.5, 10 ** -0.8
    g0 = [-1.0, 0.5]
    recs = [_rec(10, e0, h0, 0.5, 3.0, 0, 0, INF, [0, 0], g0)]
    recs.append(_rec(50, 50 ** -0.5, 50 ** -0.8, 0.5, 2.0, 200, 0, 1.0, [0, 0], [-1.0, 0.5]))
    mu1 = [0.0, 50 * 0.1]
    e1, h1 = 50 ** -0.5 / 50, max(50 ** -0.8 / 50, 1e-5)
    recs.append(_rec(50, e1, h1, 0.1, 1.5, 200, 0, 1.0, mu1, [-0.2, 0.1]))
    mu2 = [0.0, 5.0 + 50 * 5e-4]
    e2, h2 = max(e1 / 50, 1e-5), max(h1 / 50, 1e-5)
    recs.append(_rec(50, e2, h2, 5e-4, 1.4, 200, 0, 1.0, mu2, [-0.3, 5e-4]))
    recs.append(_rec(50, 50 ** -0.5, 50 ** -0.8, 2e-4, 1.3, 200, 0, 1.0, mu2, [2e-4, -1.0]))
    tr = np.array(recs)[None]
    assert O.alm_schedule_check(cfg, tr, 3) == []
    bad = tr.copy()
    bad[0, 1, 0] = 10.0                       # rho not raised
    assert any(p[2] == "rho" for p in O.alm_schedule_check(cfg, bad, 3))
    bad = tr.copy()
    bad[0, 4, 0] = 250.0                      # rho raised at a small violation
    assert any(p[2] == "rho" for p in O.alm_schedule_check(cfg, bad, 3))
    bad = tr.copy()
    bad[0, 2, 9] = 5.0 + 1e-12                # mu off by an ulp-scale amount
    assert any(p[2] == "mu" for p in O.alm_schedule_check(cfg, bad, 3))


def test_root_mode_by_hand():
    """Root mode: mu moves every iteration; rho grows only when the violation
    fell by less than 1.25x (and exceeds feas_tol), never at k = 0."""
    cfg = _cfg(update_multipliers_always=True, outer_max_iter=3)
    e0, h0 = 10 ** -0.5, 10 ** -0.8
    recs = [_rec(10, e0, h0, 1.0, 0.0, 0, 0, INF, [0], [1.0])]
    mu = [10 * 1.0]
    recs.append(_rec(10, e0, h0, 1.0, 0.0, 200, 0, 1.0, mu, [1.0]))       # k0: no growth
    mu = [mu[0] + 10 * 0.9]
    recs.append(_rec(50, 50 ** -0.5, 50 ** -0.8, 0.9, 0.0, 200, 0, 1.0, mu, [0.9]))  # 1/0.9 < 1.25
    mu = [mu[0] + 50 * 0.5]
    recs.append(_rec(50, 50 ** -0.5, 50 ** -0.8, 0.5, 0.0, 200, 0, 1.0, mu, [0.5]))  # 1.8 >= 1.25
    tr = np.array(recs)[None]
    assert O.alm_schedule_check(cfg, tr, 2) == []
    bad = tr.copy()
    bad[0, 3, 0] = 250.0
    assert O.alm_schedule_check(cfg, bad, 2) != []


def test_exits_by_hand():
    """First-feasibility patience 1: feasible at k0 (plateau start), no 1e-4
    improvement at k1 -> stop before the update, the rest not run; an early
    inner exit needs gradient norm <= eta_k or a plateau step count."""
    cfg = _cfg(first_feas_patience=1, outer_max_iter=4)
    e0, h0 = 10 ** -0.5, 10 ** -0.8
    recs = [_rec(10, e0, h0, 0.0, 5.0, 0, 0, INF, [0], [-1.0])]
    e1, h1 = e0 / 10, h0 / 10
    recs.append(_rec(10, e1, h1, 0.0, 4.0, 57, 0, h0 * 0.5, [0], [-1.0]))
    recs.append(_rec(10, e1, h1, 0.0, 4.0 - 5e-5, 30, 1, 9.0, [0], [-1.0]))
    nan = [np.nan] * 8 + [np.nan, np.nan]
    recs.append(nan[:6] + [2.0] + nan[7:])
    recs.append(nan[:6] + [2.0] + nan[7:])
    tr = np.array(recs)[None]
    assert O.alm_schedule_check(cfg, tr, 2) == []
    bad = tr.copy()
    bad[0, 1, 5] = 57.0
    bad[0, 1, 7] = 1.0                        # stopped early with a large gradient
    assert any(p[2] == "early inner exit" for p in O.alm_schedule_check(cfg, bad, 2))
    bad = tr.copy()
    bad[0, 2, 6] = 0.0                        # did not stop
    assert any(p[2] == "stop" for p in O.alm_schedule_check(cfg, bad, 2))
