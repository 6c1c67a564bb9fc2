"""The expression-program oracle pinned by: the reference's golden CROWN
bounds of even powers, closed forms of the exp / log / reciprocal / square
planes, per-operator and whole-program sampled soundness."""
import json
import math
import os

import numpy as np
import pytest

from gear_amd.problems import random_bisection_boxes
from gear_amd.program import Program, nlp_n8_m3_problem, random_nlp
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def _single(build, n=1):
    P = Program(n)
    P.output(build(P))
    return P


def test_even_power_matches_reference_golden():
    fx = json.load(open(os.path.join(HERE, "golden", "pow_crown_even.json")))
    L = np.array(fx["input_lb"])
    U = np.array(fx["input_ub"])
    for case in fx["cases"]:
        p = case["exponent"]
        for i in range(len(L)):
            lo = O.dag_crown(_single(lambda P: P.pow(0, p)), L[i:i + 1], U[i:i + 1])[0, 0]
            hi = -O.dag_crown(_single(lambda P: P.neg(P.pow(0, p))), L[i:i + 1], U[i:i + 1])[0, 0]
            assert lo == pytest.approx(case["lb"][i], rel=1e-9, abs=1e-9), (p, i)
            assert hi == pytest.approx(case["ub"][i], rel=1e-9, abs=1e-9), (p, i)


@pytest.mark.parametrize("l,u", [(-1.0, 0.5), (0.2, 3.0), (-4.0, -3.5), (1.0, 1.0 + 1e-3)])
def test_exp_parallel_tangent_closed_form(l, u):
    """Lower plane: the tangent whose slope equals the secant slope
    (convex_concave.py:580-600); on [l, u] its minimum is at x = l."""
    lb = O.dag_crown(_single(lambda P: P.exp(0)), np.array([l]), np.array([u]))[0, 0]
    if u - l > 1e-6:
        m = math.log((math.exp(u) - math.exp(l)) / (u - l))
    else:
        m = (l + u) / 2
    exact = math.exp(m) * (1 + l - m)
    # outward margins scale with the plane's terms, slope * |x| ~ e^u (1 + |l| + |u|)
    scale = math.exp(u) * (1 + abs(l) + abs(u))
    assert lb <= exact + 1e-15 and exact - lb <= 1e-13 * scale
    assert lb <= math.exp(l)


@pytest.mark.parametrize("l,u", [(0.5, 4.0), (1.0, 1.5), (2.0, 9.0)])
def test_log_planes_closed_form(l, u):
    """Lower plane: secant (minimum log(l)); upper plane: tangent at the
    logarithmic mean (u - l) / log(u / l) (convex_concave.py:64-92)."""
    lb = O.dag_crown(_single(lambda P: P.log(0)), np.array([l]), np.array([u]))[0, 0]
    assert lb <= math.log(l) and math.log(l) - lb < 1e-13
    ub = -O.dag_crown(_single(lambda P: P.neg(P.log(0))), np.array([l]), np.array([u]))[0, 0]
    m = (u - l) / math.log(u / l)
    exact = math.log(m) + (u - m) / m
    assert ub >= exact - 1e-15 and ub - exact < 1e-13 * (1 + abs(exact))


@pytest.mark.parametrize("l,u", [(0.5, 4.0), (-3.0, -0.25)])
def test_reciprocal_planes_closed_form(l, u):
    """Convex on (0, inf): tangent at the midpoint below, secant above;
    reversed on (-inf, 0) (convex_concave.py:315-375)."""
    lb = O.dag_crown(_single(lambda P: P.recip(0)), np.array([l]), np.array([u]))[0, 0]
    mid = (l + u) / 2
    if l > 0:
        exact = 1 / mid - (u - mid) / mid ** 2          # tangent, minimum at u
    else:
        exact = 1 / u                                   # secant, minimum at u
    assert lb <= exact + 1e-15 and exact - lb < 1e-13 * (1 + abs(exact))


def test_square_tangent_closed_form():
    """BoundSqr: tangent at relu(l) - relu(-u) (activations.py:199-207)."""
    for l, u, exact in [(0.5, 2.0, 0.25), (-3.0, -1.0, 1.0), (-1.0, 2.0, 0.0)]:
        lb = O.dag_crown(_single(lambda P: P.sqr(0)), np.array([l]), np.array([u]))[0, 0]
        assert lb <= exact and exact - lb < 1e-13 * (1 + abs(l) + abs(u)) ** 2


OPS = {
    "exp": (lambda P: P.exp(0), 1, False), "log": (lambda P: P.log(0), 1, True),
    "recip": (lambda P: P.recip(0), 1, True), "sqr": (lambda P: P.sqr(0), 1, False),
    "pow3": (lambda P: P.pow(0, 3), 1, False), "pow4": (lambda P: P.pow(0, 4), 1, False),
    "sin": (lambda P: P.sin(P.mulc(0, 3.0)), 1, False), "mul": (lambda P: P.mul(0, 1), 2, False),
    "div": (lambda P: P.div(0, P.addc(P.sqr(1), 1.0)), 2, False),
    "sin_of_product": (lambda P: P.sin(P.mul(0, 1)), 2, False),
    "log1p_sq": (lambda P: P.log(P.addc(P.sqr(0), 1.0)), 1, False),
}


@pytest.mark.parametrize("name", sorted(OPS))
@pytest.mark.parametrize("sign", [1, -1])
def test_operator_soundness(name, sign):
    build, n, positive = OPS[name]
    rng = np.random.default_rng(abs(hash(name)) % 1000 + (sign > 0))
    for _ in range(60):
        P = Program(n)
        node = build(P)
        P.output(node if sign > 0 else P.neg(node))
        c = rng.uniform(-3, 3, n)
        w = rng.uniform(0.01, 3, n)
        L, U = c - w / 2, c + w / 2
        if positive:
            L, U = np.abs(L) + 0.05, np.abs(L) + 0.05 + w
        lb = O.dag_crown(P, L[None], U[None])[0, 0]
        X = np.vstack([L + rng.random((800, n)) * (U - L), L, U])
        F = P.eval(X)[:, 0]
        assert np.nanmin(F) >= lb - 1e-12 * (1 + abs(lb))


@pytest.mark.parametrize("seed", range(4))
def test_random_program_soundness(seed):
    Q = random_nlp(12, 3, seed=seed)
    L, U = random_bisection_boxes(Q.xl, Q.xu, 24, seed=seed)
    lb = O.dag_crown(Q, L, U)
    rng = np.random.default_rng(seed)
    for b in range(len(L)):
        F = Q.eval(L[b] + rng.random((400, 12)) * (U[b] - L[b]))
        ok = np.isfinite(F)
        assert np.all(F[ok] >= np.broadcast_to(lb[b], F.shape)[ok] - 1e-12 * (1 + np.abs(F[ok])))


def test_pole_falls_back_to_interval():
    """1 / x over a box containing 0 has no relaxation: the row takes its
    interval bound (backward_bound.py:826-877), here -inf."""
    P = _single(lambda P: P.recip(0))
    lb = O.dag_crown(P, np.array([[-1.0]]), np.array([[2.0]]))[0, 0]
    assert lb == -math.inf
    # x + 1/y with y through 0: vacuous; x * exp(y) stays finite
    P = Program(2)
    P.output(P.add(0, P.recip(1)))
    P.output(P.mul(0, P.exp(1)))
    lb = O.dag_crown(P, np.array([[-1.0, -1.0]]), np.array([[1.0, 1.0]]))[0]
    assert lb[0] == -math.inf and np.isfinite(lb[1])


def test_nlp_n8_m3_sound():
    P = nlp_n8_m3_problem()
    L, U = random_bisection_boxes(P.xl, P.xu, 128, seed=2)
    lb = O.dag_crown(P, L, U)
    rng = np.random.default_rng(3)
    for b in range(len(L)):
        F = P.eval(L[b] + rng.random((300, 8)) * (U[b] - L[b]))
        assert np.all(F >= lb[b] - 1e-12 * (1 + np.abs(F)))
