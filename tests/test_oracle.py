"""The oracle pinned against what fixes it independently of its own
formulas: sampled soundness, exact closed forms (linear rows, McCormick on a
single bilinear term in exact rational arithmetic), brute force on tiny
boxes, and the reference's split / branch / prune expectations."""
from fractions import Fraction

import numpy as np
import pytest

from gear_amd.problems import QCQP, qcqp_dense, random_bisection_boxes
from oracle import oracle as O


def _problem(n, m, seed, box=1.0):
    return qcqp_dense(n, m, seed, box=box)


@pytest.mark.parametrize("n,m,seed", [(3, 1, 0), (8, 3, 1), (20, 5, 2), (50, 20, 3)])
def test_sound_on_samples(n, m, seed):
    p = _problem(n, m, seed)
    L, U = random_bisection_boxes(p.xl, p.xu, 64, seed=seed + 10)
    lb, lA, lbias = O.qcqp_crown(p.Q, p.c, p.d, L, U, return_A=True)
    rng = np.random.default_rng(seed)
    for b in range(len(L)):
        x = L[b] + rng.random((256, n)) * (U[b] - L[b])
        # vertices too (a corner holds the minimum of a linear plane)
        corners = np.where(rng.random((64, n)) < 0.5, L[b], U[b])
        x = np.vstack([x, corners])
        f = p.values(x)
        assert np.all(f >= lb[b] - 1e-12 * (1 + np.abs(f))), b
        # the returned plane lA x + lbias is a valid lower plane of f
        plane = x @ lA[b].T + lbias[b]
        assert np.all(f - plane >= -1e-9 * (1 + np.abs(f))), b


def test_linear_rows_exact_minimum():
    """Q = 0: the bound is the exact box minimum of c x + d, moved outward
    only by the rounding margins."""
    rng = np.random.default_rng(4)
    n, S = 12, 5
    p = QCQP(np.zeros((S, n, n)), rng.standard_normal((S, n)), rng.standard_normal(S),
             np.zeros(S), -np.ones(n), np.ones(n))
    L, U = random_bisection_boxes(p.xl, p.xu, 32, seed=5)
    lb = O.qcqp_crown(p.Q, p.c, p.d, L, U)
    for b in range(len(L)):
        for s in range(S):
            exact = sum(min(Fraction(p.c[s, i]) * Fraction(L[b, i]),
                            Fraction(p.c[s, i]) * Fraction(U[b, i])) for i in range(n))
            exact += Fraction(p.d[s])
            scale = float(sum(abs(Fraction(p.c[s, i])) for i in range(n))) + abs(p.d[s]) + 1
            assert lb[b, s] <= float(exact)
            assert float(exact) - lb[b, s] <= 1e-13 * scale


def _mccormick_box_min(xl, xu, yl, yu):
    """max over the two textbook McCormick lower planes of their box minimum,
    in exact rational arithmetic."""
    xl, xu, yl, yu = map(Fraction, (xl, xu, yl, yu))
    planes = [(yl, xl, -xl * yl), (yu, xu, -xu * yu)]
    mins = [min(a * x + b * y + c for x in (xl, xu) for y in (yl, yu)) for a, b, c in planes]
    return max(mins)


@pytest.mark.parametrize("box", [((1, 3), (1, 4)), ((-3, -1), (-4, -1)), ((-2, 3), (1, 4)),
                                 ((1, 3), (-2, 4)), ((-2, 3), (-1, 2)), ((0.25, 0.5), (-7, 9))])
def test_single_bilinear_term_is_mccormick(box):
    """f = x0 * x1: y0 = x1 exactly, so the CROWN bound is the box minimum
    of the selected McCormick plane, the larger of the two."""
    (a, b), (c, d) = box
    n, S = 2, 1
    Q = np.zeros((S, n, n))
    Q[0, 0, 1] = 1.0
    p = QCQP(Q, np.zeros((S, n)), np.zeros(S), np.zeros(S), np.array([a, c], float),
             np.array([b, d], float))
    lb = O.qcqp_crown(p.Q, p.c, p.d, p.xl[None], p.xu[None])[0, 0]
    exact = float(_mccormick_box_min(a, b, c, d))
    assert lb <= exact
    assert exact - lb <= 1e-13 * (1 + abs(exact) + 4 * max(abs(a), abs(b)) * max(abs(c), abs(d)))
    # and it never exceeds the true minimum of x0 x1 (a vertex)
    assert lb <= min(x * y for x in (a, b) for y in (c, d))


@pytest.mark.parametrize("seed", range(6))
def test_dense_row_bilinear_uses_exact_linear_interval(seed):
    """f = x_i (q^T x) with q_i = 0 and 19 other nonzero q_j: the first
    Linear layer's interval of y = q^T x is exact, [sum q_j c_j - sum |q_j|
    r_j, + ...] (IBP of one linear form over a box), and y does not depend
    on x_i, so the CROWN bound is the McCormick rectangle minimum on
    [l_i, u_i] x [y_lo, y_hi], computed here in rational arithmetic -- a
    wrong centre or a radius without |q_j| moves it far outside the margin."""
    rng = np.random.default_rng(seed)
    n, i = 20, seed % 20
    q = rng.standard_normal(n)
    q[i] = 0.0
    Q = np.zeros((1, n, n))
    Q[0, i] = q
    xl = rng.uniform(-2, 0.5, n)
    xu = xl + rng.uniform(0.1, 2.5, n)
    p = QCQP(Q, np.zeros((1, n)), np.zeros(1), np.zeros(1), xl, xu)
    lb = O.qcqp_crown(p.Q, p.c, p.d, xl[None], xu[None])[0, 0]
    F = Fraction
    c = [(F(a) + F(b)) / 2 for a, b in zip(xl, xu)]
    r = [(F(b) - F(a)) / 2 for a, b in zip(xl, xu)]
    yc = sum(F(qj) * cj for qj, cj in zip(q, c))
    yr = sum(abs(F(qj)) * rj for qj, rj in zip(q, r))
    exact = _mccormick_box_min(xl[i], xu[i], yc - yr, yc + yr)
    scale = 1 + float(sum(abs(F(qj)) for qj in q)) * 4 * float(max(map(abs, np.r_[xl, xu]))) ** 2
    assert lb <= float(exact)
    assert float(exact) - lb <= 1e-13 * scale, (lb, float(exact))


def test_brute_force_tiny():
    """n = 2: the bound is below the minimum over a dense grid."""
    p = _problem(2, 2, 7, box=2.0)
    L, U = random_bisection_boxes(p.xl, p.xu, 16, seed=3)
    lb = O.qcqp_crown(p.Q, p.c, p.d, L, U)
    for b in range(len(L)):
        g0 = np.linspace(L[b, 0], U[b, 0], 101)
        g1 = np.linspace(L[b, 1], U[b, 1], 101)
        X = np.stack(np.meshgrid(g0, g1), -1).reshape(-1, 2)
        f = p.values(X)
        assert np.all(f.min(0) >= lb[b])


def test_bound_tightens_under_bisection():
    """Plain CROWN on a child box is not worse than the lower bound of f
    over the parent: min over the two children >= parent bound is not
    guaranteed, but every child bound stays below the child's true minimum
    and the union's minimum bound tends to the true minimum as boxes
    shrink."""
    p = _problem(4, 1, 8)
    L, U = p.xl[None].copy(), p.xu[None].copy()
    gaps = []
    for depth in range(12):
        lb = O.qcqp_crown(p.Q, p.c, p.d, L, U)
        rng = np.random.default_rng(depth)
        x = L[:, None, :] + rng.random((len(L), 64, 4)) * (U - L)[:, None, :]
        f = p.values(x.reshape(-1, 4)).reshape(len(L), 64, -1)
        gaps.append(float((f.min(1)[:, 0] - lb[:, 0]).max()))
        idx = O.branch_widest(L, U)
        L, U = O.split(L, U, idx)
    assert gaps[-1] < gaps[0] / 10


def test_branch_widest_ties_lowest_index():
    L = np.zeros((3, 4))
    U = np.array([[1, 1, 1, 1], [0.5, 2, 2, 1], [1, 1, 3, 0.5]], float)
    assert list(O.branch_widest(L, U)) == [0, 1, 2]


def test_split_matches_reference_semantics():
    """split.py:26-122 with P = 2: child 0 = [mid, xU], child 1 = [xL, mid],
    laid out [all child 0; all child 1] (test_input_split_split.py:47-120:
    doubled batch, valid bounds, midpoint used)."""
    L = np.zeros((2, 3))
    U = np.ones((2, 3))
    oL, oU = O.split(L, U, np.array([0, 1], dtype=np.int32))
    assert oL.shape == (4, 3)
    np.testing.assert_array_equal(oL, [[0.5, 0, 0], [0, 0.5, 0], [0, 0, 0], [0, 0, 0]])
    np.testing.assert_array_equal(oU, [[1, 1, 1], [1, 1, 1], [0.5, 1, 1], [1, 0.5, 1]])
    assert np.all(oL <= oU)


def test_prune_predicate():
    """get_remaining_index (solving mode): any row lb > thr, objective
    lb >= thr, objective +inf with a finite threshold; invalid boxes go too."""
    inf = np.inf
    thr = np.array([1.0, 0.5, 0.5])
    lb = np.array([
        [0.0, 0.0, 0.0],     # keep
        [1.0, 0.0, 0.0],     # objective lb >= P: drop
        [0.0, 0.6, 0.0],     # constraint lb > thr: drop
        [0.0, 0.5, 0.5],     # equal on a constraint row: keep
        [inf, 0.0, 0.0],     # +inf objective: drop
        [-inf, 0.0, 0.0],    # keep
        [np.nan, 0.0, 0.0],  # NaN is not verified: keep
        [0.0, 0.0, 0.0],     # invalid box: drop
    ])
    L = np.zeros((8, 2))
    U = np.ones((8, 2))
    L[7, 1] = 2.0
    keep, kept, dm = O.prune(lb, thr, L, U)
    assert list(keep) == [True, False, False, True, False, True, True, False]
    assert kept == 4
    assert dm == 1.0
    # no incumbent yet: objective row inactive
    keep, _, dm = O.prune(lb[:4], np.array([inf, 0.5, 0.5]), L[:4], U[:4])
    assert list(keep) == [True, True, False, True]
    assert dm == inf


def test_oracle_deterministic_across_threads():
    p = _problem(16, 4, 9)
    L, U = random_bisection_boxes(p.xl, p.xu, 200, seed=1)
    a = O.qcqp_crown(p.Q, p.c, p.d, L, U, threads=1)
    b = O.qcqp_crown(p.Q, p.c, p.d, L, U, threads=4)
    np.testing.assert_array_equal(a, b)


def test_branch_sb_reference_expectations():
    """The reference's explicit SB cases (complete_verifier/tests/
    test_input_split_branching_heuristics.py): a dominant lA coefficient
    picks its input; a non-finite bound sends the batch to the widest input."""
    L = np.zeros((2, 4))
    U = np.ones((2, 4))
    lA = np.zeros((2, 1, 4))
    lA[:, :, 2] = 10.0
    idx = O.branch_sb(L, U, np.zeros((2, 1)), np.zeros(1), lA, coeff_thresh=0.0, root_w=None)
    assert idx.tolist() == [2, 2]
    idx = O.branch_sb(np.zeros((1, 4)), np.array([[0.1, 0.8, 0.2, 0.4]]),
                      np.array([[np.nan, 0.0]]), np.zeros(2),
                      np.random.default_rng(0).standard_normal((1, 2, 4)), root_w=None)
    assert idx.tolist() == [1]


def test_branch_sb_score_closed_form():
    """score_i = sum_s max(|a_si|, t) w_i / 2 (argmax, lowest index on
    ties); +inf objective thresholds count as the box's bound; the
    exhaustive guard replaces a choice narrower than ratio x the widest
    relative width."""
    rng = np.random.default_rng(3)
    B, S, n = 40, 3, 6
    L = rng.uniform(-1, 0, (B, n))
    U = L + rng.uniform(0.01, 1.0, (B, n))
    lA = rng.standard_normal((B, S, n)) * (rng.random((B, S, n)) < 0.7)
    lb = rng.standard_normal((B, S))
    thr = np.array([np.inf, 0.5, 0.2])
    idx = O.branch_sb(L, U, lb, thr, lA, coeff_thresh=1e-3, root_w=None)
    score = (np.maximum(np.abs(lA), 1e-3) * (U - L)[:, None, :] / 2).sum(1)
    assert np.array_equal(idx, score.argmax(1))
    # ties: equal widths and coefficients -> the lowest index
    idx = O.branch_sb(np.zeros((1, 3)), np.ones((1, 3)), np.zeros((1, 1)), np.zeros(1),
                      np.ones((1, 1, 3)), root_w=None)
    assert idx.tolist() == [0]
    # the guard: input 0 has the only large coefficient but is 1e4 times
    # narrower (relative to the root) than input 1
    Lg, Ug = np.zeros((1, 2)), np.array([[1e-4, 1.0]])
    lAg = np.array([[[1e6, 0.0]]])
    assert O.branch_sb(Lg, Ug, np.zeros((1, 1)), np.zeros(1), lAg, root_w=None).tolist() == [0]
    assert O.branch_sb(Lg, Ug, np.zeros((1, 1)), np.zeros(1), lAg, root_w=np.ones(2),
                       min_width_ratio=1e-3).tolist() == [1]
