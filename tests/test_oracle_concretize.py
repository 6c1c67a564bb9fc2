"""Constrained concretization oracle (oracle.constrained_concretize) pinned
by the reference's own expectations, by linear programming (scipy HiGHS)
and by the exact maximum of the one-row dual."""
import numpy as np
import pytest
from scipy.optimize import linprog

from oracle import oracle as O


def test_reference_identity_case():
    """solver/tests/test_constrained_concretize.py:82-107: the identity with
    the cut x0 + x1 + 2 <= 0 on [-1, 2] x [-2, 1] has lower bounds (-1, -2)
    and upper bounds (0, -1)."""
    xL, xU = np.array([[-1.0, -2.0]]), np.array([[2.0, 1.0]])
    A, b = np.array([[[1.0, 1.0]]]), np.array([[2.0]])
    eye = np.eye(2)[None]
    lo, flags, kept = O.constrained_concretize(eye, A, b, xL, xU)
    up = -O.constrained_concretize(-eye, A, b, xL, xU)[0]
    assert flags[0] == 0 and kept[0] == 1
    np.testing.assert_allclose(lo[0], [-1.0, -2.0], rtol=0, atol=1e-12)
    assert np.all(lo[0] <= [-1.0, -2.0])         # outward
    np.testing.assert_allclose(up[0], [0.0, -1.0], rtol=0, atol=1e-12)
    assert np.all(up[0] >= [0.0, -1.0])          # outward


def test_reference_row_selection():
    """complete_verifier/tests/test_concretize_func.py:48-108: binding rows
    are kept in order; x0 centres uncentred rows; the cap keeps the rows with
    the largest worst-case value."""
    A = np.array([[[1., 0.], [1., 1.], [0., 1.]], [[0., 1.], [1., 1.], [1., 0.]]])
    d = np.array([[0.5, -5., 0.5], [-5., 0.5, -5.]])
    keep, dk, _ = O.constraint_rows(A[0], d[0], -np.ones(2), np.ones(2))
    # d: the centred offsets, widened down by their rounding enclosure
    assert keep.tolist() == [0, 2] and np.all(dk <= 0.5) and np.allclose(dk, 0.5, atol=1e-13)
    keep, _, _ = O.constraint_rows(A[1], d[1], -np.ones(2), np.ones(2))
    assert keep.tolist() == [1]
    keep, dk, _ = O.constraint_rows(np.array([[1., 0.], [1., 0.]]), np.array([-3.5, -1.5]),
                                    np.array([1., 1.]), np.array([3., 3.]))
    assert keep.tolist() == [1] and np.allclose(dk, 0.5, atol=1e-13)   # x_0 <= 1.5 binds on [1, 3]
    keep, dk, _ = O.constraint_rows(np.array([[1., 0.], [0., 1.], [1., 1.]]),
                                    np.array([0.2, 0.5, 0.1]), -np.ones(2), np.ones(2), 2)
    assert keep.tolist() == [1, 2] and np.allclose(sorted(dk), [0.1, 0.5], atol=1e-13)


def _lp_min(c, A, b, xl, xu):
    r = linprog(c, A_ub=A, b_ub=-b, bounds=list(zip(xl, xu)), method="highs")
    return r.fun if r.status == 0 else None


def _binding_instance(rng, X, K, B, H):
    xL = -rng.random((B, X)) - 0.1
    xU = rng.random((B, X)) + 0.1
    x0 = (xL + xU) / 2
    A = rng.standard_normal((B, K, X))
    b = -(A @ x0[..., None])[..., 0] + 0.3 * rng.standard_normal((B, K))
    obj = rng.standard_normal((B, H, X))
    return obj, A, b, xL, xU


def test_one_cut_is_exact():
    """With one cut the one-row dual is the LP (strong duality): the sweep's
    bound equals the LP optimum up to its enclosure and is never above it."""
    rng = np.random.default_rng(0)
    obj, A, b, xL, xU = _binding_instance(rng, 9, 1, 40, 3)
    root, flags, _ = O.constrained_concretize(obj, A, b, xL, xU)
    n = 0
    for i in range(40):
        if flags[i]:
            continue
        for h in range(3):
            v = _lp_min(obj[i, h], A[i], b[i], xL[i], xU[i])
            assert v is not None
            assert root[i, h] <= v + 1e-12
            assert root[i, h] >= v - 1e-9 * (1 + abs(v)), (i, h, root[i, h], v)
            n += 1
    assert n > 60


def test_many_cuts_valid_and_above_naive():
    rng = np.random.default_rng(1)
    obj, A, b, xL, xU = _binding_instance(rng, 12, 6, 30, 2)
    root, flags, kept = O.constrained_concretize(obj, A, b, xL, xU)
    naive, _, _ = O.constrained_concretize(obj, np.zeros((30, 0, 12)), np.zeros((30, 0)), xL, xU)
    gain = 0
    for i in range(30):
        if flags[i] & 1:
            continue
        for h in range(2):
            v = _lp_min(obj[i, h], A[i], b[i], xL[i], xU[i])
            if v is None:
                continue
            assert root[i, h] <= v + 1e-12
            assert root[i, h] >= naive[i, h]
            gain += root[i, h] > naive[i, h] + 1e-9
    assert gain > 10


def test_redundant_rows_change_nothing():
    """complete_verifier/tests/test_concretize_func.py:68-88: rows that hold on
    the whole box do not change any bound (bit for bit)."""
    rng = np.random.default_rng(2)
    B, X, H = 8, 5, 3
    xL, xU = -rng.random((B, X)), rng.random((B, X))
    x0, eps = (xU + xL) / 2, (xU - xL) / 2
    bind = rng.standard_normal((B, 2, X))
    bind_b = -(bind @ x0[..., None])[..., 0]
    slack = rng.standard_normal((B, 6, X))
    slack_b = -(slack @ x0[..., None])[..., 0] - (np.abs(slack) * eps[:, None]).sum(-1) - 0.1
    obj = rng.standard_normal((B, H, X))
    with_slack = O.constrained_concretize(
        obj, np.concatenate([slack[:, :3], bind, slack[:, 3:]], 1),
        np.concatenate([slack_b[:, :3], bind_b, slack_b[:, 3:]], 1), xL, xU)[0]
    only = O.constrained_concretize(obj, bind, bind_b, xL, xU)[0]
    assert np.array_equal(with_slack, only)


def test_cap_between_naive_and_full():
    rng = np.random.default_rng(3)
    obj, A, b, xL, xU = _binding_instance(rng, 6, 8, 20, 3)
    full = O.constrained_concretize(obj, A, b, xL, xU, max_rows=0)[0]
    cap = O.constrained_concretize(obj, A, b, xL, xU, max_rows=2)
    naive = O.constrained_concretize(obj, A[:, :0], b[:, :0], xL, xU)[0]
    assert np.all(cap[2] <= 2)
    ok = cap[1] == 0
    assert np.all(cap[0][ok] >= naive[ok])
    for i in np.nonzero(ok)[0]:
        for h in range(3):
            v = _lp_min(obj[i, h], A[i], b[i], xL[i], xU[i])
            if v is not None:
                assert cap[0][i, h] <= v + 1e-12 and full[i, h] <= v + 1e-12


def test_infeasible_and_covered_flags():
    xL, xU = np.zeros((1, 3)), np.ones((1, 3))
    obj = np.array([[[1.0, -2.0, 0.5]]])
    naive = O.constrained_concretize(obj, np.zeros((1, 0, 3)), np.zeros((1, 0)), xL, xU)[0]
    # x0 + x1 + x2 >= 4 misses the unit box
    r, f, _ = O.constrained_concretize(obj, -np.ones((1, 1, 3)), np.array([[4.0]]), xL, xU)
    assert f[0] & 1
    # x0 <= 5 holds on the whole box: no cut used, the box bound
    r, f, k = O.constrained_concretize(obj, np.array([[[1.0, 0, 0]]]), np.array([[-5.0]]), xL, xU)
    assert f[0] == 2 and k[0] == 0 and np.array_equal(r, naive)
    # non-finite rows carry no cut
    r, f, k = O.constrained_concretize(obj, np.array([[[np.nan, 1.0, 0]]]), np.array([[0.0]]),
                                       xL, xU)
    assert np.array_equal(r, naive)


def test_one_row_multiplier_maximises_the_dual():
    """The multiplier is the maximiser of the concave one-row dual
    g(beta) = c.x0 + beta d - sum |c + beta a| eps over beta >= 0: the
    bound equals max over {0} and the breakpoints of g, enumerated."""
    rng = np.random.default_rng(4)
    for t in range(200):
        X = int(rng.integers(1, 10))
        xl = -rng.random(X) - 0.05
        xu = rng.random(X) + 0.05
        x0, e = (xl + xu) / 2, (xu - xl) / 2
        c = rng.standard_normal(X)
        a = rng.standard_normal(X)
        if t % 5 == 0:
            a[rng.integers(0, X)] = 0.0
        bb = -(a @ x0) + 0.5 * rng.standard_normal()
        r, f, _ = O.constrained_concretize(c[None, None], a[None, None], np.array([[bb]]),
                                           xl[None], xu[None], max_rows=0)
        if f[0]:
            continue
        d = a @ x0 + bb
        cand = [0.0] + [-(ci / ai) for ci, ai in zip(c, a) if ai != 0 and -(ci / ai) > 0]
        g = max(c @ x0 + be * d - np.abs(c + be * a) @ e for be in cand)
        assert abs(r[0, 0] - g) <= 1e-9 * (1 + abs(g)), (t, r[0, 0], g)
        assert r[0, 0] <= g + 1e-12


# ---------------------------------------------------------------- clipping
from fractions import Fraction as Fr


def _exact_min(a, c, lo, hi):
    v = Fr(float(c))
    for ai, l, u in zip(a, lo, hi):
        ai, l, u = Fr(float(ai)), Fr(float(l)), Fr(float(u))
        v += min(ai * l, ai * u)
    return v


def _plain_clip(lo, hi, a, c, thr):
    xh, ep = (hi + lo) / 2, (hi - lo) / 2
    dm = a @ xh - np.abs(a) @ ep + c
    cx = (thr - (dm - a * xh + np.abs(a) * ep)) / np.where(a == 0, np.nan, a)
    xu = np.where(a > 0, cx, np.inf)
    xl = np.where(a < 0, cx, -np.inf)
    return np.maximum(xl, lo), np.minimum(xu, hi)


def test_clip_removes_no_exactly_feasible_point():
    """complete_verifier/tests/test_clip_outward.py:29-70: points exactly on
    the plane's threshold stay inside the clipped box (Fraction fuzz), a box
    is emptied only when the plane's exact minimum is above the threshold,
    and the clipped volume is within 1e-12 (log, relative) of the plain
    floating-point clip."""
    rng = np.random.default_rng(0)
    bad = 0
    worst = 0.0
    for trial in range(300):
        n = int(rng.integers(1, 5))
        scale = 10.0 ** int(rng.integers(-3, 4))
        lo = rng.standard_normal(n) - 1
        hi = lo + rng.random(n) * 2 + 0.1
        a = rng.standard_normal(n) * scale
        c = rng.standard_normal() * scale
        mid = _exact_min(a, c, lo, hi)
        top = -_exact_min(-a, -c, lo, hi)
        thr = float(mid + (top - mid) * Fr(int(rng.integers(5, 95)), 100))
        oL, oU, st = O.clip_domains(a[None, None], np.array([[c]]), np.array([thr]),
                                    lo[None], hi[None])
        oL, oU = oL[0], oU[0]
        if (oL > oU).any():
            bad += Fr(thr) >= mid
            continue
        for j in range(n):
            if a[j] == 0:
                continue
            for _ in range(5):
                x = [Fr(float(lo[k])) + (Fr(float(hi[k])) - Fr(float(lo[k]))) *
                     Fr(int(rng.integers(0, 1001)), 1000) for k in range(n)]
                rest = Fr(float(c)) + sum(Fr(float(a[k])) * x[k] for k in range(n) if k != j)
                x[j] = (Fr(thr) - rest) / Fr(float(a[j]))
                if not (Fr(float(lo[j])) <= x[j] <= Fr(float(hi[j]))):
                    continue
                if not all(Fr(float(oL[k])) <= x[k] <= Fr(float(oU[k])) for k in range(n)):
                    bad += 1
        rL, rU = _plain_clip(lo, hi, a, c, thr)
        wn, wo = np.maximum(oU - oL, 0), np.maximum(rU - rL, 0)
        if (wo > 0).all():
            worst = max(worst, abs(float(np.log(wn).sum() - np.log(wo).sum())))
    assert bad == 0
    assert worst < 1e-12, worst


def test_clip_certified_needs_margin():
    """complete_verifier/tests/test_clip_outward.py:73-86."""
    lo, hi = np.zeros((1, 2)), np.ones((1, 2))
    a, c = np.array([[[1.0, 1.0]]]), np.array([[0.1]])
    lb = (a[0, 0] @ ((hi + lo) / 2)[0] - np.abs(a[0, 0]) @ ((hi - lo) / 2)[0]) + 0.1
    oL, oU, st = O.clip_domains(a, c, np.array([np.nextafter(lb, -np.inf)]), lo, hi)
    assert (oL <= oU).all() and st[0] == 1
    oL, oU, st = O.clip_domains(a, c, np.array([lb - 1e-6]), lo, hi)
    assert (oL > oU).any() and st[0] == 2


def test_clip_keeps_every_unverified_point():
    """Several planes at once (the rows of a QCQP's CROWN planes): no sampled
    point with every plane at or below its threshold leaves the clipped box;
    an infinite threshold (no incumbent yet) clips nothing; boxes of the
    same parent share its planes (b % P)."""
    rng = np.random.default_rng(5)
    P, S, n, B = 4, 5, 6, 16
    A = rng.standard_normal((P, S, n))
    c = rng.standard_normal((P, S))
    thr = np.abs(rng.standard_normal(S)) * 2
    lo = -rng.random((B, n)) - 0.2
    hi = rng.random((B, n)) + 0.2
    oL, oU, st = O.clip_domains(A, c, thr, lo, hi)
    shrunk = 0
    for b in range(B):
        X = lo[b] + rng.random((4000, n)) * (hi[b] - lo[b])
        keep = ((X @ A[b % P].T + c[b % P]) <= thr).all(1)
        inside = ((X >= oL[b]) & (X <= oU[b])).all(1)
        assert not (keep & ~inside).any(), b
        shrunk += (oL[b] > lo[b]).any() or (oU[b] < hi[b]).any()
    assert shrunk > 0
    thr_inf = thr.copy()
    thr_inf[:] = np.inf
    oL, oU, st = O.clip_domains(A, c, thr_inf, lo, hi)
    assert np.array_equal(oL, lo) and np.array_equal(oU, hi)
