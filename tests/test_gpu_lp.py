"""LP dual certificate kernel (gh_lp_dual_bound_f64) on the CROWN planes of
QCQP boxes: below the exact weak-duality value of its own multipliers, equal
to the oracle's evaluation of them, below the LP optimum (HiGHS), and below
sampled feasible objective values."""
from fractions import Fraction

import numpy as np
import pytest
import torch
from scipy.optimize import linprog

from gear_amd.problems import qcqp_n50_m20_problem, qcqp_dense, random_bisection_boxes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _exact_D(c, c0, G, h, l, u, y):
    F = Fraction
    y = [max(F(float(v)), F(0)) for v in y]
    r = [F(float(c[j])) + sum(F(float(G[i, j])) * y[i] for i in range(G.shape[0]))
         for j in range(G.shape[1])]
    t = [min(r[j] * F(float(l[j])), r[j] * F(float(u[j]))) for j in range(G.shape[1])]
    return F(float(c0)) - sum(F(float(h[i])) * y[i] for i in range(G.shape[0])) + sum(t)


# (n, m, boxes): 16 lanes per box up to n, m = 16 (a partly filled last
# workgroup at 37 boxes), one wavefront per box up to 64, one workgroup above
@pytest.mark.parametrize("which", ["small", "n16_m16", "n5_m1", "n13_m6_b37", "n17_m4",
                                   "n6_m17", "n50_m20"])
def test_lp_certificate(cuda, which):
    from gear_amd.crown import BoundedQCQP
    from gear_amd.lp import qcqp_lp_bound
    sizes = {"small": (10, 3, 64), "n16_m16": (16, 16, 64), "n5_m1": (5, 1, 64),
             "n13_m6_b37": (13, 6, 37), "n17_m4": (17, 4, 64), "n6_m17": (6, 17, 64)}
    if which in sizes:
        n, m, nbox = sizes[which]
        p = qcqp_dense(n, m, seed=4)
    else:
        p, nbox = qcqp_n50_m20_problem(), 64
    L, U = random_bisection_boxes(p.xl, p.xu, nbox, seed=5)
    net = BoundedQCQP(p, cuda)
    xL = torch.as_tensor(L, device=cuda)
    xU = torch.as_tensor(U, device=cuda)
    lb, lA, lbias = net.bound(xL, xU, return_A=True)
    thr = torch.as_tensor(np.concatenate([[np.inf], p.rhs[1:] + 1e-6]), device=cuda)
    lp, y = qcqp_lp_bound(lA, lbias, thr, xL, xU, iters=500, return_y=True)
    torch.cuda.synchronize()
    lp, y = lp.cpu().numpy(), y.cpu().numpy()
    lAn, lbn, thrn = lA.cpu().numpy(), lbias.cpu().numpy(), thr.cpu().numpy()
    gaps = []
    checked_exact = 0
    for b in range(len(L)):
        c, c0, G = lAn[b, 0], lbn[b, 0], lAn[b, 1:]
        h = thrn[1:] - lbn[b, 1:]
        o = O.lp_dual_value(c, c0, G, h, L[b], U[b], y[b])
        assert abs(o - lp[b]) <= 1e-10 * (1 + abs(o)), b
        if checked_exact < 8:
            assert Fraction(float(lp[b])) <= _exact_D(c, c0, G, h, L[b], U[b], y[b]), b
            checked_exact += 1
        # the plane-only value (y = 0) is a floor of the kept best
        assert lp[b] >= O.lp_dual_value(c, c0, G, h, L[b], U[b], np.zeros(G.shape[0])) - 1e-12
        res = linprog(c, A_ub=G, b_ub=h, bounds=list(zip(L[b], U[b])), method="highs")
        if res.status == 0:
            opt = res.fun + c0
            assert lp[b] <= opt + 1e-9 * (1 + abs(opt)), b
            gaps.append((opt - lp[b]) / (1 + abs(opt)))
        elif res.status == 2:      # the planes already prove the box empty
            gaps.append(0.0)
    gaps = np.array(gaps)
    print(f"{which}: {len(L)} boxes, PDHG gap to the LP optimum: median {np.median(gaps):.2e}, "
          f"max {gaps.max():.2e}; LP over CROWN lb: mean {np.mean(lp - lb.cpu().numpy()[:, 0]):.4f}",
          flush=True)
    # soundness against the QCQP itself at sampled feasible points
    rng = np.random.default_rng(0)
    for b in range(0, len(L), 8):
        X = L[b] + rng.random((4000, p.n)) * (U[b] - L[b])
        F = p.values(X)
        feas = np.all(F[:, 1:] <= p.rhs[1:] + 1e-6, axis=1)
        if feas.any():
            assert lp[b] <= F[feas, 0].min() + 1e-9


# 16 lanes per box; one wavefront per box (the dual step's row sums split
# over 4 and 3 lane groups); one workgroup per box
@pytest.mark.parametrize("n,m", [(12, 4), (24, 6), (40, 20), (70, 5)])
def test_lp_certificate_cuts_with_binding_constraints(cuda, n, m):
    """Linear objective and linear constraints, so the planes are exact
    (lA[b] = c, lbias[b] = 0): the LP value rises above the plane-only box
    minimum and reaches the LP optimum."""
    from gear_amd.lp import qcqp_lp_bound
    rng = np.random.default_rng(3 + n)
    c = rng.standard_normal((m + 1, n))
    rhs = np.concatenate([[0.0], -np.abs(rng.standard_normal(m)) * 0.5])
    L, U = random_bisection_boxes(-np.ones(n), np.ones(n), 16, seed=1)
    xL = torch.as_tensor(L, device=cuda)
    xU = torch.as_tensor(U, device=cuda)
    lA = torch.as_tensor(np.broadcast_to(c, (len(L), m + 1, n)).copy(), device=cuda)
    lbias = torch.zeros(len(L), m + 1, dtype=torch.float64, device=cuda)
    thr = torch.as_tensor(np.concatenate([[np.inf], rhs[1:]]), device=cuda)
    lp = qcqp_lp_bound(lA, lbias, thr, xL, xU, iters=2000).cpu().numpy()
    rises = solved = 0
    for b in range(len(L)):
        G = c[1:]
        h = rhs[1:]
        res = linprog(c[0], A_ub=G, b_ub=h, bounds=list(zip(L[b], U[b])), method="highs")
        base = O.lp_dual_value(c[0], 0.0, G, h, L[b], U[b], np.zeros(m))
        if res.status == 2:
            continue
        solved += 1
        opt = res.fun
        assert lp[b] <= opt + 1e-9 * (1 + abs(opt))
        assert opt - lp[b] <= 1e-3 * (1 + abs(opt)), (b, opt, lp[b])
        if lp[b] > base + 1e-6:
            rises += 1
    print(f"n={n} m={m}: the constraints raise the certificate in {rises} of {solved} "
          f"feasible boxes", flush=True)
    assert solved >= 4 and rises >= solved // 2


def test_bab_with_lp_pruning_keeps_the_optimum(cuda):
    """B&B with the LP certificate in the prune closes the gap (the plain
    CROWN prune does not within the same iteration budget), never certifies
    above either run's incumbent, and bounds fewer boxes (the in-loop ALM
    searches different boxes in the two runs, so their incumbents agree only
    to the ALM's accuracy)."""
    from gear_amd.bab import BaBConfig, QCQPBranchAndBound
    from gear_amd.problems import QCQP
    rng = np.random.default_rng(8)
    n, m = 6, 3
    A = rng.standard_normal((n, n))
    Q = np.zeros((m + 1, n, n))
    Q[0] = A @ A.T / n - 0.5 * np.eye(n)           # nonconvex objective
    c = rng.standard_normal((m + 1, n))
    rhs = np.concatenate([[0.0], -np.abs(rng.standard_normal(m)) * 0.3])
    p = QCQP(Q=Q, c=c, d=np.zeros(m + 1), rhs=rhs, xl=-np.ones(n), xu=np.ones(n))
    out = {}
    for iters in (0, 300):
        bb = QCQPBranchAndBound(p, BaBConfig(batch=256, cap=1 << 16, dual_lp_iters=iters,
                                             max_iters=400, time_limit=60.0), device=cuda)
        bb.solve()
        out[iters] = (bb.P, bb.D, bb.stats.boxes_bounded)
        print(f"dual_lp_iters={iters}: P {bb.P:.8f} D {bb.D:.8f} boxes {bb.stats.boxes_bounded}",
              flush=True)
    from gear_amd.bab import is_primal_dual_converged
    (P0, D0, n0), (P1, D1, n1) = out[0], out[300]
    assert D1 <= min(P0, P1) + 1e-9 and D0 <= min(P0, P1) + 1e-9
    assert abs(P1 - P0) <= 1e-3 * (1 + abs(P0))
    assert is_primal_dual_converged(P1, D1, 1e-5, 1e-4)
    assert D0 <= D1 + 1e-9 * max(1.0, abs(P0))
    assert n1 <= n0


# 16 lanes per box, one wavefront per box, one workgroup per box
@pytest.mark.parametrize("n,m", [(10, 3), (17, 4), (70, 5)])
def test_lp_stops_at_the_objective_threshold(cuda, n, m):
    """With a finite thr[0] a box stops at the first checkpoint whose
    certified value reaches it: the prune decision (lb >= thr[0]) equals the
    one of the full run, the boxes that stay below run every iteration
    (bit-identical bounds), and every bound is the oracle's value of its own
    multipliers."""
    from gear_amd.crown import BoundedQCQP
    from gear_amd.lp import qcqp_lp_bound
    p = qcqp_dense(n, m, seed=7)
    L, U = random_bisection_boxes(p.xl, p.xu, 48, seed=3)
    xL = torch.as_tensor(L, device=cuda)
    xU = torch.as_tensor(U, device=cuda)
    if n <= 64:
        _, lA, lbias = BoundedQCQP(p, cuda).bound(xL, xU, return_A=True)
    else:   # the FP64 bound covers n <= 64: seeded planes of the same shape
        rng = np.random.default_rng(9)
        lA = torch.as_tensor(rng.standard_normal((len(L), m + 1, n)) / np.sqrt(n), device=cuda)
        lbias = torch.as_tensor(rng.standard_normal((len(L), m + 1)) * 0.3, device=cuda)
    thr = torch.as_tensor(np.concatenate([[np.inf], p.rhs[1:] + 1e-6]), device=cuda)
    full = qcqp_lp_bound(lA, lbias, thr, xL, xU, iters=400).cpu().numpy()
    fin = full[np.isfinite(full)]
    assert len(fin) > 8
    thr0 = float(np.median(fin))
    thr_t = thr.clone()
    thr_t[0] = thr0
    early, y = qcqp_lp_bound(lA, lbias, thr_t, xL, xU, iters=400, return_y=True)
    torch.cuda.synchronize()
    early, y = early.cpu().numpy(), y.cpu().numpy()
    assert np.array_equal(early >= thr0, full >= thr0)
    below = early < thr0
    assert below.any() and (~below).any()
    assert np.array_equal(early[below], full[below])
    assert np.all(early[~below] <= full[~below])
    lAn, lbn, thrn = lA.cpu().numpy(), lbias.cpu().numpy(), thr.cpu().numpy()
    for b in range(len(L)):
        c, c0, G = lAn[b, 0], lbn[b, 0], lAn[b, 1:]
        h = thrn[1:] - lbn[b, 1:]
        o = O.lp_dual_value(c, c0, G, h, L[b], U[b], y[b])
        if np.isfinite(early[b]):
            assert abs(o - early[b]) <= 1e-10 * (1 + abs(o)), b
    print(f"n={n} m={m}: {int((~below).sum())} of {len(L)} boxes stopped at thr0 {thr0:.6f}",
          flush=True)


def test_noncontiguous_inputs_equal_contiguous(cuda):
    """Strided views (transposed storage for the boxes, a row slice of the
    planes) give the same certificates as contiguous copies: each copy the
    binding makes stays alive until the kernel is queued."""
    from gear_amd.crown import BoundedQCQP
    from gear_amd.lp import qcqp_lp_bound
    p = qcqp_dense(10, 3, seed=4)
    L, U = random_bisection_boxes(p.xl, p.xu, 64, seed=5)
    net = BoundedQCQP(p, cuda)
    xL = torch.as_tensor(L, device=cuda)
    xU = torch.as_tensor(U, device=cuda)
    lb, lA, lbias = net.bound(xL, xU, return_A=True)
    thr = torch.as_tensor(np.concatenate([[np.inf], p.rhs[1:] + 1e-6]), device=cuda)
    ref = qcqp_lp_bound(lA, lbias, thr, xL, xU, iters=300)
    tL = xL.t().contiguous().t()        # same values, column-major storage
    tU = xU.t().contiguous().t()
    assert not tL.is_contiguous() and not tU.is_contiguous()
    wide = torch.cat([lA, lA], dim=1)[:, :lA.shape[1]]  # a strided view of the same planes
    assert not wide.is_contiguous()
    got = qcqp_lp_bound(wide, lbias, thr, tL, tU, iters=300)
    torch.cuda.synchronize()
    assert torch.equal(got, ref)
