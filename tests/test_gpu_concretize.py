"""Constrained concretization kernel against the oracle, bit for bit."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _case(rng, B, H, K, X, special=False):
    xL = -rng.random((B, X)) - 0.05
    xU = rng.random((B, X)) + 0.05
    x0 = (xL + xU) / 2
    A = rng.standard_normal((B, K, X))
    b = -(A @ x0[..., None])[..., 0] + 0.4 * rng.standard_normal((B, K))
    obj = rng.standard_normal((B, H, X))
    if special and K:
        A[::7, 0, 0] = np.nan                       # a non-finite cut
        A[1::5, -1] = 0.0                           # an all-zero cut row
        b[2::9, 0] = 1e3                            # misses the box: infeasible
        A[3::4, :, X // 2] = 0.0                    # zero coefficients: q = +-inf / nan
        obj[4::6, :, 0] = 0.0
        xU[5::8, 0] = xL[5::8, 0]                   # a flat coordinate
        b[6::10] = -1e3                             # every cut redundant
    return obj, A, b, xL, xU


@pytest.mark.parametrize("B,H,K,X,max_rows,special", [
    (300, 3, 1, 9, 16, False), (300, 4, 6, 12, 16, True), (257, 21, 21, 50, 16, True),
    (64, 5, 21, 50, 0, True), (100, 2, 40, 64, 16, True), (50, 1, 3, 1, 0, False),
    (80, 3, 0, 7, 16, False), (128, 7, 12, 33, 4, True)])
def test_matches_oracle(cuda, B, H, K, X, max_rows, special):
    from gear_amd.concretize import constrained_concretize
    print(f"case B={B} H={H} K={K} X={X} cap={max_rows}", flush=True)
    rng = np.random.default_rng(B + K + X)
    obj, A, b, xL, xU = _case(rng, B, H, K, X, special)
    r0, f0, k0 = O.constrained_concretize(obj, A, b, xL, xU, max_rows=max_rows)
    d = lambda a: torch.as_tensor(a, device=cuda)
    r, f, k = constrained_concretize(d(obj), d(A), d(b), d(xL), d(xU), max_rows=max_rows)
    torch.cuda.synchronize()
    assert np.array_equal(f.cpu().numpy(), f0)
    assert np.array_equal(k.cpu().numpy(), k0)
    got = r.cpu().numpy()
    assert np.array_equal(got.view(np.int64), r0.view(np.int64)), \
        f"max diff {np.nanmax(np.abs(got - r0))}"
    if special:
        assert (f0 & 1).any() and (f0 & 2).any() and (k0 > 0).any()


def test_bound_form_matches_oracle(cuda):
    """The bound form on real CROWN planes of a QCQP with the constraint
    rows' planes as cuts: equal to the oracle bit for bit, never below the
    box bound, above it on some rows, +inf on infeasible domains."""
    from gear_amd.concretize import constrained_bound
    from gear_amd.crown import BoundedQCQP
    from gear_amd.problems import qcqp_dense, random_bisection_boxes
    p = qcqp_dense(20, 6, seed=4)
    L, U = random_bisection_boxes(p.xl, p.xu, 400, seed=3)
    lb0, lA, lbias = O.qcqp_crown(p.Q, p.c, p.d, L, U, return_A=True)
    thr = np.concatenate([[np.inf], p.rhs[1:] + 1e-6])
    cutA = lA[:, 1:].copy()
    cutb = lbias[:, 1:] - thr[1:]
    cutb[::11] -= 50.0            # a few boxes with stronger cuts
    cutb[5::37, 0] += 1e4         # and a few infeasible ones
    ref, f0, k0 = O.constrained_bound(lA, lbias, cutA, cutb, L, U, lb0)
    d = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=cuda)
    lb = d(lb0)
    got, f, k = constrained_bound(d(lA), d(lbias), d(cutA), d(cutb), d(L), d(U), lb)
    torch.cuda.synchronize()
    g = got.cpu().numpy()
    assert np.array_equal(f.cpu().numpy(), f0) and np.array_equal(k.cpu().numpy(), k0)
    assert np.array_equal(g.view(np.int64), ref.view(np.int64))
    assert np.all(g >= lb0)
    assert (g > lb0).any()
    assert np.all(np.isinf(g[(f0 & 1) == 1]))
    assert ((f0 & 1) == 1).any()


@pytest.mark.parametrize("P,S,n,B", [(64, 21, 50, 128), (7, 3, 5, 14), (1, 1, 1, 9),
                                     (16, 4, 64, 32)])
def test_clip_matches_oracle(cuda, P, S, n, B):
    from gear_amd.concretize import clip_domains
    print(f"case P={P} S={S} n={n} B={B}", flush=True)
    rng = np.random.default_rng(P + S + n)
    A = rng.standard_normal((P, S, n))
    c = rng.standard_normal((P, S))
    thr = np.abs(rng.standard_normal(S)) * 3
    thr[0] = np.inf if S > 2 else thr[0]
    lo = -rng.random((B, n)) - 0.2
    hi = rng.random((B, n)) + 0.2
    if P > 2:
        A[1, 0, 0] = np.nan                  # non-finite planes: left as is
        c[2, -1] = 1e3                       # a row that certifies its boxes
    oL0, oU0, st0 = O.clip_domains(A, c, thr, lo, hi)
    d = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=cuda)
    xL, xU = d(lo), d(hi)
    oL, oU, st = clip_domains(d(A), d(c), d(thr), xL, xU, out=(xL, xU))    # in place
    torch.cuda.synchronize()
    assert np.array_equal(st.cpu().numpy(), st0)
    assert np.array_equal(oL.cpu().numpy().view(np.int64), oL0.view(np.int64))
    assert np.array_equal(oU.cpu().numpy().view(np.int64), oU0.view(np.int64))
    assert (st0 == 1).any()


def _tight_instance():
    from gear_amd.problems import QCQP
    rng = np.random.default_rng(8)
    n, m = 6, 3
    A = rng.standard_normal((n, n))
    Q = np.zeros((m + 1, n, n))
    Q[0] = A @ A.T / n - 0.5 * np.eye(n)
    c = rng.standard_normal((m + 1, n))
    rhs = np.concatenate([[0.0], -np.abs(rng.standard_normal(m)) * 0.3])
    return QCQP(Q=Q, c=c, d=np.zeros(m + 1), rhs=rhs, xl=-np.ones(n), xu=np.ones(n))


def test_clip_complete_loop_matches_oracle(cuda):
    """Twelve iterations of the reordered loop with clip_type 'complete' on
    the device reproduce the oracle-driven loop bit for bit: frontier boxes,
    keys and stored cuts; the constrained bound raises bounds and clipping
    shrinks children along the way."""
    from gear_amd.bab import BaBConfig, QCQPBranchAndBound
    p = _tight_instance()
    n, S = p.n, p.S
    bb = QCQPBranchAndBound(p, BaBConfig(batch=16, cap=8192, clip="complete"), device=cuda)
    bb.root(alm=False)
    bb.set_incumbent(-0.5, None)
    thr = np.concatenate([[-0.5], p.rhs[1:] + 1e-6])
    FL, FU = p.xl[None].copy(), p.xu[None].copy()
    Flb = O.qcqp_crown(p.Q, p.c, p.d, FL, FU)
    Fpay = np.zeros((1, S * (n + 1)))
    Fpay[0, S * n:] = -1.0
    raised = shrunk = 0
    for it in range(12):
        keys = (Flb - thr).max(axis=1)
        order = np.argsort(-keys, kind="stable")
        FL, FU, Flb, Fpay = FL[order], FU[order], Flb[order], Fpay[order]
        nb = min(32, len(FL))
        pL, pU, pk, pp = FL[-nb:], FU[-nb:], Flb[-nb:], Fpay[-nb:]
        FL, FU, Flb, Fpay = FL[:-nb], FU[:-nb], Flb[:-nb], Fpay[:-nb]
        lb0, lA, lbias = O.qcqp_crown(p.Q, p.c, p.d, pL, pU, return_A=True)
        lb, _, _ = O.constrained_bound(lA, lbias, pp[:, :S * n].reshape(nb, S, n), pp[:, S * n:],
                                       pL, pU, lb0)
        raised += int((lb > lb0).sum())
        act = np.isfinite(thr)
        cuts = np.concatenate([np.where(act[None, :, None], lA, 0.0).reshape(nb, S * n),
                               np.where(act[None, :], lbias - thr, -1.0)], 1)
        lb = np.maximum(lb, pk)
        cL0, cU0 = O.split(pL, pU, O.branch_widest(pL, pU))
        cL, cU, _ = O.clip_domains(lA, lbias, thr, cL0, cU0)
        shrunk += int(((cL > cL0).any(1) | (cU < cU0).any(1)).sum())
        ck = O.child_keys(lb, lA, lbias, cL, cU)
        keep, _, _ = O.prune(ck, thr, cL, cU)
        FL = np.vstack([FL, cL[keep]])
        FU = np.vstack([FU, cU[keep]])
        Flb = np.vstack([Flb, ck[keep]])
        Fpay = np.vstack([Fpay, cuts[np.nonzero(keep)[0] % nb]])
        assert bb.step() == nb
        f = bb.frontier
        c = len(f)
        assert c == len(FL), it
        assert np.array_equal(f.L[:c].cpu().numpy(), FL)
        assert np.array_equal(f.U[:c].cpu().numpy(), FU)
        assert np.array_equal(f.lb[:c].cpu().numpy().view(np.int64), Flb.view(np.int64))
        assert np.array_equal(f.pay[:c].cpu().numpy().view(np.int64), Fpay.view(np.int64))
    assert raised > 0 and shrunk > 0


def test_clip_complete_closes_with_fewer_boxes(cuda):
    """On a tightly constrained instance the clip_type 'complete' loop
    reaches the 1e-4 gap with a sound bound and fewer bounded domains than
    box concretization alone (the in-loop ALM searches different boxes in
    the two runs, so their incumbents agree only to the ALM's accuracy)."""
    from gear_amd.bab import BaBConfig, QCQPBranchAndBound, is_primal_dual_converged
    p = _tight_instance()
    out = {}
    for clip in ("none", "complete"):
        bb = QCQPBranchAndBound(p, BaBConfig(batch=256, cap=1 << 16, clip=clip, max_iters=400,
                                             time_limit=60.0), device=cuda)
        bb.solve()
        out[clip] = (bb.P, bb.D, bb.stats.boxes_bounded)
        print(f"clip={clip}: P {bb.P:.8f} D {bb.D:.8f} boxes {bb.stats.boxes_bounded}", flush=True)
    (P0, D0, n0), (P1, D1, n1) = out["none"], out["complete"]
    assert abs(P1 - P0) <= 1e-3 * (1 + abs(P0))
    assert D1 <= min(P0, P1) + 1e-9 and D0 <= min(P0, P1) + 1e-9
    assert is_primal_dual_converged(P1, D1, 1e-5, 1e-4)
    assert n1 < n0
