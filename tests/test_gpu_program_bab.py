"""The B&B loop over expression programs (gear_amd.program_bab) against the
oracle-driven loop: sort, pick, bound with the rows' planes (never below the
stored key), widest branch, bisect, child keys from the objective plane,
prune, append -- bit for bit over the iterations; and the loop's certified
lower bound against sampled feasible points."""
import numpy as np
import pytest
import torch

from gear_amd.problems import random_bisection_boxes
from gear_amd.program import Program, nlp_n8_m3_problem, random_nlp
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64).view(np.int64)
    b = np.ascontiguousarray(b, dtype=np.float64).view(np.int64)
    return np.array_equal(a, b)


def _pole_program():
    """x0 / x1 with x1 crossing 0 in part of the tree (undefined rows) and a
    constraint through sin."""
    P = Program(3, name="pole")
    P.output(P.add(P.div(0, P.addc(1, 0.3)), P.sqr(2)))
    P.output(P.sin(P.mul(0, 2)), rhs=0.5)
    P.xl, P.xu = np.array([-1.0, -1.0, -1.0]), np.array([1.0, 1.0, 1.0])
    return P


def _incumbent(P, seed):
    """The best sampled feasible objective (a valid incumbent)."""
    rng = np.random.default_rng(seed)
    X = P.xl + rng.random((20000, P.n)) * (P.xu - P.xl)
    F = P.eval(X)
    feas = np.all(F[:, 1:] <= np.asarray(P.rhs[1:]), axis=1) & np.isfinite(F[:, 0])
    return float(F[feas, 0].min()) if feas.any() else np.inf


@pytest.mark.parametrize("which,batch,iters", [("nlp", 16, 8), ("random", 8, 10), ("pole", 4, 12),
                                               ("infeasible", 4, 3)])
def test_program_bab_iterations_match_oracle_loop(cuda, which, batch, iters):
    from gear_amd.bab import BaBConfig
    from gear_amd.program_bab import ProgramBranchAndBound, program_alm_config
    print(f"case program bab {which}", flush=True)
    P = {"nlp": nlp_n8_m3_problem, "random": lambda: random_nlp(6, 2, seed=14),
         "pole": _pole_program,
         # a constraint row's root bound is above its rhs: the root is dropped
         "infeasible": lambda: random_nlp(6, 2, seed=3)}[which]()
    bb = ProgramBranchAndBound(P, BaBConfig(batch=batch, cap=1 << 14, alm=program_alm_config(),
                                            alm_interval=0), device=cuda)
    bb.root(alm=False)
    inc = _incumbent(P, 1)
    if np.isfinite(inc):
        bb.set_incumbent(inc, None)
    thr = np.concatenate([[inc], np.asarray(P.rhs[1:], dtype=np.float64) + 1e-6])
    FL, FU = P.xl[None].astype(np.float64), P.xu[None].astype(np.float64)
    Flb = O.dag_crown(P, FL, FU)
    keep, _, _ = O.prune(Flb, thr, FL, FU)          # the root is filtered too
    FL, FU, Flb = FL[keep], FU[keep], Flb[keep]
    assert len(bb.frontier) == len(FL)
    pruned = 0
    for it in range(iters):
        if len(FL) == 0:
            break
        keys = (Flb - thr).max(axis=1)
        order = np.argsort(-keys, kind="stable")
        FL, FU, Flb = FL[order], FU[order], Flb[order]
        nb = min(2 * batch, len(FL))
        pL, pU, pk = FL[-nb:], FU[-nb:], Flb[-nb:]
        FL, FU, Flb = FL[:-nb], FU[:-nb], Flb[:-nb]
        lb, lA, lbias, und = O.dag_crown_planes(P, pL, pU)
        lb = np.maximum(lb, pk)
        cL, cU = O.split(pL, pU, O.branch_widest(pL, pU))
        ck = O.child_keys(lb, lA, lbias, cL, cU)
        keep, kept, _ = O.prune(ck, thr, cL, cU)
        pruned += len(keep) - kept
        FL = np.vstack([FL, cL[keep]])
        FU = np.vstack([FU, cU[keep]])
        Flb = np.vstack([Flb, ck[keep]])
        assert bb.step() == nb
        c = len(bb.frontier)
        assert c == len(FL), it
        assert _bits_equal(bb.frontier.L[:c].cpu().numpy(), FL)
        assert _bits_equal(bb.frontier.U[:c].cpu().numpy(), FU)
        assert _bits_equal(bb.frontier.lb[:c].cpu().numpy(), Flb)
        if len(FL) == 0:
            break
    print(f"  frontier {len(FL)}, pruned {pruned}", flush=True)
    if which in ("random", "pole"):
        assert pruned > 0
    if which == "infeasible":
        assert len(bb.frontier) == 0


def test_program_bab_solve_is_sound(cuda):
    """solve() with the root ALM incumbent: D never above a sampled feasible
    objective, and P (replayed in FP64) feasible."""
    from gear_amd.bab import BaBConfig
    from gear_amd.alm import replay_float64
    from gear_amd.program_bab import ProgramBranchAndBound, program_alm_config
    P = nlp_n8_m3_problem()
    bb = ProgramBranchAndBound(P, BaBConfig(batch=64, cap=1 << 15, max_iters=40,
                                            alm=program_alm_config(), alm_interval=0),
                               device=cuda)
    bb.solve()
    print(f"nlp n8 m3: P {bb.P:.6f} D {bb.D:.6f} boxes {bb.stats.boxes_bounded} "
          f"stop {bb.stats.stop}", flush=True)
    inc = _incumbent(P, 2)
    assert bb.D <= inc + 1e-9 * (1 + abs(inc))
    if bb.x_best is not None:
        assert replay_float64(bb.p, bb.x_best, 1e-6)[2]
        assert bb.D <= bb.P + 1e-9 * (1 + abs(bb.P))
