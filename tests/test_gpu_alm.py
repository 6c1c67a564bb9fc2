"""Fused ALM-PGD kernel (gh_alm_pgd_f64): the outer schedule after every
outer iteration equals the oracle's replay of the reference's rules (both
kernel forms, BaB and root mode); the unique optimum of a convex box QP
(scipy L-BFGS-B as the reference value); agreement with the autograd
implementation of the same iteration; every returned point feasible when
replayed in FP64 at 1e-6."""
import dataclasses
import time

import numpy as np
import pytest
import torch
from scipy.optimize import minimize

from gear_amd.alm import ALMConfig, ALMPGDSolver, replay_float64
from gear_amd.problems import QCQP, DeviceQCQP, qcqp_n50_m20_problem, qcqp_dense
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _box(p, dev):
    L = torch.as_tensor(p.xl, dtype=torch.float64, device=dev)[None]
    U = torch.as_tensor(p.xu, dtype=torch.float64, device=dev)[None]
    return L, U


# n <= 16 with S <= 8 runs the one-wavefront register kernel, larger n the
# workgroup kernel (rows over 8 or 4 wavefronts)
@pytest.mark.parametrize("n,m,seed", [(8, 3, 1), (16, 7, 4), (50, 20, 3), (90, 10, 2)])
@pytest.mark.parametrize("root", [False, True])
def test_schedule_matches_reference_rules(cuda, n, m, seed, root):
    print(f"case alm schedule n={n} m={m} root={root}", flush=True)
    p = qcqp_n50_m20_problem() if (n, m, seed) == (50, 20, 3) else qcqp_dense(n, m, seed=seed)
    dev = DeviceQCQP(p, cuda)
    cfg = ALMConfig(num_restarts=24, br_budget=0, update_multipliers_always=root)
    res = ALMPGDSolver(cfg, backend="hip").solve_with_restarts(dev, *_box(p, cuda), trace=True)
    tr = res.trace.cpu().numpy()
    probs = O.alm_schedule_check(cfg, tr, p.S)
    assert probs == [], probs[:5]
    flags = tr[:, 1:, 6]
    steps = tr[:, 1:, 5]
    print(f"  stopped early: {int((flags == 1).sum())} of {tr.shape[0]}, "
          f"mean inner steps {np.nanmean(np.where(flags < 2, steps, np.nan)):.1f}", flush=True)
    # the rules were exercised: rho grew somewhere and the multipliers moved
    assert np.nanmax(tr[:, :, 0]) > cfg.rho0 or root
    if res.x is not None:
        assert replay_float64(p, res.x, 1e-6)[2]


def test_convex_box_qp_reaches_optimum(cuda):
    rng = np.random.default_rng(0)
    n = 12
    A = rng.standard_normal((n, n)) / np.sqrt(n)
    Q = (A @ A.T + 0.1 * np.eye(n))[None]
    c = rng.standard_normal((1, n))
    p = QCQP(Q=Q, c=c, d=np.zeros(1), rhs=np.zeros(1), xl=-np.ones(n), xu=np.ones(n))
    dev = DeviceQCQP(p, cuda)
    # the iteration itself, without the early exits (the outer ones stop at a
    # 1e-4 gain, the inner plateau test at a 1e-3 relative change of L)
    cfg = ALMConfig(num_restarts=8, br_budget=0, obj_patience=-1, first_feas_patience=-1,
                    inner_plateau_check=0)
    res = ALMPGDSolver(cfg).solve_with_restarts(dev, *_box(p, cuda))
    f = lambda x: float(x @ Q[0] @ x + c[0] @ x)
    g = lambda x: (Q[0] + Q[0].T) @ x + c[0]
    ref = minimize(f, np.zeros(n), jac=g, method="L-BFGS-B", bounds=[(-1, 1)] * n,
                   options={"ftol": 1e-15, "gtol": 1e-12})
    print(f"kernel {res.obj:.10f} scipy {ref.fun:.10f}", flush=True)
    assert abs(res.obj - ref.fun) <= 1e-6 * (1 + abs(ref.fun))
    # with the presets' exits the objective is within their 1e-4 tolerance
    res2 = ALMPGDSolver(ALMConfig(num_restarts=8, br_budget=0)).solve_with_restarts(
        dev, *_box(p, cuda))
    assert res2.obj <= ref.fun + 1e-3 * (1 + abs(ref.fun))


@pytest.mark.parametrize("n,m,seed", [(50, 20, 3), (20, 8, 1), (90, 10, 2), (8, 3, 1), (15, 6, 2),
                                      (5, 2, 3), (16, 7, 4)])
def test_kernel_matches_autograd_iteration(cuda, n, m, seed):
    """The fused kernel and the op-by-op autograd form run the same
    iteration: from the same starts, after two outer iterations of 15 inner
    steps (the multipliers and rho updated in between) their iterates agree
    to 1e-9 of the box width -- longer runs drift apart by rounding alone.
    With the presets' full schedule both return FP64-feasible points."""
    p = qcqp_n50_m20_problem() if (n, m, seed) == (50, 20, 3) else qcqp_dense(n, m, seed=seed)
    dev = DeviceQCQP(p, cuda)
    L, U = _box(p, cuda)
    short = ALMConfig(num_restarts=16, br_budget=0, outer_max_iter=2, inner_steps=15,
                      obj_patience=-1,
                      first_feas_patience=-1, inner_plateau_check=0)
    hip = ALMPGDSolver(short, backend="hip")
    lo, hi, x0 = hip.starts(L, U, 16)
    xh = x0.clone()
    hip._run_hip(dev, lo, hi, xh, False)
    _, _, xt = ALMPGDSolver(short, backend="torch")._run_torch(dev, lo, hi, x0.clone())
    torch.cuda.synchronize()
    err = float(((xh - xt).abs() / (hi - lo)).max())
    print(f"n={n} m={m}: iterate difference after 2 x 15 steps {err:.3g} of the width",
          flush=True)
    assert err <= 1e-9
    cfg = ALMConfig(num_restarts=32, br_budget=0)
    t0 = time.perf_counter()
    rh = ALMPGDSolver(cfg, backend="hip").solve_with_restarts(dev, L, U)
    torch.cuda.synchronize()
    th = time.perf_counter() - t0
    t0 = time.perf_counter()
    rt = ALMPGDSolver(cfg, backend="torch").solve_with_restarts(dev, L, U)
    torch.cuda.synchronize()
    tt = time.perf_counter() - t0
    print(f"  presets: kernel {rh.obj:.8f} ({th * 1e3:.1f} ms), autograd {rt.obj:.8f} "
          f"({tt * 1e3:.1f} ms)", flush=True)
    for r in (rh, rt):
        assert r.x is not None
        obj, viol, feas = replay_float64(p, r.x, 1e-6)
        assert feas, viol
        assert abs(obj - r.obj) <= 1e-9 * (1 + abs(obj))


def test_many_boxes_one_launch(cuda):
    """B boxes x R restarts in one launch (R from the shared B*R budget):
    each restart stays in its box."""
    p = qcqp_dense(16, 4, seed=5)
    dev = DeviceQCQP(p, cuda)
    from gear_amd.problems import random_bisection_boxes
    Ln, Un = random_bisection_boxes(p.xl, p.xu, 40, seed=2)
    L = torch.as_tensor(Ln, device=cuda)
    U = torch.as_tensor(Un, device=cuda)
    s = ALMPGDSolver(ALMConfig(br_budget=160, outer_max_iter=5, inner_steps=50))
    res = s.solve_with_restarts(dev, L, U, trace=True)
    assert res.trace.shape[0] == 160
    if res.x is not None:
        x = res.x.cpu().numpy()
        assert np.any(np.all((x >= Ln - 1e-15) & (x <= Un + 1e-15), axis=1))
        assert replay_float64(p, res.x, 1e-6)[2]
