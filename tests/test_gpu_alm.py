"""Fused ALM-PGD kernel (gh_alm_pgd_f64): the unique optimum of a convex box
QP (scipy L-BFGS-B as the reference value), and agreement with the
autograd implementation of the same iteration on nonconvex QCQPs, with every
returned point feasible when replayed in FP64 at 1e-6."""
import time

import numpy as np
import pytest
import torch
from scipy.optimize import minimize

from gear_amd.alm import ALMConfig, ALMPGDSolver, replay_float64
from gear_amd.problems import QCQP, DeviceQCQP, qcqp_n50_m20_problem, qcqp_dense

pytestmark = pytest.mark.gpu


def _box(p, dev):
    L = torch.as_tensor(p.xl, dtype=torch.float64, device=dev)[None]
    U = torch.as_tensor(p.xu, dtype=torch.float64, device=dev)[None]
    return L, U


def test_convex_box_qp_reaches_optimum(cuda):
    rng = np.random.default_rng(0)
    n = 12
    A = rng.standard_normal((n, n)) / np.sqrt(n)
    Q = (A @ A.T + 0.1 * np.eye(n))[None]
    c = rng.standard_normal((1, n))
    p = QCQP(Q=Q, c=c, d=np.zeros(1), rhs=np.zeros(1), xl=-np.ones(n), xu=np.ones(n))
    dev = DeviceQCQP(p, cuda)
    res = ALMPGDSolver(ALMConfig(num_restarts=8)).solve_with_restarts(dev, *_box(p, cuda))
    f = lambda x: float(x @ Q[0] @ x + c[0] @ x)
    g = lambda x: (Q[0] + Q[0].T) @ x + c[0]
    ref = minimize(f, np.zeros(n), jac=g, method="L-BFGS-B", bounds=[(-1, 1)] * n,
                   options={"ftol": 1e-15, "gtol": 1e-12})
    print(f"kernel {res.obj:.10f} scipy {ref.fun:.10f}", flush=True)
    assert abs(res.obj - ref.fun) <= 1e-6 * (1 + abs(ref.fun))


# n <= 16 with S <= 4 / S <= 8 runs the one-wavefront register kernel
@pytest.mark.parametrize("n,m,seed", [(50, 20, 3), (20, 8, 1), (90, 10, 2), (8, 3, 1), (15, 6, 2),
                                      (5, 2, 3), (16, 7, 4)])
def test_kernel_matches_autograd_iteration(cuda, n, m, seed):
    p = qcqp_n50_m20_problem() if (n, m, seed) == (50, 20, 3) else qcqp_dense(n, m, seed=seed)
    dev = DeviceQCQP(p, cuda)
    cfg = ALMConfig(num_restarts=32)
    L, U = _box(p, cuda)
    hip = ALMPGDSolver(cfg, backend="hip")
    tch = ALMPGDSolver(cfg, backend="torch")
    hip.solve_with_restarts(dev, L, U)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rh = hip.solve_with_restarts(dev, L, U)
    torch.cuda.synchronize()
    th = time.perf_counter() - t0
    t0 = time.perf_counter()
    rt = tch.solve_with_restarts(dev, L, U)
    torch.cuda.synchronize()
    tt = time.perf_counter() - t0
    print(f"n={n} m={m}: kernel {rh.obj:.8f} ({th * 1e3:.1f} ms), autograd {rt.obj:.8f} "
          f"({tt * 1e3:.1f} ms)", flush=True)
    for r in (rh, rt):
        assert r.x is not None
        obj, viol, feas = replay_float64(p, r.x, 1e-6)
        assert feas, viol
        assert abs(obj - r.obj) <= 1e-9 * (1 + abs(obj))
    assert abs(rh.obj - rt.obj) <= 1e-2 * (1 + abs(rt.obj))


def test_many_boxes_one_launch(cuda):
    """B boxes x R restarts in one launch: each restart stays in its box."""
    p = qcqp_dense(16, 4, seed=5)
    dev = DeviceQCQP(p, cuda)
    from gear_amd.problems import random_bisection_boxes
    Ln, Un = random_bisection_boxes(p.xl, p.xu, 40, seed=2)
    L = torch.as_tensor(Ln, device=cuda)
    U = torch.as_tensor(Un, device=cuda)
    s = ALMPGDSolver(ALMConfig(num_restarts=4, outer_max_iter=5, inner_steps=50))
    res = s.solve_with_restarts(dev, L, U)
    if res.x is not None:
        x = res.x.cpu().numpy()
        assert np.any(np.all((x >= Ln - 1e-15) & (x <= Un + 1e-15), axis=1))
        assert replay_float64(p, res.x, 1e-6)[2]
