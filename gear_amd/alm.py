"""ALM-PGD primal search over boxes (batched restarts), on the device.

Follows the reference's HPR augmented Lagrangian with a projected, clipped
Adam inner loop (complete_verifier/nlp_solver/alm.py):
  * L(x) = phi(x) + 0.5 rho ||relu(g(x) + mu / rho)||^2  (:655-676);
  * AdamClipping_NLP.step: Adam moments, the raw lr times m / denom,
    clipped per coordinate to lr_vec (divided by the first-moment bias
    correction when scale_eps_with_bias), projected onto the box (:134-160);
  * lr decays linearly from 1.0 to 0.1 over the inner steps (:603-608);
  * multipliers mu = max(mu + rho g, 0); rho *= kappa where the violation
    did not shrink enough (:1860-1907), capped at max_rho;
  * restart 0 starts at the box centre, the others uniformly in the box
    (:1337-1363).
The incumbent is the best point whose replayed FP64 violation is within
feas_tol; its objective is replayed in FP64 (gear/report.py:114-141).
"""
from dataclasses import dataclass

import ctypes

import torch


@dataclass
class ALMConfig:
    rho0: float = 1.0
    kappa: float = 1.1
    max_rho: float = 1e6
    outer_max_iter: int = 20
    inner_steps: int = 200
    inner_lr: float = 0.01
    feas_tol: float = 1e-6
    beta1: float = 0.9
    beta2: float = 0.999
    adam_eps: float = 1e-8
    scale_eps_with_bias: bool = True
    num_restarts: int = 64
    seed: int = 0


@dataclass
class ALMResult:
    x: torch.Tensor          # (n,) best feasible point or None
    obj: float               # its FP64 objective (inf if none)
    viol: float              # its max violation


def qcqp_values(dev, x):
    """f_s(x) for x (N, n) on the device problem dev -> (N, S)."""
    S, n = dev.S, dev.n
    y = x @ dev.Q.reshape(S * n, n).t()
    return (x.unsqueeze(1) * y.view(-1, S, n)).sum(-1) + x @ dev.c.t() + dev.d


class _GhAlmConfig(ctypes.Structure):
    """gh_alm_config of include/gearhip.h."""
    _fields_ = [(k, ctypes.c_double) for k in ("rho0", "kappa", "max_rho", "inner_lr", "feas_tol",
                                               "beta1", "beta2", "adam_eps", "lr0")] + \
               [(k, ctypes.c_int32) for k in ("outer_max_iter", "inner_steps",
                                              "scale_eps_with_bias", "pad")]


class ALMPGDSolver:
    """backend "hip": the fused kernel gh_alm_pgd_f64 (one wavefront per
    restart); "torch": the same iteration as device tensor operations with
    autograd (kept as a second implementation the kernel is tested
    against)."""

    def __init__(self, cfg: ALMConfig = None, backend="hip"):
        self.cfg = cfg or ALMConfig()
        if backend not in ("hip", "torch"):
            raise ValueError(backend)
        self.backend = backend
        self._plans = {}

    def _plan(self, dev):
        key = id(dev)
        if key not in self._plans:
            from . import _lib
            nb = int(_lib.lib().gh_alm_plan_bytes(dev.n, dev.S))
            plan = torch.empty(nb // 8, dtype=torch.float64, device=dev.Q.device)
            _lib.check(_lib.lib().gh_alm_prepare(dev.n, dev.S, dev.Q.data_ptr(), plan.data_ptr(),
                                                 _lib.stream_ptr()), "gh_alm_prepare")
            self._plans[key] = (dev, plan)
        return self._plans[key][1]

    def _run_hip(self, dev, lo, hi, x, lr0):
        from . import _lib
        cfg = self.cfg
        N, n = x.shape
        c = _GhAlmConfig(cfg.rho0, cfg.kappa, cfg.max_rho, cfg.inner_lr, cfg.feas_tol, cfg.beta1,
                         cfg.beta2, cfg.adam_eps, lr0, cfg.outer_max_iter, cfg.inner_steps,
                         int(cfg.scale_eps_with_bias), 0)
        best_x = torch.empty_like(x)
        best_obj = torch.empty(N, dtype=torch.float64, device=x.device)
        _lib.check(_lib.lib().gh_alm_pgd_f64(N, n, dev.S, self._plan(dev).data_ptr(),
                                             dev.c.data_ptr(), dev.d.data_ptr(),
                                             dev.rhs.data_ptr(), lo.data_ptr(), hi.data_ptr(),
                                             x.data_ptr(), best_x.data_ptr(), best_obj.data_ptr(),
                                             ctypes.byref(c), _lib.stream_ptr()), "gh_alm_pgd_f64")
        return best_x, best_obj

    def solve_with_restarts(self, dev, x_L, x_U, num_restarts=None, x_init=None):
        """Search every box x_L/x_U (B, n) with R restarts each; returns the
        best feasible point over all of them."""
        cfg = self.cfg
        R = num_restarts or cfg.num_restarts
        B, n = x_L.shape
        gen = torch.Generator(device=x_L.device)
        gen.manual_seed(cfg.seed)
        lo = x_L.repeat_interleave(R, 0)
        hi = x_U.repeat_interleave(R, 0)
        width = hi - lo
        u = torch.rand(lo.shape, generator=gen, device=lo.device, dtype=lo.dtype)
        x = lo + u * width
        x.view(B, R, n)[:, 0] = (x_L + x_U) / 2 if x_init is None else x_init
        N = x.shape[0]
        if self.backend == "hip":
            lr0 = cfg.inner_lr * float(width.max()) if N else 0.0
            best_x, best_obj = self._run_hip(dev, lo.contiguous(), hi.contiguous(),
                                             x.contiguous(), lr0)
            if N == 0:
                return ALMResult(None, float("inf"), float("inf"))
            i = int(torch.argmin(best_obj))
            obj = float(best_obj[i])
            if obj == float("inf"):
                return ALMResult(None, float("inf"), float("inf"))
            return ALMResult(best_x[i].detach().clone(), obj, 0.0)
        m = dev.S - 1
        rhs = dev.rhs[1:]
        mu = torch.zeros(N, m, dtype=x.dtype, device=x.device)
        rho = torch.full((N,), cfg.rho0, dtype=x.dtype, device=x.device)
        lr_vec0 = cfg.inner_lr * width
        lr0 = cfg.inner_lr * float(width.max()) if N else 0.0
        best_obj = torch.full((N,), float("inf"), dtype=x.dtype, device=x.device)
        best_x = x.clone()
        prev_viol = torch.full((N,), float("inf"), dtype=x.dtype, device=x.device)
        for k in range(cfg.outer_max_iter):
            ex = torch.zeros_like(x)
            ex2 = torch.zeros_like(x)
            for t in range(cfg.inner_steps):
                xr = x.detach().requires_grad_(True)
                f = qcqp_values(dev, xr)
                g = f[:, 1:] - rhs
                pen = 0.5 * rho * torch.relu(g + mu / rho[:, None]).pow(2).sum(1)
                L = f[:, 0] + pen
                (grad,) = torch.autograd.grad(L.sum(), xr)
                grad = torch.nan_to_num(grad, nan=0.0, posinf=0.0, neginf=0.0)
                with torch.no_grad():
                    step = t + 1
                    bc1 = 1 - cfg.beta1 ** step
                    bc2 = 1 - cfg.beta2 ** step
                    ex.mul_(cfg.beta1).add_(grad, alpha=1 - cfg.beta1)
                    ex2.mul_(cfg.beta2).addcmul_(grad, grad, value=1 - cfg.beta2)
                    denom = (ex2.sqrt() / bc2 ** 0.5).add_(cfg.adam_eps)
                    decay = 1.0 - 0.9 * t / max(cfg.inner_steps, 1)
                    lr = lr0 * decay
                    max_step = lr_vec0 * decay
                    if cfg.scale_eps_with_bias:
                        max_step = max_step / bc1
                    upd = (-(ex / denom) * lr).clamp(min=-max_step, max=max_step)
                    x = torch.minimum(torch.maximum(x + upd, lo), hi)
            with torch.no_grad():
                f = qcqp_values(dev, x)
                g = f[:, 1:] - rhs
                viol = torch.relu(g).amax(1) if m else torch.zeros(N, dtype=x.dtype,
                                                                   device=x.device)
                feas = viol <= cfg.feas_tol
                better = feas & (f[:, 0] < best_obj)
                best_obj = torch.where(better, f[:, 0], best_obj)
                best_x = torch.where(better[:, None], x, best_x)
                mu = torch.clamp(mu + rho[:, None] * g, min=0.0)
                bad = viol > 0.25 * prev_viol
                rho = torch.where(bad, (rho * cfg.kappa).clamp(max=cfg.max_rho), rho)
                prev_viol = viol
        i = int(torch.argmin(best_obj))
        obj = float(best_obj[i])
        if obj == float("inf"):
            return ALMResult(None, float("inf"), float("inf"))
        xb = best_x[i].detach().clone()
        return ALMResult(xb, obj, 0.0)


def replay_float64(problem, x, tol=1e-6):
    """(objective, max violation, feasible) of x replayed in FP64 on the
    host (gear/report.py:114-141: a 1e-12 relative cushion on the
    tolerance)."""
    import numpy as np
    xv = np.asarray(x.detach().cpu().double().numpy() if hasattr(x, "detach") else x,
                    dtype=np.float64)[None]
    f = problem.values(xv)[0]
    g = f[1:] - problem.rhs[1:]
    viol = float(max(0.0, g.max())) if len(g) else 0.0
    feas = viol <= tol * (1 + 1e-12)
    return float(f[0]), viol, feas
