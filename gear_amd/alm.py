"""ALM-PGD primal search over boxes (batched restarts), on the device.

Follows the reference's HPR augmented Lagrangian with a projected, clipped
Adam inner loop (complete_verifier/nlp_solver/alm.py), each restart as the
reference runs one restart of one domain:
  * L(x) = phi(x) + 0.5 rho ||relu(g(x) + mu / rho)||^2  (:655-676);
  * AdamClipping_NLP.step: Adam moments, lr times m / denom, clipped per
    coordinate to inner_lr * width (divided by the first-moment bias
    correction when scale_eps_with_bias), projected onto the box (:134-160);
    lr = inner_lr * (largest edge of the restart's box), decayed linearly
    from 1.0 to 0.1 over the inner steps (:585-608, utils.py:121-147);
  * the inner loop stops at gradient norm <= eta_k (:764-776) or when L
    stops changing (:678-699);
  * best point, objective-patience and first-feasibility exits
    (:1650-1849), then the mu / rho / eps_k / eta_k schedule (:1851-1914):
    BaB mode by default (update_multipliers_always false, as the GEAR
    presets set it), root mode otherwise;
  * restart 0 starts at the box centre, the others uniformly in the box
    (:1337-1363); R = br_budget // B restarts per box when no count is
    given (resolve_alm_restart_count, :271-285).
The defaults are the GEAR presets' values (gear/presets/gear-f0.yaml:121,
137-155, 198).  The incumbent is the best point whose replayed FP64
violation is within feas_tol; its objective is replayed in FP64
(gear/report.py:114-141).
"""
from dataclasses import dataclass

import ctypes

import torch


@dataclass
class ALMConfig:
    rho0: float = 10.0
    kappa: float = 5.0
    max_rho: float = 1e6
    alpha: float = 0.5
    beta: float = 0.8
    eps_final: float = 1e-5
    eta_final: float = 1e-5
    outer_max_iter: int = 20
    inner_steps: int = 200
    inner_lr: float = 1.0
    inner_lr_outer_decay: float = 1.0
    inner_lr_min_factor: float = 1e-6
    inner_plateau_check: int = 10
    inner_plateau_rtol: float = 1e-3
    inner_plateau_patience: int = 2
    obj_patience: int = 5
    obj_improvement_tol: float = 1e-4
    first_feas_patience: int = 1
    update_multipliers_always: bool = False
    feas_tol: float = 1e-6
    beta1: float = 0.9
    beta2: float = 0.999
    adam_eps: float = 1e-8
    scale_eps_with_bias: bool = True
    num_restarts: int = 1000       # bab_num_restarts, when no shared budget is set
    br_budget: int = 20480         # bab_br_budget: B * R restarts per call
    # cap on the multiply-adds of a call at full length, restarts x rows x
    # n^2 x outer x inner iterations: the reference bounds an ALM call by
    # wall time instead (bab_timeout), which a fused launch cannot check
    work_budget: float = 2e12
    seed: int = 0


@dataclass
class ALMResult:
    x: torch.Tensor          # (n,) best feasible point or None
    obj: float               # its FP64 objective (inf if none)
    viol: float              # its max violation
    trace: torch.Tensor = None   # (N, outer_max_iter + 1, width), when asked for


def qcqp_values(dev, x):
    """f_s(x) for x (N, n) on the device problem dev -> (N, S) (a device
    problem with its own values(x), such as an expression program's, is
    evaluated by it)."""
    if hasattr(dev, "values"):
        return dev.values(x)
    S, n = dev.S, dev.n
    y = x @ dev.Q.reshape(S * n, n).t()
    return (x.unsqueeze(1) * y.view(-1, S, n)).sum(-1) + x @ dev.c.t() + dev.d


_DOUBLES = ("rho0", "kappa", "max_rho", "inner_lr", "feas_tol", "beta1", "beta2", "adam_eps",
            "alpha", "beta", "eps_final", "eta_final", "inner_lr_outer_decay",
            "inner_lr_min_factor", "obj_improvement_tol", "inner_plateau_rtol")
_INTS = ("outer_max_iter", "inner_steps", "scale_eps_with_bias", "update_multipliers_always",
         "obj_patience", "first_feas_patience", "inner_plateau_check", "inner_plateau_patience")


class _GhAlmConfig(ctypes.Structure):
    """gh_alm_config of include/gearhip.h."""
    _fields_ = [(k, ctypes.c_double) for k in _DOUBLES] + [(k, ctypes.c_int32) for k in _INTS]

    @classmethod
    def of(cls, cfg):
        return cls(*[float(getattr(cfg, k)) for k in _DOUBLES],
                   *[int(getattr(cfg, k)) for k in _INTS])


def restart_count(cfg, batch, explicit=None, step_work=0):
    """resolve_alm_restart_count (alm.py:271-285): an explicit count, else
    the shared budget br_budget // batch, else num_restarts; then at most
    work_budget / (batch * step_work * outer * inner) restarts per box
    (step_work: the multiply-adds of one restart's inner step)."""
    batch = max(1, int(batch))
    if explicit is not None:
        R = max(1, int(explicit))
    elif cfg.br_budget > 0:
        R = max(1, int(cfg.br_budget) // batch)
    else:
        R = max(1, int(cfg.num_restarts))
    if step_work > 0 and cfg.work_budget > 0:
        steps = max(1, cfg.outer_max_iter) * max(1, cfg.inner_steps)
        R = max(1, min(R, int(cfg.work_budget // (batch * step_work * steps))))
    return R


class ALMPGDSolver:
    """backend "hip": the fused kernel gh_alm_pgd_f64 (a restart per
    workgroup, its rows held in registers or LDS); "torch": the same
    iteration as device tensor operations with autograd, every restart's
    products in one GEMM per step (the plain baseline and a second
    implementation the kernel is tested against); "auto": the kernel up to
    n = 64, above it the GEMM form, which reads Q once per step for all
    restarts where the kernel reads it once per restart."""

    KERNEL_MAX_N = 64

    def __init__(self, cfg: ALMConfig = None, backend="auto"):
        self.cfg = cfg or ALMConfig()
        if backend not in ("hip", "torch", "auto"):
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

    def _run_hip(self, dev, lo, hi, x, want_trace):
        from . import _lib
        cfg = self.cfg
        N, n = x.shape
        best_x = torch.empty_like(x)
        best_obj = torch.empty(N, dtype=torch.float64, device=x.device)
        trace = None
        if want_trace:
            w = int(_lib.lib().gh_alm_trace_width(dev.S))
            trace = torch.empty(N, cfg.outer_max_iter + 1, w, dtype=torch.float64,
                                device=x.device)
        c = _GhAlmConfig.of(cfg)
        _lib.check(_lib.lib().gh_alm_pgd_f64(N, n, dev.S, self._plan(dev).data_ptr(),
                                             dev.c.data_ptr(), dev.d.data_ptr(),
                                             dev.rhs.data_ptr(), lo.data_ptr(), hi.data_ptr(),
                                             x.data_ptr(), best_x.data_ptr(), best_obj.data_ptr(),
                                             ctypes.byref(c), _lib.ptr(trace),
                                             _lib.stream_ptr()), "gh_alm_pgd_f64")
        return best_x, best_obj, trace

    def starts(self, x_L, x_U, R, x_init=None):
        """(lo, hi, x) (B R, n): each box repeated R times, restart 0 at the
        centre (or x_init), the others uniform in the box."""
        B, n = x_L.shape
        gen = torch.Generator(device=x_L.device)
        gen.manual_seed(self.cfg.seed)
        lo = x_L.repeat_interleave(R, 0).contiguous()
        hi = x_U.repeat_interleave(R, 0).contiguous()
        u = torch.rand(lo.shape, generator=gen, device=lo.device, dtype=lo.dtype)
        x = lo + u * (hi - lo)
        x.view(B, R, n)[:, 0] = (x_L + x_U) / 2 if x_init is None else x_init
        return lo, hi, x.contiguous()

    def solve_with_restarts(self, dev, x_L, x_U, num_restarts=None, x_init=None, trace=False):
        """Search every box x_L/x_U (B, n) with R restarts each; returns the
        best feasible point over all of them."""
        R = restart_count(self.cfg, x_L.shape[0], num_restarts,
                          0 if hasattr(dev, "values") else dev.S * dev.n * dev.n)
        lo, hi, x = self.starts(x_L, x_U, R, x_init)
        backend = self.backend
        if backend == "auto":
            backend = ("hip" if not hasattr(dev, "values") and dev.n <= self.KERNEL_MAX_N
                       else "torch")
        if backend == "hip":
            if hasattr(dev, "values"):
                raise ValueError("the fused ALM kernel covers QCQPs; use backend='torch'")
            best_x, best_obj, tr = self._run_hip(dev, lo, hi, x, trace)
        else:
            best_x, best_obj, _ = self._run_torch(dev, lo, hi, x)
            tr = None
        if x.shape[0] == 0:
            return ALMResult(None, float("inf"), float("inf"), tr)
        i = int(torch.argmin(best_obj))
        obj = float(best_obj[i])
        if obj == float("inf"):
            return ALMResult(None, float("inf"), float("inf"), tr)
        return ALMResult(best_x[i].detach().clone(), obj, 0.0, tr)

    def _run_torch(self, dev, lo, hi, x):
        """The kernel's iteration as op-by-op tensor code: restarts are
        batched, each with its own exits (masks).  Returns (best_x,
        best_obj, last iterates)."""
        cfg = self.cfg
        N, n = x.shape
        m = dev.S - 1
        rhs = dev.rhs[1:]
        dt, dv = x.dtype, x.device
        inf = torch.full((N,), float("inf"), dtype=dt, device=dv)
        width = hi - lo
        edge = width.amax(1).clamp(min=1e-10) if n else inf
        mu = torch.zeros(N, m, dtype=dt, device=dv)
        rho = torch.full((N,), cfg.rho0, dtype=dt, device=dv)
        rp = (lambda r, a: r.sqrt() if a == 0.5 else r.pow(a))
        eps, eta = 1.0 / rp(rho, cfg.alpha), 1.0 / rp(rho, cfg.beta)
        with torch.no_grad():
            f = qcqp_values(dev, x)
        viol = torch.relu(f[:, 1:] - rhs).amax(1) if m else torch.zeros(N, dtype=dt, device=dv)
        best, best_viol, best_x = f[:, 0].clone(), viol.clone(), x.clone()
        best_feas = viol <= cfg.feas_tol
        prev_best, glob_best, prev_viol = inf.clone(), inf.clone(), inf.clone()
        no_imp = torch.zeros(N, dtype=torch.int64, device=dv)
        plat = torch.full((N,), -1, dtype=torch.int64, device=dv)
        alive = torch.ones(N, dtype=torch.bool, device=dv)
        for k in range(cfg.outer_max_iter):
            if not bool(alive.any()):
                break
            scale = max(cfg.inner_lr_outer_decay ** k, cfg.inner_lr_min_factor)
            eff = cfg.inner_lr * max(scale, 0.0)
            base_lr = (eff * edge)[:, None]
            lr_vec = eff * width
            ex = torch.zeros_like(x)
            ex2 = torch.zeros_like(x)
            run = alive.clone()
            prev_L = torch.zeros(N, dtype=dt, device=dv)
            have_prev = False
            count = torch.zeros(N, dtype=torch.int64, device=dv)
            for t in range(cfg.inner_steps):
                if not bool(run.any()):
                    break
                xr = x.detach().requires_grad_(True)
                f = qcqp_values(dev, xr)
                g = f[:, 1:] - rhs
                phi = f[:, 0]
                if m:
                    pen = torch.relu(g + mu / rho[:, None]).pow(2).sum(1)
                    L = phi + 0.5 * rho * pen
                else:
                    L = phi
                ffin = torch.isfinite(phi) & torch.isfinite(L)
                if m:
                    ffin &= torch.isfinite(g).all(1)
                if cfg.inner_plateau_check > 0 and t > 0 and t % cfg.inner_plateau_check == 0:
                    with torch.no_grad():
                        cur = L - 0.5 / rho * (mu * mu).sum(1) if m else L.detach()
                        cur = cur.detach()
                        if have_prev:
                            rel = (cur - prev_L).abs() / (prev_L.abs() + 1e-8)
                            count = torch.where(rel < cfg.inner_plateau_rtol, count + 1, 0)
                            run &= count < cfg.inner_plateau_patience
                        prev_L, have_prev = cur, True
                (grad,) = torch.autograd.grad(L.sum(), xr)
                with torch.no_grad():
                    bad = ~torch.isfinite(grad).all(1) | ~ffin
                    grad = torch.nan_to_num(grad, nan=0.0, posinf=0.0, neginf=0.0)
                    step = t + 1
                    bc1 = 1 - cfg.beta1 ** step
                    bc2 = 1 - cfg.beta2 ** step
                    nex = ex * cfg.beta1 + (1 - cfg.beta1) * grad
                    nex2 = ex2 * cfg.beta2 + (1 - cfg.beta2) * grad * grad
                    denom = nex2.sqrt() / bc2 ** 0.5 + cfg.adam_eps
                    fac = 1.0 + (0.1 - 1.0) * min(1.0, t / max(cfg.inner_steps, 1))
                    max_step = lr_vec / bc1 if cfg.scale_eps_with_bias else lr_vec
                    upd = torch.maximum(torch.minimum(-(nex / denom) * (base_lr * fac), max_step),
                                        -max_step)
                    nx = torch.minimum(torch.maximum(x + upd, lo), hi)
                    sel = run[:, None]
                    x = torch.where(sel, nx, x)
                    ex = torch.where(sel, nex, ex)
                    ex2 = torch.where(sel, nex2, ex2)
                    gn = torch.where(bad, inf, grad.norm(dim=1))
                    run &= ~(gn <= eta)
            with torch.no_grad():
                f = qcqp_values(dev, x)
                g = f[:, 1:] - rhs
                phi = f[:, 0]
                viol = torch.relu(g).amax(1) if m else torch.zeros(N, dtype=dt, device=dv)
                feas = viol <= cfg.feas_tol
                improved = alive & feas & (~best_feas | (phi < best))
                better = alive & ~feas & ~best_feas & (viol < best_viol)
                upd = improved | better
                best = torch.where(upd, phi, best)
                best_viol = torch.where(upd, viol, best_viol)
                best_x = torch.where(upd[:, None], x, best_x)
                best_feas = best_feas | improved
                stop = torch.zeros(N, dtype=torch.bool, device=dv)
                if cfg.obj_patience >= 0:
                    imp = best < prev_best - cfg.obj_improvement_tol
                    no_imp = torch.where(alive, torch.where(imp, 0, no_imp + 1), no_imp)
                    prev_best = torch.where(alive, torch.minimum(prev_best, best), prev_best)
                    stop |= alive & (no_imp >= cfg.obj_patience) & best_feas
                if cfg.first_feas_patience >= 0:
                    act = alive & ~stop & best_feas
                    reset = act & ((plat < 0) | (best < glob_best - cfg.obj_improvement_tol))
                    plat = torch.where(reset, k, plat)
                    glob_best = torch.where(reset, best, glob_best)
                    stop |= act & ~reset & (k - plat >= cfg.first_feas_patience)
                go = alive & ~stop
                if cfg.update_multipliers_always:
                    moves = go
                else:
                    moves = go & (viol <= eps)
                mu = torch.where(moves[:, None], torch.clamp(mu + rho[:, None] * g, min=0.0), mu)
                if cfg.update_multipliers_always:
                    if k >= 1:
                        ratio = prev_viol / viol.clamp(min=1e-12)
                        inc = go & (ratio < 1.25) & (viol > cfg.feas_tol)
                        rho = torch.where(inc, (cfg.kappa * rho).clamp(max=cfg.max_rho), rho)
                    eps = torch.where(go, 1.0 / rp(rho, cfg.alpha), eps)
                    eta = torch.where(go, 1.0 / rp(rho, cfg.beta), eta)
                    prev_viol = torch.where(go, viol, prev_viol)
                else:
                    good = go & (viol <= eps)
                    badm = go & ~(viol <= eps)
                    eps = torch.where(good, eps / rho, eps)
                    eta = torch.where(good, eta / rho, eta)
                    inc = badm & (viol > 1e-3)
                    rho = torch.where(inc, (cfg.kappa * rho).clamp(max=cfg.max_rho), rho)
                    eps = torch.where(badm, 1.0 / rp(rho, cfg.alpha), eps)
                    eta = torch.where(badm, 1.0 / rp(rho, cfg.beta), eta)
                eps = torch.where(go, eps.clamp(min=cfg.eps_final), eps)
                eta = torch.where(go, eta.clamp(min=cfg.eta_final), eta)
                alive = go
        best_obj = torch.where(best_feas, best, inf)
        return best_x, best_obj, x


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
