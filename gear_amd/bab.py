"""Batched branch and bound over input boxes on one device.

One iteration follows the reference's reordered order, the one its GEAR
presets select (reorder_bab: true, compare_with_old_bounds: true;
complete_verifier/input_split/batch_branch_and_bound.py:275-652): sort the
frontier (least verified last) -> pick out the last 2 `batch` domains ->
bound them (CROWN, never below the key they were stored with) -> branch on
the widest variable -> bisect -> key each child by its parent's bound and
the parent's objective plane over the child box (:189-225) -> prune against
the thresholds -> append the survivors.  The regular order (bisect first,
then bound the children) is kept behind BaBConfig.reorder = False.  Then
the global lower bound is the frontier minimum capped at the discarded-
threshold minimum (:620-634), and the gap test is is_primal_dual_converged
with rel_tol (complete_verifier/solving_utils.py:17-49).  ALM-PGD refreshes the incumbent every `alm_interval` iterations
(nlp_in_bab.py:1078-1231).
"""
import time
from dataclasses import dataclass, field

import torch

from .alm import ALMConfig, ALMPGDSolver, replay_float64
from .crown import (BoundedQCQP, DeviceFrontier, branch_sb, child_keys, input_split_branching,
                    input_split_parallel)


def is_primal_dual_converged(P, D, abs_tol=1e-5, rel_tol=0.0):
    """solving_utils.py:17-49."""
    if not (abs(P) < float("inf") and abs(D) < float("inf")):
        return False
    gap = P - D
    if gap < -abs_tol:
        return False
    if gap <= abs_tol:
        return True
    return gap / max(1.0, abs(P), abs(D)) <= rel_tol


def cut_rows(lA, lbias, thr):
    """Cut rows (B, S (n + 1)) lA x + (lbias - thr) <= 0 of planes
    lA (B, S, n), lbias (B, S); a row with thr = +inf (the objective before
    an incumbent exists) is the inactive 0 x - 1 <= 0
    (construct_cut_constraints, complete_verifier/input_split/utils.py:
    113-129)."""
    B, S, n = lA.shape
    out = torch.empty(B, S * (n + 1), dtype=torch.float64, device=lA.device)
    act = torch.isfinite(thr)
    out[:, :S * n].view(B, S, n).copy_(torch.where(act[None, :, None], lA, 0.0))
    out[:, S * n:].copy_(torch.where(act[None, :], lbias - thr, -1.0))
    return out


def no_cuts_row(n, S, device):
    """The stored cuts of a domain without any: every row inactive."""
    nc = torch.zeros(1, S * (n + 1), dtype=torch.float64, device=device)
    nc[0, S * n:] = -1.0
    return nc


def reordered_step(f, bound, thr, nb, n, S, planes=True, lp=None, clip=False,
                   clip_max_rows=16, sync=True, in_place=False, branching="naive",
                   root_width=None):
    """One reordered B&B step on the last nb domains of frontier f
    (batch_branch_and_bound.py:395-560 with reorder_bab): bound them
    (CROWN; with clip 'complete' concretized over the box and the cuts they
    carry, :420-428), never below the key they were stored with
    (compare_with_old_bounds, bounding.py:361-363), [lp(lb, lA, lbias, L, U):
    the LP certificate merge], branch on the widest input and bisect,
    [clip the children by the parents' planes, :523-534], key each child
    by its parent's bound and objective plane (:189-225), prune and append
    (children of a verified parent inherit its verified rows, so this one
    prune drops them).  bound(L, U, return_A=, plane_rows=) is the bound
    of the served form (FP64 or FP32-sound).  in_place: bound(..., rows=)
    accepts a row index, so without clip and LP the picked domains are
    bound and split where the frontier stores them (no gather).
    branching 'sb': the split input comes from the smart-branching score of
    the domains' planes (branch_sb, root_width for its exhaustive guard)
    instead of the widest input; it reads every row's plane."""
    sb = branching == "sb" and planes
    if branching not in ("naive", "sb"):
        raise ValueError(f"branching {branching!r}")
    if in_place and not clip and lp is None and f.in_place_room(nb):
        idx, pkey = f.pick_rows(nb)
        L, U = f.storage()
        if planes:
            lb, lA, lbias = bound(L, U, return_A=True, plane_rows=S if sb else 1, rows=idx)
        else:
            lb, lA, lbias = bound(L, U, rows=idx), None, None
        lb = torch.maximum(lb, pkey)
        choice = branch_sb(L, U, lb, thr, lA, root_width=root_width, rows=idx) if sb else None
        return f.split_append(lb, L, U, thr, lA, lbias, sync=sync, rows=idx, idx=choice)
    pL, pU, pkey = f.pick_out(nb)
    if planes:
        # the LP, the cuts and smart branching read every row's plane, the
        # child keys the objective's
        lb, lA, lbias = bound(pL, pU, return_A=True, plane_rows=S if (lp or clip or sb) else 1)
    else:
        lb, lA, lbias = bound(pL, pU), None, None
    cuts = None
    if clip:
        from .concretize import constrained_bound
        stored = f.pick_payload(nb)
        constrained_bound(lA, lbias, stored[:, :S * n].view(nb, S, n), stored[:, S * n:],
                          pL, pU, lb, max_rows=clip_max_rows)
        # the domains' planes become their children's cuts
        cuts = cut_rows(lA, lbias, thr)
    lb = torch.maximum(lb, pkey)
    if lp is not None and lA is not None:
        lp(lb, lA, lbias, pL, pU)
    choice = branch_sb(pL, pU, lb, thr, lA, root_width=root_width) if sb else None
    if not clip:
        # branch, bisect, key, prune and append without materializing the
        # children (the same frontier as branch_split_keys + filter_append)
        return f.split_append(lb, pL, pU, thr, lA, lbias, sync=sync, idx=choice)
    from .concretize import clip_domains
    idx = (choice.long().unsqueeze(1) if sb else
           input_split_branching(None, None, pL, pU, None, None, "naive"))
    cL, cU, _ = input_split_parallel(pL, pU, i_idx=idx)
    clip_domains(lA, lbias, thr, cL, cU, out=(cL, cU))
    keys = child_keys(lb, lA, lbias, cL, cU)
    return f.filter_append(keys, thr, cL, cU, sync=sync, payload=cuts)


@dataclass
class BaBConfig:
    batch: int = 16384             # 2 batch domains bounded per iteration
    cap: int = 1 << 18             # frontier capacity
    rel_tol: float = 1e-4
    abs_tol: float = 1e-5
    constr_viol_tol: float = 1e-6
    max_iters: int = 1000
    time_limit: float = 60.0
    # ALM-PGD in the loop (in_bab of the GEAR presets, nlp_in_bab.py:
    # 1078-1231): every alm_interval iterations on the alm_boxes least
    # verified domains, the interval doubling (to alm_max_interval) after
    # alm_interval_patience calls without a better incumbent and reset by
    # one; 0: ALM only at the root
    frontier_slack: int = None     # physical rows past cap (None: the frontier's default)
    # split input: 'naive' (widest) or 'sb' (the GEAR presets' smart
    # branching, branching_heuristics.py:25-125, with its exhaustive guard
    # against the root widths)
    branching: str = "naive"
    alm_interval: int = 5
    alm_max_interval: int = 20
    alm_interval_patience: int = 3
    alm_boxes: int = 64
    sort_interval: int = 1         # sort the frontier every k iterations (0: never)
    # LP dual certificate on the bounded domains' CROWN planes: PDHG
    # iterations, 0 to skip (bounding.py:489-520, dual_lp_pdhg_iters)
    dual_lp_iters: int = 0
    # the reordered loop of the GEAR presets (reorder_bab: true): bound the
    # picked domains, split them, give the children keys from the parents'
    # planes; False: the regular loop (split, then bound the children)
    reorder: bool = True
    # clip_type 'complete' of the GEAR presets (clip_input_domain): each
    # domain stores its parent's planes as cuts, its bound is concretized
    # over the box and those cuts, and children are clipped by the parent's
    # planes (batch_branch_and_bound.py:420-428, 523-534; presets
    # gear-f0.yaml:466-472); at most clip_max_rows binding cuts per domain
    # (concretize_max_constraint_rows).  "none": box concretization only.
    clip: str = "none"
    clip_max_rows: int = 16
    alm: ALMConfig = field(default_factory=ALMConfig)


@dataclass
class BaBStats:
    iterations: int = 0
    boxes_bounded: int = 0
    bound_time: float = 0.0
    wall: float = 0.0
    stop: str = ""               # converged, empty, iterations, time, frontier_full
    history: list = field(default_factory=list)


class QCQPBranchAndBound:
    """Device-resident B&B over one QCQP instance."""

    def __init__(self, problem, cfg: BaBConfig = None, device="cuda", precision="fp64"):
        self._init_common(problem, cfg or BaBConfig(),
                          BoundedQCQP(problem, device, precision=precision))

    def _init_common(self, problem, cfg, net):
        """State shared by the loops: the bound `net` (bound(L, U, return_A,
        plane_rows, rows) and the device problem net.dev), the frontier, the
        thresholds and the ALM."""
        self.p = problem
        self.cfg = cfg
        self.net = net
        self.device = self.net.device
        if self.cfg.clip not in ("none", "complete"):
            raise ValueError(f"clip {self.cfg.clip!r}: 'none' or 'complete'")
        self.clip = self.cfg.clip == "complete"
        if self.clip and (not self.cfg.reorder or self.net.large or problem.n > 64):
            raise ValueError("clip 'complete' runs in the reordered loop with n <= 64")
        n, S = problem.n, problem.S
        self.frontier = DeviceFrontier(n, S, self.cfg.cap, self.device,
                                       payload=S * (n + 1) if self.clip else 0,
                                       slack=self.cfg.frontier_slack)
        # the cut rows of a domain without any: inactive (0 x - 1 <= 0)
        self._no_cuts = no_cuts_row(n, S, self.device) if self.clip else None
        self.alm = ALMPGDSolver(self.cfg.alm)
        self.P = float("inf")
        self.x_best = None
        self.D = -float("inf")
        self.stats = BaBStats()
        thr = torch.as_tensor(problem.rhs, dtype=torch.float64).clone()
        thr[0] = float("inf")
        thr[1:] += self.cfg.constr_viol_tol
        self.thr = thr.to(self.device)
        self._glb = torch.empty(1, dtype=torch.float64, device=self.device)
        self.root_width = torch.as_tensor(problem.xu, dtype=torch.float64, device=self.device) - \
            torch.as_tensor(problem.xl, dtype=torch.float64, device=self.device)

    def set_incumbent(self, obj, x):
        if obj < self.P:
            self.P, self.x_best = obj, x
            self.thr[0] = obj

    def run_alm(self, x_L, x_U, restarts=None):
        res = self.alm.solve_with_restarts(self.net.dev, x_L, x_U, num_restarts=restarts)
        if res.x is not None:
            obj, viol, feas = replay_float64(self.p, res.x, self.cfg.constr_viol_tol)
            if feas:
                self.set_incumbent(obj, res.x)
        return res

    def root(self, alm=True):
        L = torch.as_tensor(self.p.xl, dtype=torch.float64, device=self.device)[None]
        U = torch.as_tensor(self.p.xu, dtype=torch.float64, device=self.device)[None]
        if alm:
            self.run_alm(L, U)
        lb = self.net.bound(L, U)
        self.frontier.filter_append(lb, self.thr, L, U, payload=self._no_cuts)
        self.update_dual()

    def presplit(self, leaves, alm=True):
        """Start from the root box bisected on the widest variable level by
        level into `leaves` boxes (a power of two), all bound and filtered in
        one batch -- the presplit-domains start of the reference
        (batch_branch_and_bound.py:2964-3061)."""
        L = torch.as_tensor(self.p.xl, dtype=torch.float64, device=self.device)[None]
        U = torch.as_tensor(self.p.xu, dtype=torch.float64, device=self.device)[None]
        if alm:
            self.run_alm(L, U)
        while L.shape[0] < leaves:
            idx = input_split_branching(self.net, None, L, U, None, None, "naive")
            L, U, _ = input_split_parallel(L, U, i_idx=idx)
        lb = self.net.bound(L, U)
        self.frontier.filter_append(lb, self.thr, L, U, payload=self._no_cuts)
        self.stats.boxes_bounded += L.shape[0]
        self.update_dual()

    def update_dual(self):
        if len(self.frontier) == 0:
            self.D = float(self.frontier.disc_min.item())
            if self.D == float("inf"):
                self.D = self.P
        else:
            self.D = float(self.frontier.global_lb(self._glb).item())
        return self.D

    def step(self, batch=None, sync=True):
        """One iteration; returns the number of domains bounded."""
        if self.cfg.reorder:
            return self.step_reordered(batch, sync)
        return self.step_regular(batch, sync)

    def step_regular(self, batch=None, sync=True):
        """Sort, pick `batch` parents, split them, bound the 2 batch
        children, prune and append (batch_branch_and_bound.py:275-652)."""
        if self.cfg.branching != "naive":
            raise ValueError("smart branching runs in the reordered loop (it reads the "
                             "picked domains' planes)")
        f = self.frontier
        si = self.cfg.sort_interval
        if si > 0 and self.stats.iterations % si == 0:
            f.sort(self.thr, tag=self.P)
        # children (2 nb) go where the nb parents were plus free rows
        nb = min(batch or self.cfg.batch, len(f), f.cap - len(f))
        if nb <= 0:
            return 0
        pL, pU, _ = f.pick_out(nb)
        idx = input_split_branching(self.net, None, pL, pU, None, None, "naive")
        cL, cU, _ = input_split_parallel(pL, pU, i_idx=idx)
        if self.cfg.dual_lp_iters > 0 and not self.net.large and self.P < float("inf"):
            lb, lA, lbias = self.net.bound(cL, cU, return_A=True)
            self.raise_by_lp(lb, lA, lbias, cL, cU)
        else:
            lb = self.net.bound(cL, cU)
        f.filter_append(lb, self.thr, cL, cU, sync=sync)
        self.stats.iterations += 1
        self.stats.boxes_bounded += 2 * nb
        return 2 * nb

    def step_reordered(self, batch=None, sync=True):
        """The reordered iteration (batch_branch_and_bound.py:500-560 with
        reorder_bab): sort, then reordered_step on 2 `batch` domains.  Like
        the reference (batch_branch_and_bound.py:1387-1389) it picks
        2 `batch` domains, so an iteration bounds as many domains as the
        regular one."""
        f = self.frontier
        si = self.cfg.sort_interval
        if si > 0 and self.stats.iterations % si == 0:
            f.sort(self.thr, tag=self.P)
        want = 2 * (batch or self.cfg.batch)
        nb = min(want, len(f), f.cap - len(f))
        if nb < want and f.unsynced:
            # dead entries of unsynced appends take room: drop them (a sync)
            f.compact_order()
            if si > 0:
                f.sort(self.thr, tag=self.P)
            nb = min(want, len(f), f.cap - len(f))
        if nb <= 0:
            return 0
        lp = self.raise_by_lp if (self.cfg.dual_lp_iters > 0 and self.P < float("inf")) else None
        large = getattr(self.net, "large", False)
        reordered_step(f, self.net.bound, self.thr, nb, self.p.n, self.p.S, planes=not large,
                       lp=lp, clip=self.clip, clip_max_rows=self.cfg.clip_max_rows, sync=sync,
                       in_place=not large, branching=self.cfg.branching,
                       root_width=self.root_width)
        self.stats.iterations += 1
        self.stats.boxes_bounded += nb
        return nb

    def cut_rows(self, lA, lbias):
        return cut_rows(lA, lbias, self.thr)

    def raise_by_lp(self, lb, lA, lbias, cL, cU):
        """Solving-mode merge of the LP certificate into the objective row
        (bounding.py:698-723 _merge_dual_ascent_bound): a child whose
        certified LP value reaches P - near_tol takes it as its bound, so
        the prune drops it; other children keep their CROWN bound."""
        from .lp import qcqp_lp_bound
        cfg = self.cfg
        lp = qcqp_lp_bound(lA, lbias, self.thr, cL, cU, iters=cfg.dual_lp_iters)
        near = 0.5 * max(cfg.abs_tol, cfg.rel_tol * min(max(abs(self.P), 1.0), 1e300))
        hit = lp >= (self.P - near)
        lb[:, 0] = torch.where(hit, torch.maximum(lb[:, 0], lp), lb[:, 0])
        return lb

    def solve(self):
        cfg = self.cfg
        t0 = time.time()
        self.root()
        every, fails = cfg.alm_interval, 0
        self.stats.stop = "iterations"
        while len(self.frontier) > 0 and self.stats.iterations < cfg.max_iters:
            if is_primal_dual_converged(self.P, self.D, cfg.abs_tol, cfg.rel_tol):
                self.stats.stop = "converged"
                break
            if time.time() - t0 > cfg.time_limit:
                self.stats.stop = "time"
                break
            if cfg.alm_interval and self.stats.iterations > 0 and \
                    self.stats.iterations % every == 0:
                k = min(len(self.frontier), cfg.alm_boxes)
                s = len(self.frontier) - k
                wL, wU = self.frontier.rows(s, s + k)[:2]
                P0 = self.P
                self.run_alm(wL, wU)
                # adaptive interval (nlp_in_bab.py:1200-1215)
                if self.P < P0:
                    every, fails = cfg.alm_interval, 0
                else:
                    fails += 1
                    if fails >= cfg.alm_interval_patience:
                        every, fails = min(every * 2, cfg.alm_max_interval), 0
            if self.step() == 0:
                # no room for one batch of children: the frontier is full
                self.stats.stop = "frontier_full"
                break
            self.update_dual()
            self.stats.history.append((time.time() - t0, self.P, self.D, len(self.frontier)))
        self.stats.wall = time.time() - t0
        if len(self.frontier) == 0:
            self.stats.stop = "empty"
            self.D = min(self.P, float(self.frontier.disc_min.item()))
        return self
