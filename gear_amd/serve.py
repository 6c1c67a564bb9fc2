"""Serving loop: QCQP instances arrive over time and are solved to a
relative gap on one device with continuous batching.

Each admitted instance is a session with its own HIP stream and a
device-resident B&B frontier (the reference's per-instance loop,
complete_verifier/input_split/batch_branch_and_bound.py:275-652, with the
frontier of branching_domains.py kept on the device).  Admission is by an
HBM budget: an arriving instance waits until its estimated footprint fits
beside the sessions already on the device.  Every scheduler tick launches
one B&B step of every active session on its stream, without host syncs,
then reads each session's survivor count and global lower bound once the
device work is done, so the sessions' kernels overlap on the device and new
arrivals join the next tick.

An instance's trajectory depends only on the instance: its batch size is a
function of (n, m), its ALM restarts are seeded, and the kernels are
deterministic for a given batch.  So its certificate is the same alone and
in any mix (tests/test_gpu_serve.py).  A session ends when
is_primal_dual_converged(P, D, abs_tol, rel_tol) holds
(complete_verifier/solving_utils.py:17-49), when its frontier empties
(the instance is solved, or infeasible if P = +inf), or at its time limit.
P is the incumbent replayed in FP64 against the constraints at 1e-6
(gear/report.py:114-141); D is the sound lower bound of the frontier.
"""
import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from .alm import ALMConfig, ALMPGDSolver, replay_float64
from .bab import is_primal_dual_converged, no_cuts_row, reordered_step
from .crown import (BoundedQCQP, DeviceFrontier, input_split_branching,
                    input_split_parallel)
from .problems import flops_per_box, qcqp_dense


@dataclass
class Request:
    rid: int
    arrival: float        # seconds after the stream starts
    problem: object


def poisson_stream(count, rate, seed=0, n_median=10.0, n_sigma=0.8, n_min=5, n_max=1000,
                   m_ratio=0.4):
    """`count` requests with exponential inter-arrival times (mean 1 / rate);
    instances from the config-2 generator (dense QCQP, coefficients
    N(0,1)/sqrt(n), box [-1, 1]^n) with n = clip(round(lognormal(ln
    n_median, n_sigma)), n_min, n_max) and m = max(1, round(m_ratio n))."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=count))
    t = t - t[0]
    out = []
    for k in range(count):
        n = int(np.clip(np.round(rng.lognormal(math.log(n_median), n_sigma)), n_min, n_max))
        m = max(1, int(round(m_ratio * n)))
        out.append(Request(k, float(t[k]), qcqp_dense(n, m, seed=10_000 * (seed + 1) + k)))
    return out


def config5_stream(count, rate, seed=6, arrival_seed=11, m=20):
    """The serving workload: `count` requests with exponential inter-arrival
    times (mean 1 / rate); instance i is the config-2 generator's dense
    QCQP (coefficients N(0,1)/sqrt(n), box [-1, 1]^n) with n =
    clip(round(lognormal(ln 60, 1.0)), 5, 1000), m = 20 constraints and seed
    seed + i."""
    from .problems import serving_sizes
    rng = np.random.default_rng(arrival_seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=count))
    t = t - t[0]
    ns = serving_sizes(count, seed=seed)
    return [Request(k, float(t[k]), qcqp_dense(int(ns[k]), m, seed=seed + k))
            for k in range(count)]


@dataclass
class ServeConfig:
    precision: str = "fp32s"      # bound precision of the served form
    hbm_budget: int = 128 << 30   # bytes the sessions may hold together (of 288 GB)
    rel_tol: float = 1e-4
    abs_tol: float = 1e-5
    constr_viol_tol: float = 1e-6
    time_limit: float = 10.0      # per instance, from admission
    cap: int = 1 << 21            # frontier capacity per instance (rows)
    session_bytes: int = 2 << 30  # and at most this many bytes of frontier
    step_flops: float = 2e10      # bound work per step that sets the batch size
    max_batch: int = 8192
    min_batch: int = 16
    alm_restarts: int = None        # None: the ALM's shared restart budget
    # the presets' ALM with a smaller work cap per call, for latency
    alm: ALMConfig = field(default_factory=lambda: ALMConfig(work_budget=2e11))
    # ALM on the least verified frontier boxes every alm_interval steps, and
    # every alm_interval_noinc steps while there is no incumbent
    # (run_bab_alm_step, complete_verifier/input_split/nlp_in_bab.py:1078-1231)
    alm_interval: int = 50
    alm_interval_noinc: int = 5
    alm_boxes: int = 8
    alm_box_restarts: int = None    # None: the ALM's shared restart budget
    # FP64 LP dual certificate on the bounded domains' FP32-sound planes,
    # merged in solving mode (bounding.py:489-520, 698-723); PDHG
    # iterations, 0 to skip
    dual_lp_iters: int = 300
    # the reordered step of the GEAR presets (reorder_bab: true): bound the
    # picked domains and key their children by the parents' planes
    reorder: bool = True
    # clip_type of the presets' clip_input_domain: "complete" stores the
    # parent's planes as cuts with each domain, concretizes over them and
    # clips children (n <= 64 instances); "none": box concretization
    clip: str = "none"
    clip_max_rows: int = 16


def batch_size(n, S, cfg):
    """Parents per step: a power of two giving about cfg.step_flops of bound
    work (2 children per parent), within [min_batch, max_batch]."""
    b = cfg.step_flops / (2.0 * flops_per_box(n, S))
    b = 1 << max(0, int(math.floor(math.log2(max(b, 1.0)))))
    return int(min(cfg.max_batch, max(cfg.min_batch, b)))


def _row_bytes(n, S, cfg):
    row = (2 * n + S) * 8
    if cfg.clip == "complete" and n <= 64:
        row += S * (n + 1) * 8             # the stored cut rows
    return row


def frontier_cap(n, S, cfg):
    """Frontier rows of a session: at most cfg.cap rows and cfg.session_bytes
    bytes, at least 4096 rows."""
    row = _row_bytes(n, S, cfg)
    return int(max(4096, min(cfg.cap, cfg.session_bytes // row)))


def footprint(n, S, cfg):
    """Device bytes a session holds: frontier, one step's children and
    bound / filter workspaces and planes, program data and plans."""
    row = _row_bytes(n, S, cfg)
    nb = batch_size(n, S, cfg)
    child = 2 * nb * (row + 64) + 2 * nb * 8
    if cfg.reorder:                    # 2 nb parents and their 4 nb children
        child = 2 * child + 2 * nb * (S * n + S) * 8
    prog = S * n * n * 8 * 3 + S * (n + 2) * 8
    large = 0
    if n > 64:
        large = S * n * n * 4 * 2 + 2 * nb * ((3 * n + 5 * S * n) * 4 + 4 * S * 8 + 8)
    cap = frontier_cap(n, S, cfg)
    # physical rows (cap and half again of slack), the order array, and the
    # transient second copy while live rows are compacted
    front = (cap + max(cap // 2, 4096)) * row * 2 + cap * 16
    return front + child + prog + large + (8 << 20)


class Session:
    """One instance's B&B on its own stream."""

    def __init__(self, req: Request, cfg: ServeConfig, device, bound_fn=None, alm_backend="auto"):
        self.req = req
        self.cfg = cfg
        self.p = p = req.problem
        self.device = device
        self.stream = torch.cuda.Stream(device=device)
        self.batch = batch_size(p.n, p.S, cfg)
        self.bytes = footprint(p.n, p.S, cfg)
        self.alm_backend = alm_backend
        self.bound_fn = bound_fn
        self.P = math.inf
        self.x_best = None
        self.D = -math.inf
        self.steps = 0
        self.boxes = 0
        self.status = "pending"
        self.t_admit = self.t_done = None
        self._pending = False

    # -- setup: data, incumbent, root split (on the session's stream)
    def start(self, now):
        cfg, p = self.cfg, self.p
        self.t_admit = now
        with torch.cuda.stream(self.stream):
            self.planes = False
            if self.bound_fn is None:
                net = BoundedQCQP(p, self.device, precision=cfg.precision)
                self.bound, self.dev = net.bound, net.dev
                self.planes = not net.large
            else:
                self.dev = p.to_device(self.device)
                self.bound = self.bound_fn(self.dev)
                # the same algorithm as the served form: planes (child keys,
                # LP certificate) where the served bound has them
                self.planes = p.n <= 64
            self.clip = cfg.clip == "complete" and self.planes and p.n <= 64
            self.frontier = DeviceFrontier(p.n, p.S, frontier_cap(p.n, p.S, cfg), self.device,
                                           payload=p.S * (p.n + 1) if self.clip else 0)
            thr = torch.as_tensor(p.rhs, dtype=torch.float64).clone()
            thr[0] = math.inf
            thr[1:] += cfg.constr_viol_tol
            self.thr = thr.to(self.device)
            L = torch.as_tensor(p.xl, dtype=torch.float64, device=self.device)[None]
            U = torch.as_tensor(p.xu, dtype=torch.float64, device=self.device)[None]
            self.alm = ALMPGDSolver(cfg.alm, backend=self.alm_backend)
            self._incumbent(L, U, cfg.alm_restarts)
            # root split into `batch` leaves (presplit start)
            while L.shape[0] < self.batch:
                idx = input_split_branching(None, None, L, U, None, None, "naive")
                L, U, _ = input_split_parallel(L, U, i_idx=idx)
            lb = self.bound(L, U)
            self.frontier.filter_append(lb, self.thr, L, U, sync=False,
                                        payload=no_cuts_row(p.n, p.S, self.device)
                                        if self.clip else None)
            self.boxes += L.shape[0]
            self._host = torch.empty(3, dtype=torch.float64, pin_memory=True)
            self._glb = torch.empty(1, dtype=torch.float64, device=self.device)
            self._total = torch.empty(1, dtype=torch.int64, device=self.device)
            self._host_total = torch.empty(1, dtype=torch.int64, pin_memory=True)
            self._total.copy_(self.frontier.total)
            self._host_total.copy_(self._total, non_blocking=True)
            self._host[1:2].copy_(self.frontier.disc_min, non_blocking=True)
            self._host[0] = -math.inf
            self._pending = True
            self._have_glb = False
            self._nb = 0
            self._bounded = 0
            self.event = torch.cuda.Event()
            self.event.record(self.stream)
        self.status = "active"

    def _incumbent(self, L, U, restarts):
        """ALM-PGD over boxes L, U; a point whose FP64 replay is feasible and
        better becomes the incumbent (and the objective threshold)."""
        res = self.alm.solve_with_restarts(self.dev, L, U, num_restarts=restarts)
        if res.x is not None:
            obj, viol, feas = replay_float64(self.p, res.x, self.cfg.constr_viol_tol)
            if feas and obj < self.P:
                self.P, self.x_best = obj, res.x
                self.thr[0] = obj

    def _raise_by_lp(self, lb, lA, lbias, cL, cU):
        """Solving-mode merge of the LP certificate (bounding.py:698-723):
        children whose certified LP value reaches P - near_tol take it."""
        from .lp import qcqp_lp_bound
        cfg = self.cfg
        lp = qcqp_lp_bound(lA, lbias, self.thr, cL, cU, iters=cfg.dual_lp_iters)
        near = 0.5 * max(cfg.abs_tol, cfg.rel_tol * min(max(abs(self.P), 1.0), 1e300))
        hit = lp >= (self.P - near)
        lb[:, 0] = torch.where(hit, torch.maximum(lb[:, 0], lp), lb[:, 0])

    def _lp(self):
        return self.planes and self.cfg.dual_lp_iters > 0 and self.P < math.inf

    def _step_regular(self, f):
        """Pick `batch` parents, bisect, bound the children [LP], prune and
        append; returns the parents picked."""
        nb = min(self.batch, len(f), f.cap - len(f))
        if nb > 0:
            pL, pU, _ = f.pick_out(nb)
            idx = input_split_branching(None, None, pL, pU, None, None, "naive")
            cL, cU, _ = input_split_parallel(pL, pU, i_idx=idx)
            if self._lp():
                lb, lA, lbias = self.bound(cL, cU, return_A=True)
                self._raise_by_lp(lb, lA, lbias, cL, cU)
            else:
                lb = self.bound(cL, cU)
            f.filter_append(lb, self.thr, cL, cU, sync=False)
            self._bounded = 2 * nb
        return nb

    def _step_reordered(self, f):
        """Pick 2 `batch` domains, bound them (never below their keys) [LP],
        bisect, key the children by the parents' planes, prune and append;
        returns the domains picked."""
        nb = min(2 * self.batch, len(f), f.cap - len(f))
        if nb > 0:
            # the library bound takes a row index; a given bound_fn may not
            reordered_step(f, self.bound, self.thr, nb, self.p.n, self.p.S, planes=self.planes,
                           lp=self._raise_by_lp if self._lp() else None, clip=self.clip,
                           clip_max_rows=self.cfg.clip_max_rows, sync=False,
                           in_place=self.planes)
            self._bounded = nb
        return nb

    def launch(self):
        """Enqueue the global bound of the current frontier, the frontier
        sort and one step.  An append not yet accounted (the presplit of a
        session started in this tick) is committed first: the global bound
        and the pick must see every survivor, or a frontier that still holds
        boxes would read as empty and the bound as the incumbent."""
        f = self.frontier
        cfg = self.cfg
        if self._pending:
            self.event.synchronize()
            f.commit(int(self._host_total[0]))
            self._pending = False
            if self._nb:
                self.steps += 1
                self.boxes += self._bounded
            self._nb = 0
        with torch.cuda.stream(self.stream):
            f.global_lb(self._glb)
            f.sort(self.thr, tag=self.P)
            self._host[0:1].copy_(self._glb, non_blocking=True)
            self._have_glb = True
            if cfg.reorder:
                nb = self._step_reordered(f)
            else:
                nb = self._step_regular(f)
            self._nb = nb
            if nb > 0:
                self._total.copy_(f.total)
                self._host_total.copy_(self._total, non_blocking=True)
                self._pending = True
            self._host[1:2].copy_(f.disc_min, non_blocking=True)
            self.event.record(self.stream)

    def finish(self, now):
        """Read the step's results; returns True when the session is done."""
        self.event.synchronize()
        f = self.frontier
        if self._pending:
            f.commit(int(self._host_total[0]))
            self._pending = False
            if self._nb:
                self.steps += 1
                self.boxes += self._bounded
        disc = float(self._host[1])
        if self._have_glb:
            self.D = float(self._host[0])
        cfg = self.cfg
        if is_primal_dual_converged(self.P, self.D, cfg.abs_tol, cfg.rel_tol):
            self.status = "converged"
        elif len(f) == 0:
            # every domain verified: the bound closes on the incumbent
            self.D = min(self.P, disc)
            self.status = "converged" if self.P < math.inf else "infeasible"
        elif now - self.t_admit > cfg.time_limit or len(f) >= f.cap:
            self.status = "timeout"
        if self.status != "active":
            self.t_done = now
            return True
        every = cfg.alm_interval_noinc if self.P == math.inf else cfg.alm_interval
        if every > 0 and self.steps > 0 and self.steps % every == 0 and self._nb:
            k = min(cfg.alm_boxes, len(f))
            s0 = len(f) - k
            with torch.cuda.stream(self.stream):
                wL, wU = f.rows(s0, s0 + k)[:2]
                self._incumbent(wL, wU, cfg.alm_box_restarts)
        return False

    def result(self):
        r = self.req
        done = self.t_done if self.t_done is not None else math.inf
        return {"rid": r.rid, "n": self.p.n, "m": self.p.m, "arrival": r.arrival,
                "admit": self.t_admit, "done": done, "status": self.status,
                "time_to_gap": (done - r.arrival) if self.status in ("converged", "infeasible")
                else math.inf,
                "P": self.P, "D": self.D, "steps": self.steps, "boxes": self.boxes,
                "batch": self.batch}


def _percentile(vals, q):
    v = np.sort(np.asarray(vals, dtype=np.float64))
    if v.size == 0:
        return math.nan
    return float(v[min(v.size - 1, int(math.ceil(q / 100.0 * v.size)) - 1)])


def summarize(results, wall):
    """Percentiles of the time from arrival to certificate over all
    requests; a request without one (time limit, or never started) counts
    with its time until it was given up, so those percentiles are lower
    bounds and `censored` says how many such requests there are."""
    ttg = []
    for r in results:
        if r["status"] in ("converged", "infeasible"):
            ttg.append(r["done"] - r["arrival"])
        else:
            end = r["done"] if r["done"] < math.inf else wall
            ttg.append(max(0.0, end - r["arrival"]))
    boxes = sum(r["boxes"] for r in results)
    return {"requests": len(results),
            "converged": sum(r["status"] == "converged" for r in results),
            "timeout": sum(r["status"] == "timeout" for r in results),
            "infeasible": sum(r["status"] == "infeasible" for r in results),
            "censored": sum(r["status"] not in ("converged", "infeasible") for r in results),
            "p50_time_to_gap_s": _percentile(ttg, 50),
            "p99_time_to_gap_s": _percentile(ttg, 99),
            "subdomain_bounds_per_s": boxes / wall if wall > 0 else math.nan,
            "boxes": boxes, "wall_s": wall}


def serve(requests, cfg: ServeConfig = None, device="cuda", realtime=True, log=None):
    """Run the stream; returns (per-request results, summary).  With
    realtime, requests are admitted no earlier than their arrival time on
    the wall clock; otherwise all are available at once."""
    cfg = cfg or ServeConfig()
    device = torch.device(device)
    queue = sorted(requests, key=lambda r: r.arrival)
    active, results = [], []
    used = 0
    t0 = time.perf_counter()
    clock = (lambda: time.perf_counter() - t0) if realtime else (lambda: math.inf)
    wall_clock = lambda: time.perf_counter() - t0
    while queue or active:
        now = clock()
        # admission by HBM budget, in arrival order
        while queue and queue[0].arrival <= now:
            need = footprint(queue[0].problem.n, queue[0].problem.S, cfg)
            if active and used + need > cfg.hbm_budget:
                break
            req = queue.pop(0)
            s = Session(req, cfg, device)
            s.start(wall_clock())
            used += s.bytes
            active.append(s)
            if log:
                log(f"admit {req.rid} n={req.problem.n} m={req.problem.m} batch={s.batch} "
                    f"t={wall_clock():.3f}")
        if not active:
            if queue:
                time.sleep(max(0.0, min(queue[0].arrival - clock(), 0.05)))
            continue
        for s in active:
            s.launch()
        still = []
        for s in active:
            if s.finish(wall_clock()):
                used -= s.bytes
                results.append(s.result())
                if log:
                    r = results[-1]
                    log(f"done {r['rid']} {r['status']} steps={r['steps']} t={r['done']:.3f}")
                s.frontier = None
            else:
                still.append(s)
        active = still
    wall = wall_clock()
    results.sort(key=lambda r: r["rid"])
    return results, summarize(results, wall)


def serve_plain_sequential(requests, cfg: ServeConfig = None, device="cuda", budget_s=None,
                           log=None):
    """The plain FP64 form, one instance at a time in arrival order: the
    op-by-op PyTorch bound (gear_amd.plain), the autograd ALM and the same
    B&B steps, each instance run to its gap or time limit before the next
    starts.  Requests not started within budget_s are reported as not
    converged."""
    from .plain import plain_qcqp_crown
    cfg = cfg or ServeConfig()
    device = torch.device(device)
    queue = sorted(requests, key=lambda r: r.arrival)
    results = []
    t0 = time.perf_counter()
    now = lambda: time.perf_counter() - t0

    def plain_bound(d):
        return lambda L, U, **kw: plain_qcqp_crown(d, L, U, **kw)

    for req in queue:
        if budget_s is not None and now() > budget_s:
            results.append({"rid": req.rid, "n": req.problem.n, "m": req.problem.m,
                            "arrival": req.arrival, "admit": None, "done": math.inf,
                            "status": "not_started", "time_to_gap": math.inf, "P": math.inf,
                            "D": -math.inf, "steps": 0, "boxes": 0, "batch": 0})
            continue
        wait = req.arrival - now()
        if wait > 0:
            time.sleep(wait)
        s = Session(req, cfg, device, bound_fn=plain_bound, alm_backend="torch")
        s.start(now())
        while True:
            s.launch()
            if s.finish(now()):
                break
        results.append(s.result())
        if log:
            r = results[-1]
            log(f"plain done {r['rid']} {r['status']} steps={r['steps']} t={r['done']:.3f}")
    wall = now()
    results.sort(key=lambda r: r["rid"])
    return results, summarize(results, wall)
