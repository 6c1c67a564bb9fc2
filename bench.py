"""Benchmark: subdomain bounds per second of the batched B&B loop on one GPU,
with the expression-program, large-n FP32-sound and serving sections beside
it (each its own JSON object on the same line).

Workload: seeded nonconvex QCQP, n = 50, m = 20, coefficients N(0,1)/sqrt(n),
x in [-1, 1]^50; the loop bounds (CROWN, FP64), branches on the widest
variable, prunes against the ALM-PGD incumbent and the constraint rows, and
appends survivors to a device frontier of at most 2^18 domains.  One step =
one B&B iteration of the reordered loop: the picked domains (the largest
power of two up to 32,768 whose growth fits the frontier over the run) are
bounded, bisected, keyed and filtered into the frontier.  The loop starts
from the root bisected (widest variable first) into one step's worth of
leaves; the root ALM-PGD search runs once before that.

Prints one JSON line (see the keys below).  The CPU baseline times the FP64
oracle on the host cores over a bounded sample of the same children; the
plain FP64 GPU path (op-by-op PyTorch) is timed on the same children too.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
# HBM bytes per bound launch of 8,192 boxes read in place (the loop at the 2^18
# frontier cap), measured with rocprofv3 FETCH_SIZE (x2.00) + WRITE_SIZE (x1.00),
# scale factors from a 256 MiB copy (profiles/qcqp_crown_f64_n50_m20.md); scaled
# to the launch size below.
MEASURED_TRAFFIC_BYTES_8K = 35.6e6
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the expression-program bound on the 256-instance
# workload (rocprofv3 FETCH_SIZE x2.00 + WRITE_SIZE x1.00,
# profiles/dag_crown_uneven256.md).
MEASURED_DAG_TRAFFIC_BYTES = 32.58e9


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=0,
                    help="domains picked and bounded per step (0: the largest power of two "
                         "up to 32,768 whose growth over the run fits the 2^18 frontier)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-alm", action="store_true")
    ap.add_argument("--no-dag", action="store_true", help="skip the expression-program section")
    ap.add_argument("--dag-only", action="store_true",
                    help="run only the expression-program section (256 uneven instances)")
    ap.add_argument("--no-large", action="store_true", help="skip the n=500 FP32-sound section")
    ap.add_argument("--large-only", action="store_true",
                    help="run only the n=500, m=300 FP32-sound section")
    ap.add_argument("--large-boxes", type=int, default=65536)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving section")
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the serving section (Poisson stream of QCQP instances)")
    ap.add_argument("--serve-requests", type=int, default=12)
    ap.add_argument("--serve-rate", type=float, default=0.25, help="arrivals per second")
    ap.add_argument("--serve-plain-budget", type=float, default=60.0,
                    help="seconds the plain sequential form may run")
    ap.add_argument("--no-clip", action="store_true", help="skip the clip_type 'complete' section")
    ap.add_argument("--clip-only", action="store_true",
                    help="only the clip_type 'complete' section (kernels and time to gap)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 5 times (a known-size "
                         "reference for the HBM byte counters) and exit")
    return ap.parse_args()


def copy_probe(mib):
    import torch
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copy probe: {mib} MiB read + {mib} MiB written per launch", flush=True)


def bench_uneven_dag(reps=20, cpu=True):
    """Expression programs: 256 seeded instances (n lognormal in [5, 200],
    m = n // 4, random expression rows of depth 2-5), each with its own
    batch of 64 * 2^k boxes (k uniform in 0..5), bounded in one persistent
    launch replayed from a captured graph with the autotuned wave count."""
    import numpy as np
    import torch

    from gear_amd.dag import DagPlan
    from gear_amd.problems import random_bisection_boxes
    from gear_amd.program import uneven_nlp_instances

    dev = torch.device("cuda:0")
    insts = uneven_nlp_instances()
    rng = np.random.default_rng(0)
    counts = [int(64 * 2 ** rng.integers(0, 6)) for _ in insts]
    boxes = [random_bisection_boxes(P.xl, P.xu, B, seed=k)
             for k, (P, B) in enumerate(zip(insts, counts))]
    t_plan = time.perf_counter()
    plan = DagPlan(insts, device=dev)
    t_plan = time.perf_counter() - t_plan
    xL = [torch.as_tensor(L, device=dev) for L, _ in boxes]
    xU = [torch.as_tensor(U, device=dev) for _, U in boxes]
    lbs = [torch.empty(B, P.R, dtype=torch.float64, device=dev) for B, P in zip(counts, insts)]
    cnt = torch.tensor(counts, dtype=torch.int64, device=dev)
    batches = plan.make_batches(xL, xU, lbs, [cnt[k:k + 1] for k in range(len(insts))])
    total = sum(counts)
    waves = plan.tune_waves(batches, total)
    g = plan.graph(batches, waves)
    stream = torch.cuda.current_stream()
    for _ in range(3):
        g.replay()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        g.replay()
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    state = sum(c * b for c, b in zip(counts, plan.state_bytes))
    io = sum(c * b for c, b in zip(counts, plan.io_bytes))
    flops = sum(c * f for c, f in zip(counts, plan.flops))
    trans = sum(c * t for c, t in zip(counts, plan.transcendentals))
    achieved = flops / (ms * 1e-3) / 1e12
    # algorithmic HBM bytes: each box in (2 n doubles) and its bounds out (S)
    alg_bytes = sum(c * 8 * (2 * P.n + P.R) for c, P in zip(counts, insts))
    out = {
        "value": round(total / (ms * 1e-3), 1),
        "unit": "subdomain bounds/s (every row of the instance)",
        "row_bounds_per_s": round(sum(c * P.R for c, P in zip(counts, insts)) / (ms * 1e-3), 1),
        "ms_per_launch": round(ms, 4),
        "dtype": "f64",
        "config": {"workload": "256 expression-program instances, n lognormal(ln 40, 0.8) "
                               "clipped to [5, 200], m = n // 4, boxes per instance "
                               "64 * 2^U{0..5}, one persistent launch (graph replay)",
                   "instances": len(insts), "boxes": total, "rows": sum(P.R for P in insts),
                   "waves": waves, "plan_s": round(t_plan, 3), "jobs": plan.jobs},
        "roofline": {
            "bound": "valu",
            "kernel": "dag_crown_kernel",
            "achieved": round(achieved, 3),
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
            "traffic": round(MEASURED_DAG_TRAFFIC_BYTES) if MEASURED_DAG_TRAFFIC_BYTES else None,
            "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE, profiles/dag_crown_uneven256.md",
            "algorithmic_flops": flops,
            "transcendentals": trans,
            "algorithmic_bytes": alg_bytes,
            "algorithmic": "FP64 operations by the counting rule, intermediate-bound passes "
                           "included (gh_dag_plan_work); bytes: boxes in and bounds out",
            "state_bytes": state + io,
            "state": "per-box state the backward passes read and write (A, planes, bounds): "
                     "the workspace traffic a cache-resident schedule would avoid",
        },
        "parity": "bounds identical to the FP64 oracle (tests/test_gpu_dag.py)",
    }
    if cpu:
        from oracle import oracle as O
        cores = min(16, os.cpu_count() or 1)
        tc = 0.0
        nb = 0
        for k, P in enumerate(insts):
            L, U = boxes[k]
            t = time.perf_counter()
            O.dag_crown(P, L[:32], U[:32], threads=cores)
            tc += time.perf_counter() - t
            nb += min(32, len(L))
        out["cpu_baseline"] = {"value": round(nb / tc, 1), "unit": "subdomain bounds/s",
                               "cores": cores, "kind": "port",
                               "sample": f"first 32 boxes of each of the 256 instances "
                                         f"({nb} boxes, {tc:.2f} s wall), FP64 oracle"}
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def bench_nlp_config1(reps=5):
    """The n = 8, m = 3 NLP (x^T Q x + sum c_i exp(0.3 x_i) + 4 bilinear
    terms, three quadratic constraints) over 1,024 boxes of seeded random
    bisection of [-2, 2]^8 (default_rng(2)): the FP64 CPU restatement timed
    on all host cores and on one thread (best of 5 after one warm-up), and
    the HIP expression-program bound of the same boxes (identical bounds,
    tests/test_gpu_dag.py)."""
    import torch

    from gear_amd.dag import DagPlan
    from gear_amd.problems import random_bisection_boxes
    from gear_amd.program import nlp_n8_m3_problem
    from oracle import oracle as O

    P = nlp_n8_m3_problem()
    L, U = random_bisection_boxes(P.xl, P.xu, 1024, seed=2)
    out = {"config": {"workload": "nlp n=8 m=3 (exp, bilinear, quadratic constraints), 1024 "
                                  "boxes of seeded random bisection of [-2, 2]^8",
                      "boxes": 1024, "rows": P.R},
           "host": {"cpu_model": cpu_model(), "nproc": os.cpu_count()}}
    cpu = {}
    for threads in (os.cpu_count() or 1, 1):
        O.dag_crown(P, L, U, threads=threads)
        best = float("inf")
        for _ in range(reps):
            t = time.perf_counter()
            O.dag_crown(P, L, U, threads=threads)
            best = min(best, time.perf_counter() - t)
        cpu[f"threads_{threads}"] = {"value": round(1024 / best, 1), "unit": "subdomain bounds/s",
                                     "ms_per_call": round(best * 1e3, 3)}
    out["cpu_fp64_oracle"] = cpu
    dev = torch.device("cuda:0")
    plan = DagPlan([P], device=dev)
    xL, xU = [torch.as_tensor(L, device=dev)], [torch.as_tensor(U, device=dev)]
    plan.bound(xL, xU)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        plan.bound(xL, xU)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / 20
    out["gpu_fp64"] = {"value": round(1024 / (ms * 1e-3), 1), "unit": "subdomain bounds/s",
                       "ms_per_call": round(ms, 4),
                       "note": "one call bounds the 1,024 boxes (plan made once)"}
    return out


def bench_precision_large(boxes=65536, reps=2, cpu=True, sample=16, chunk=512):
    """Dense QCQP n = 500, m = 300 (config-4 generator), FP32-sound bound of
    one batch of 65,536 boxes of seeded bisection (fused FP32 MFMA kernels,
    FP64 relaxation), checked box by box against the FP64 bound of the same
    boxes on the device (the op-by-op FP64 path, itself compared with the
    FP64 oracle on a sample): violations and looseness over the whole batch."""
    import numpy as np
    import torch

    from gear_amd.crown import BoundedQCQP
    from gear_amd.plain import plain_qcqp_crown
    from gear_amd.problems import qcqp_n500_m300_problem, random_bisection_boxes

    dev = torch.device("cuda:0")
    p = qcqp_n500_m300_problem()
    n, S = p.n, p.S
    L, U = random_bisection_boxes(p.xl, p.xu, boxes, seed=7)
    net = BoundedQCQP(p, dev, precision="fp32s")
    xL = torch.as_tensor(L, device=dev)
    xU = torch.as_tensor(U, device=dev)
    lb = net.bound(xL[:1024], xU[:1024])
    lb = torch.empty(boxes, S, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        net.bound(xL, xU, out=lb)
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    from gear_amd.problems import flops_per_box
    fpb = flops_per_box(n, S)
    achieved = fpb * boxes / (ms * 1e-3) / 1e12
    # FP64 bounds of every box on the device, in chunks
    t64 = time.perf_counter()
    viol = 0
    loose_max, loose_sum = 0.0, 0.0
    lb64_all = torch.empty_like(lb)
    for c0 in range(0, boxes, chunk):
        lb64 = plain_qcqp_crown(net.dev, xL[c0:c0 + chunk], xU[c0:c0 + chunk])
        lb64_all[c0:c0 + chunk] = lb64
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t64
    lb32 = lb
    # the FP64 path agrees with the oracle to rounding (below): a violation
    # is an FP32-sound bound above the FP64 one by more than that
    tol = 1e-11 * (1.0 + lb64_all.abs())
    viol = int((lb32 > lb64_all + tol).sum())
    loose = (lb64_all - lb32) / (lb64_all.abs() + 1.0)
    out = {
        "value": round(boxes / (ms * 1e-3), 1),
        "unit": "subdomain bounds/s",
        "ms_per_batch": round(ms, 3),
        "dtype": "f32 (dense products) + f64 (relaxation, error terms)",
        "config": {"workload": "dense QCQP n=500, m=300, coefficients N(0,1)/sqrt(n), one batch "
                               "of 65,536 boxes of seeded random bisection of [-1, 1]^500",
                   "boxes": boxes, "n": n, "m": S - 1},
        "roofline": {"bound": "mfma",
                     "kernel": "large_fwd_kernel + large_bwd_kernel (fused FP32 MFMA)",
                     "achieved": round(achieved, 2), "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                     "flops_per_box": fpb,
                     "note": "whole bound per batch; the kernels also form the magnitude "
                             "product |Q| |x0| (2 S n^2 per box more, not counted)"},
        "near_tie_boxes": int(net.last_flag.sum()),
        "soundness": {"boxes": boxes, "violations_vs_fp64_gpu": viol,
                      "looseness_max": float(loose.max()), "looseness_mean": float(loose.mean()),
                      "fp64_gpu_s": round(t64, 2),
                      "fp64_gpu": "op-by-op FP64 CROWN of the same boxes on the device"},
    }
    if cpu:
        from oracle import oracle as O
        cores = os.cpu_count() or 1
        k = min(sample, boxes)
        t = time.perf_counter()
        lbo = O.qcqp_crown(p.Q, p.c, p.d, L[:k], U[:k], threads=cores)
        t = time.perf_counter() - t
        g64 = lb64_all[:k].cpu().numpy()
        out["soundness"]["fp64_gpu_vs_oracle_max_rel"] = float(
            np.max(np.abs(g64 - lbo) / (1.0 + np.abs(lbo))))
        out["soundness"]["violations_vs_fp64_oracle_sample"] = int(
            np.sum(lb32[:k].cpu().numpy() > lbo))
        out["cpu_baseline"] = {"value": round(k / t, 2), "unit": "subdomain bounds/s",
                               "cores": cores, "kind": "port",
                               "sample": f"{k} of the boxes, FP64 oracle ({t:.2f} s wall)"}
    return out


def bench_serving(requests=12, rate=0.25, plain_budget=60.0, plain=True, time_limit=3.0):
    """Serving: a Poisson stream of the serving workload (config5_stream:
    the config-2 generator with n lognormal(ln 60, 1) in [5, 1000], m = 20),
    admitted by HBM budget and solved to a 1e-4 relative gap with
    continuous batching, FP32-sound bounds and FP64-replayed incumbents;
    beside it the plain FP64 form (op-by-op bound, GEMM-form ALM, the same
    planes, keys and LP certificate) one instance at a time.  The arrival
    rate loads the plain form at about 80 % (its service time is at most
    the per-instance time limit)."""
    import torch

    from gear_amd.serve import ServeConfig, config5_stream, serve, serve_plain_sequential

    dev = torch.device("cuda:0")
    reqs = config5_stream(requests, rate)
    cfg = ServeConfig(time_limit=time_limit)
    # warm the kernels once (not timed): one small and one large instance
    serve(config5_stream(2, 100.0, seed=99), ServeConfig(time_limit=0.5), device=dev,
          realtime=False)
    res, summ = serve(reqs, cfg, device=dev, realtime=True)
    out = {"value": round(summ["subdomain_bounds_per_s"], 1),
           "unit": "subdomain bounds/s",
           "p50_time_to_gap_s": summ["p50_time_to_gap_s"],
           "p99_time_to_gap_s": summ["p99_time_to_gap_s"],
           "converged": summ["converged"], "timeout": summ["timeout"],
           "infeasible": summ["infeasible"], "censored": summ["censored"],
           "wall_s": round(summ["wall_s"], 3),
           "config": {"workload": f"{requests} QCQP instances (config-2 generator, n lognormal("
                                  f"ln 60, 1.0) in [5, 1000], m = 20, seeds 6 + i), Poisson "
                                  f"arrivals at {rate}/s, rel gap 1e-4, time limit "
                                  f"{time_limit} s per instance",
                      "n": [r["n"] for r in res], "precision": cfg.precision,
                      "hbm_budget_bytes": cfg.hbm_budget},
           "time_to_gap_s": [round(r["time_to_gap"], 4) if r["time_to_gap"] < float("inf")
                             else None for r in res],
           "D_at_end": [r["D"] for r in res], "P_at_end": [r["P"] for r in res]}
    if plain:
        pres, psumm = serve_plain_sequential(reqs, cfg, device=dev, budget_s=plain_budget)
        out["plain_fp64_sequential"] = {
            "value": round(psumm["subdomain_bounds_per_s"], 1), "unit": "subdomain bounds/s",
            "p50_time_to_gap_s": psumm["p50_time_to_gap_s"],
            "p99_time_to_gap_s": psumm["p99_time_to_gap_s"],
            "converged": psumm["converged"], "timeout": psumm["timeout"],
            "censored": psumm["censored"],
            "not_started": sum(r["status"] == "not_started" for r in pres),
            "wall_s": round(psumm["wall_s"], 3), "budget_s": plain_budget,
            "D_at_end": [r["D"] for r in pres]}
        same = [abs(a["D"] - b["D"]) <= 3e-4 * max(1.0, abs(a["P"]))
                for a, b in zip(res, pres) if a["status"] == b["status"] == "converged"]
        out["plain_fp64_sequential"]["certificates_agree"] = f"{sum(same)}/{len(same)}"
    return out


def _tight_qcqp():
    """A small QCQP with binding constraints (rhs < 0) and a nonconvex
    objective: the instance where the constrained bound and clipping
    matter (seeded, default_rng(8))."""
    import numpy as np
    from gear_amd.problems import QCQP
    rng = np.random.default_rng(8)
    n, m = 6, 3
    A = rng.standard_normal((n, n))
    Q = np.zeros((m + 1, n, n))
    Q[0] = A @ A.T / n - 0.5 * np.eye(n)
    c = rng.standard_normal((m + 1, n))
    rhs = np.concatenate([[0.0], -np.abs(rng.standard_normal(m)) * 0.3])
    return QCQP(Q=Q, c=c, d=np.zeros(m + 1), rhs=rhs, xl=-np.ones(n), xu=np.ones(n))


def bench_clip(boxes=32768, reps=10):
    """clip_type 'complete': (1) the constrained bound and the clip kernels
    on one B&B step's worth of n = 50, m = 20 domains whose cuts are their
    parents' CROWN planes; (2) time and domains to the 1e-4 gap on a tight
    instance with and without it."""
    import numpy as np
    import torch
    from gear_amd.bab import BaBConfig, QCQPBranchAndBound
    from gear_amd.concretize import clip_domains, constrained_bound
    from gear_amd.crown import BoundedQCQP, input_split_branching, input_split_parallel
    from gear_amd.problems import qcqp_n50_m20_problem, random_bisection_boxes
    dev = torch.device("cuda:0")
    p = qcqp_n50_m20_problem()
    n, S = p.n, p.S
    net = BoundedQCQP(p, dev)
    L, U = random_bisection_boxes(p.xl, p.xu, boxes // 2, seed=5)
    pL, pU = torch.as_tensor(L, device=dev), torch.as_tensor(U, device=dev)
    _, plA, plbias = net.bound(pL, pU, return_A=True)
    # an incumbent at the 30th percentile of the parents' objective bounds
    # makes the objective row a cut too
    thr = torch.as_tensor(np.concatenate([[0.0], p.rhs[1:] + 1e-6]), device=dev)
    lb_p = net.bound(pL, pU)
    thr[0] = torch.quantile(lb_p[:, 0], 0.3)
    idx = input_split_branching(None, None, pL, pU, None, None, "naive")
    cL, cU, _ = input_split_parallel(pL, pU, i_idx=idx)
    K = cL.shape[0] // 2
    par = torch.arange(2 * K, device=dev) % K
    cutA = plA[par].contiguous()
    act = torch.isfinite(thr)
    cutb = torch.where(act[None, :], plbias[par] - thr, -1.0).contiguous()
    lb, lA, lbias = net.bound(cL, cU, return_A=True)
    stream = torch.cuda.current_stream()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps * 1e-3

    lb_c = lb.clone()
    _, flags, kept = constrained_bound(lA, lbias, cutA, cutb, cL, cU, lb_c)
    t_cb = timed(lambda: constrained_bound(lA, lbias, cutA, cutb, cL, cU, lb.clone()))
    t_box = timed(lambda: net.bound(cL, cU, return_A=True))
    oL, oU = cL.clone(), cU.clone()
    _, _, st = clip_domains(plA, plbias, thr, cL, cU, out=(oL, oU))
    t_clip = timed(lambda: clip_domains(plA, plbias, thr, cL, cU, out=(oL, oU)))
    shrunk = int((((oL > cL).any(1)) | ((oU < cU).any(1))).sum())
    raised = int((lb_c[:, 0] > lb[:, 0]).sum())
    B = cL.shape[0]
    cut_bytes = S * (n + 1) * 8
    kernels = {
        "domains": B, "n": n, "cuts_per_domain": S,
        "constrained_bound": {"ms": round(t_cb * 1e3, 3), "domains_per_s": round(B / t_cb, 1),
                              "cuts_kept_mean": round(float(kept.float().mean()), 2),
                              "objective_bounds_raised": raised,
                              "infeasible_domains": int((flags & 1).sum()),
                              "bytes_per_domain": S * n * 8 + S * 8 + cut_bytes + 2 * n * 8 + S * 8,
                              "bound": "VALU (sort network and FP64 compare per cut and row)"},
        "crown_planes_same_domains": {"ms": round(t_box * 1e3, 3),
                                      "domains_per_s": round(B / t_box, 1)},
        "clip": {"ms": round(t_clip * 1e3, 3), "domains_per_s": round(B / t_clip, 1),
                 "shrunk": shrunk, "emptied": int((st >= 2).sum()),
                 "bytes_per_domain": 2 * 2 * n * 8},
    }
    # time and domains to the 1e-4 gap on the tight instance
    tight = _tight_qcqp()
    solve = {}
    for clip in ("none", "complete"):
        bb = QCQPBranchAndBound(tight, BaBConfig(batch=256, cap=1 << 16, clip=clip,
                                                 max_iters=400, time_limit=15.0), device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bb.solve()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        from gear_amd.bab import is_primal_dual_converged
        solve[clip] = {"P": bb.P, "D": bb.D,
                       "converged": bool(is_primal_dual_converged(bb.P, bb.D, 1e-5, 1e-4)),
                       "domains_bounded": bb.stats.boxes_bounded,
                       "iterations": bb.stats.iterations, "stop": bb.stats.stop,
                       "wall_s": round(wall, 3)}
    return {"kernels": kernels, "time_to_gap_tight_n6_m3": solve,
            "note": "clip_type 'complete' (the GEAR presets' clip_input_domain): domains carry "
                    "their parent's planes as cuts; the time-to-gap runs use the same loop, "
                    "batch 256, iteration cap 400, time limit 15 s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.dag_only:
        print(json.dumps(bench_uneven_dag(cpu=not args.no_cpu)), flush=True)
        return
    if args.large_only:
        print(json.dumps(bench_precision_large(args.large_boxes, cpu=not args.no_cpu)), flush=True)
        return
    if args.clip_only:
        print(json.dumps(bench_clip()), flush=True)
        return
    if args.serve_only:
        print(json.dumps(bench_serving(args.serve_requests, args.serve_rate,
                                       args.serve_plain_budget, plain=not args.no_plain)),
              flush=True)
        return
    import numpy as np
    import torch

    from gear_amd.bab import BaBConfig, QCQPBranchAndBound
    from gear_amd.crown import input_split_branching, input_split_parallel
    from gear_amd.problems import qcqp_n50_m20_problem, flops_per_box, bytes_per_box

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    p = qcqp_n50_m20_problem()
    n, S = p.n, p.S
    cap = 1 << 18
    # a reordered step picks `picks` domains, bounds them and appends at most
    # 2 picks children: the frontier grows by at most `picks` per step
    picks = args.batch
    if picks <= 0:
        picks = 32768
        while picks > 1024 and picks * (args.warmup + args.steps + 1) > cap:
            picks //= 2
    need = picks * (args.warmup + args.steps + 1)
    if need > cap:
        raise SystemExit(f"picks * (warmup + steps + 1) = {need} exceeds the frontier cap {cap}")
    args.batch = picks // 2      # BaBConfig.batch: half the domains picked per step
    # physical rows: cap and as many again, so the run's appends need no compaction
    bb = QCQPBranchAndBound(p, BaBConfig(batch=args.batch, cap=cap, frontier_slack=cap),
                            device=dev)
    t_alm = time.time()
    # presplit start: `picks` leaves so every step bounds a full batch
    bb.presplit(2 * args.batch, alm=not args.no_alm)
    torch.cuda.synchronize()
    t_alm = time.time() - t_alm
    stream = torch.cuda.current_stream()

    # bound-kernel timing: events on the stream the kernel is launched on
    events = []
    orig_bound = bb.net.bound

    def timed_bound(x_L, x_U, **kw):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        out = orig_bound(x_L, x_U, **kw)
        e1.record(stream)
        rows = kw.get("rows")
        events.append((e0, e1, x_L.shape[0] if rows is None else rows.shape[0]))
        return out

    bb.net.bound = timed_bound
    for _ in range(args.warmup):
        bb.step()
    events.clear()
    sample_parents = None
    if len(bb.frontier) >= args.batch:
        s0 = len(bb.frontier) - args.batch
        sample_parents = bb.frontier.rows(s0, s0 + args.batch)[:2]
    import gc
    gc.collect()
    gc.disable()    # no collector pause inside the timed steps
    torch.cuda.synchronize()
    boxes = 0
    step_ms = []
    t0 = time.perf_counter()
    # steps are issued without reading each step's survivor count (the
    # frontier's dead-slot form); per-step host times are issue times
    for _ in range(args.steps):
        ts = time.perf_counter()
        boxes += bb.step(sync=False)
        step_ms.append(round((time.perf_counter() - ts) * 1e3, 3))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    gc.enable()
    bb.frontier.compact_order()
    bb.update_dual()
    wall = t1 - t0
    kt = [(a.elapsed_time(b) * 1e-3, nb) for a, b, nb in events]
    k_time = sum(t for t, _ in kt)
    k_boxes = sum(nb for _, nb in kt)
    avg_launch = k_time / max(len(kt), 1)
    avg_boxes = k_boxes / max(len(kt), 1)
    fpb = flops_per_box(n, S)
    achieved_tf = fpb * avg_boxes / avg_launch / 1e12
    value = boxes / wall

    # the same loop with FP32-sound bounds (fresh tree, same start)
    bb32 = QCQPBranchAndBound(p, BaBConfig(batch=args.batch, cap=cap, frontier_slack=cap),
                              device=dev, precision="fp32s")
    if bb.x_best is not None:
        bb32.set_incumbent(bb.P, bb.x_best)
    bb32.presplit(2 * args.batch, alm=False)
    for _ in range(args.warmup):
        bb32.step()
    torch.cuda.synchronize()
    gc.collect()
    gc.disable()
    boxes32 = 0
    t32 = time.perf_counter()
    for _ in range(args.steps):
        boxes32 += bb32.step(sync=False)
    torch.cuda.synchronize()
    t32 = time.perf_counter() - t32
    gc.enable()
    bb32.frontier.compact_order()
    bb32.update_dual()

    result = {
        "metric": "subdomain bounds/sec + p50/p99 time-to-1e-4 gap, served FP32-sound vs plain FP64",
        "value": round(value, 1),
        "unit": "subdomain bounds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded QCQP generator, default_rng(3))",
        "config": {"workload": "qcqp_n50_m20 batched B&B, reordered step: sort -> bound the "
                               "picked domains (CROWN + planes) -> widest branch -> bisect -> "
                               "child keys -> prune -> append, frontier cap 2^18",
                   "n": n, "m": S - 1, "domains_bounded_per_step": 2 * args.batch,
                   "children_per_step": 4 * args.batch},
        "roofline": {
            "bound": "mfma",
            "kernel": "qcqp_crown_f64_kernel",
            "achieved": round(achieved_tf, 3),
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved_tf / FP64_PEAK_TFLOPS, 4),
            "traffic": round(MEASURED_TRAFFIC_BYTES_8K * avg_boxes / 8192),
            "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE, profiles/qcqp_crown_f64_n50_m20.md",
            "flops_per_box": fpb,
            "bytes_per_box": bytes_per_box(n, S),
            "avg_launch_ms": round(avg_launch * 1e3, 4),
            "boxes_per_launch": avg_boxes,
        },
        "step_ms": step_ms,
        "bab": {"frontier": len(bb.frontier), "P": bb.P, "D": bb.D,
                "root_alm_s": round(t_alm, 3)},
        "bab_fp32_sound": {"value": round(boxes32 / t32, 1), "unit": "subdomain bounds/s",
                           "ms_per_step": round(t32 / args.steps * 1e3, 4),
                           "frontier": len(bb32.frontier), "D": bb32.D,
                           "note": "the same loop and start with FP32-sound bounds"},
    }

    if sample_parents is not None and not args.no_plain:
        from gear_amd.plain import plain_qcqp_crown
        pL, pU = sample_parents
        idx = input_split_branching(None, None, pL, pU, None, None, "naive")
        cL, cU, _ = input_split_parallel(pL, pU, i_idx=idx)
        plain_qcqp_crown(bb.net.dev, cL, cU)
        torch.cuda.synchronize()
        tp = time.perf_counter()
        reps = 3
        for _ in range(reps):
            plain_qcqp_crown(bb.net.dev, cL, cU)
        torch.cuda.synchronize()
        tp = (time.perf_counter() - tp) / reps
        fused = orig_bound(cL, cU)
        torch.cuda.synchronize()
        tf = time.perf_counter()
        for _ in range(reps):
            orig_bound(cL, cU, out=fused)
        torch.cuda.synchronize()
        tf = (time.perf_counter() - tf) / reps
        # FP32-sound form on the same children
        from gear_amd.crown import BoundedQCQP
        net32 = BoundedQCQP(p, dev, precision="fp32s")
        lb32 = net32.bound(cL, cU)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        for _ in range(reps):
            net32.bound(cL, cU, out=lb32)
        torch.cuda.synchronize()
        t3 = (time.perf_counter() - t3) / reps
        # LP dual certificate (FP64, 500 PDHG iterations) on the same children
        from gear_amd.lp import qcqp_lp_bound
        _, lA_c, lbias_c = orig_bound(cL, cU, return_A=True)
        thr_c = bb.thr
        qcqp_lp_bound(lA_c, lbias_c, thr_c, cL, cU, iters=500)
        torch.cuda.synchronize()
        tl = time.perf_counter()
        lp_c = qcqp_lp_bound(lA_c, lbias_c, thr_c, cL, cU, iters=500)
        torch.cuda.synchronize()
        tl = time.perf_counter() - tl
        raised = int((lp_c > fused[:, 0] + 1e-9 * (1 + fused[:, 0].abs())).sum())
        result["lp_certificate"] = {"value": round(cL.shape[0] / tl, 1),
                                    "unit": "subdomain LP certificates/s",
                                    "pdhg_iters": 500, "ms": round(tl * 1e3, 3),
                                    "boxes_raised_above_crown": raised,
                                    "flops_per_box": 500 * 4 * (S - 1) * n}
        del lA_c, lbias_c
        viol = int((lb32 > fused).sum())
        loose = ((fused - lb32) / (fused.abs() + 1.0))
        result["plain_fp64_gpu"] = {"value": round(cL.shape[0] / tp, 1),
                                    "unit": "subdomain bounds/s",
                                    "fused_same_boxes": round(cL.shape[0] / tf, 1),
                                    "speedup": round(tp / tf, 2)}
        result["fp32_sound"] = {"value": round(cL.shape[0] / t3, 1),
                                "unit": "subdomain bounds/s",
                                "vs_plain_fp64": round(tp / t3, 2),
                                "vs_fused_fp64": round(tf / t3, 2),
                                "violations_vs_fp64": viol,
                                "near_tie_boxes_settled_fp64": int(net32.last_flag.sum()),
                                "looseness_max": float(loose.max()),
                                "looseness_mean": float(loose.mean())}

    if sample_parents is not None and not args.no_cpu:
        from oracle import oracle as O
        pL, pU = sample_parents
        idx = input_split_branching(None, None, pL, pU, None, None, "naive")
        cL, cU, _ = input_split_parallel(pL, pU, i_idx=idx)
        L = cL.cpu().numpy()
        U = cU.cpu().numpy()
        cores = min(16, os.cpu_count() or 1)
        O.qcqp_crown(p.Q, p.c, p.d, L[:64], U[:64], threads=cores)
        tc = time.perf_counter()
        O.qcqp_crown(p.Q, p.c, p.d, L, U, threads=cores)
        tc = time.perf_counter() - tc
        result["cpu_baseline"] = {"value": round(L.shape[0] / tc, 1),
                                  "unit": "subdomain bounds/s", "cores": cores, "kind": "port",
                                  "sample": f"{L.shape[0]} children of one step, FP64 oracle "
                                            f"({tc:.2f} s wall)"}
    if not args.no_cpu:
        result["nlp_cpu_config1"] = bench_nlp_config1()
    if not args.no_dag:
        result["uneven_dag"] = bench_uneven_dag(cpu=not args.no_cpu)
    if not args.no_large:
        result["precision_large"] = bench_precision_large(args.large_boxes, cpu=not args.no_cpu)
    if not args.no_clip:
        result["clip_complete"] = bench_clip()
    if not args.no_serve:
        result["serving"] = bench_serving(args.serve_requests, args.serve_rate,
                                          args.serve_plain_budget, plain=not args.no_plain)
        sv = result["serving"]
        head = {"served_fp32s_p50_time_to_gap_s": sv["p50_time_to_gap_s"],
                "served_fp32s_p99_time_to_gap_s": sv["p99_time_to_gap_s"],
                "served_fp32s_subdomain_bounds_per_s": sv["value"]}
        if "plain_fp64_sequential" in sv:
            pl = sv["plain_fp64_sequential"]
            head.update({"plain_fp64_p50_time_to_gap_s": pl["p50_time_to_gap_s"],
                         "plain_fp64_p99_time_to_gap_s": pl["p99_time_to_gap_s"],
                         "plain_fp64_subdomain_bounds_per_s": pl["value"],
                         "served_vs_plain_bounds_per_s": round(sv["value"] / pl["value"], 2)
                         if pl["value"] else None})
        result["headline"] = head
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
