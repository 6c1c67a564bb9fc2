"""Serving loop: certificates are sound, each instance's result is the same
alone and in the mix, HBM-budget admission, and the plain sequential form
agrees."""
import math

import numpy as np
import pytest
import torch

from gear_amd.alm import replay_float64
from gear_amd.serve import (ServeConfig, poisson_stream, serve, serve_plain_sequential)

pytestmark = pytest.mark.gpu


def _check_certificate(req, r, cfg):
    p = req.problem
    assert r["D"] <= r["P"] + cfg.abs_tol + 1e-9 * abs(r["P"])
    if r["status"] == "converged":
        gap = r["P"] - r["D"]
        assert gap <= max(cfg.abs_tol, cfg.rel_tol * max(1.0, abs(r["P"]), abs(r["D"]))) + 1e-12
    # D is below every sampled feasible objective value
    rng = np.random.default_rng(req.rid)
    X = p.xl + rng.random((20000, p.n)) * (p.xu - p.xl)
    F = p.values(X)
    feas = np.all(F[:, 1:] <= p.rhs[1:] + 1e-9, axis=1)
    if feas.any():
        assert r["D"] <= F[feas, 0].min() + 1e-9 * (1 + abs(r["D"]))


def test_served_stream_sound_certificates(cuda):
    reqs = poisson_stream(8, rate=50.0, seed=1, n_median=6, n_sigma=0.3)
    cfg = ServeConfig(time_limit=20.0)
    print("case stream", [(q.problem.n, q.problem.m) for q in reqs], flush=True)
    res, summ = serve(reqs, cfg, device=cuda, realtime=False)
    print(summ, flush=True)
    assert summ["requests"] == 8
    for q, r in zip(reqs, res):
        print(r, flush=True)
        assert r["status"] in ("converged", "infeasible", "timeout")
        _check_certificate(q, r, cfg)
    assert summ["converged"] >= 6


def test_same_result_alone_and_in_the_mix(cuda):
    reqs = poisson_stream(4, rate=100.0, seed=3, n_median=7, n_sigma=0.3)
    cfg = ServeConfig(time_limit=20.0)
    mixed, _ = serve(reqs, cfg, device=cuda, realtime=False)
    for q, m in zip(reqs, mixed):
        alone, _ = serve([q], cfg, device=cuda, realtime=False)
        a = alone[0]
        print(q.rid, "alone", a["P"], a["D"], a["steps"], "mixed", m["P"], m["D"], m["steps"],
              flush=True)
        if a["status"] == "timeout" or m["status"] == "timeout":
            continue  # a time limit cuts the two runs at different steps
        for key in ("status", "P", "D", "steps", "boxes"):
            assert a[key] == m[key], (q.rid, key, a[key], m[key])


def test_budget_admission_serializes(cuda):
    reqs = poisson_stream(3, rate=100.0, seed=4, n_median=6, n_sigma=0.2)
    cfg = ServeConfig(time_limit=20.0, hbm_budget=1)     # room for one session at a time
    lines = []
    res, summ = serve(reqs, cfg, device=cuda, realtime=False, log=lines.append)
    active, peak = 0, 0
    for ln in lines:
        active += 1 if ln.startswith("admit") else (-1 if ln.startswith("done") else 0)
        peak = max(peak, active)
    assert peak == 1
    assert summ["requests"] == 3
    for q, r in zip(reqs, res):
        _check_certificate(q, r, cfg)


def test_plain_sequential_agrees(cuda):
    reqs = poisson_stream(2, rate=100.0, seed=5, n_median=6, n_sigma=0.2)
    cfg = ServeConfig(time_limit=30.0)
    fast, _ = serve(reqs, cfg, device=cuda, realtime=False)
    plain, _ = serve_plain_sequential(reqs, cfg, device=cuda)
    for q, a, b in zip(reqs, fast, plain):
        print(q.rid, "served", a["status"], a["P"], a["D"], "plain", b["status"], b["P"], b["D"],
              flush=True)
        _check_certificate(q, b, cfg)
        if a["status"] == b["status"] == "converged":
            scale = max(1.0, abs(a["P"]))
            assert abs(a["D"] - b["D"]) <= 3e-4 * scale


def test_served_bound_is_never_above_a_feasible_point(cuda):
    """A request whose presplit keeps few of its boxes (request 7 of the
    seed-11 stream, n = 5): the served lower bound D must stay below every
    feasible objective value (sampled), whether or not the gap closes.  The
    first step of a session is launched in the tick that starts it, so this
    also checks that the presplit's survivors are accounted before the
    global bound is read (reading them late made the frontier look empty
    and D equal to the incumbent)."""
    from gear_amd.serve import ServeConfig, poisson_stream, serve
    r = poisson_stream(24, 8.0, seed=11, n_median=8.0, n_sigma=0.6)[7]
    p = r.problem
    rng = np.random.default_rng(0)
    X = rng.uniform(-1.0, 1.0, size=(200000, p.n))
    F = p.values(X)
    feas = (F[:, 1:] <= p.rhs[1:]).all(1)
    assert feas.any()
    fmin = float(F[feas, 0].min())
    req = type(r)(r.rid, 0.0, p)
    res, _ = serve([req], ServeConfig(time_limit=1.5, cap=1 << 18), device=cuda, realtime=False)
    x = res[0]
    print(f"status {x['status']} P {x['P']:.6f} D {x['D']:.6f} sampled feasible min {fmin:.6f}",
          flush=True)
    assert x["D"] <= fmin + 1e-9
    assert x["steps"] > 0


def test_serving_workload_sizes_sound(cuda):
    """The serving workload's sizes (n lognormal(ln 60, 1) in [5, 1000],
    m = 20): sessions above n = 64 run the large-n FP32-sound bound beside
    the small ones in the same ticks; every certificate is sound."""
    from gear_amd.serve import config5_stream
    reqs = config5_stream(5, rate=100.0)
    ns = [q.problem.n for q in reqs]
    print("case serving sizes", ns, flush=True)
    assert max(ns) > 64 and min(ns) <= 64
    cfg = ServeConfig(time_limit=2.0)
    res, summ = serve(reqs, cfg, device=cuda, realtime=False)
    print(summ, flush=True)
    for q, r in zip(reqs, res):
        print(q.problem.n, r["status"], r["P"], r["D"], r["steps"], r["boxes"], flush=True)
        assert r["steps"] > 0 and r["boxes"] > 0
        _check_certificate(q, r, cfg)


def test_plain_form_runs_the_same_algorithm(cuda):
    """The plain FP64 sequential form takes planes (child keys and the LP
    certificate) where the served form does: on an n <= 64 instance both
    run the reordered loop with planes and their certificates agree."""
    reqs = poisson_stream(1, rate=100.0, seed=7, n_median=6, n_sigma=0.1)
    cfg = ServeConfig(time_limit=20.0)
    fast, _ = serve(reqs, cfg, device=cuda, realtime=False)
    plain, _ = serve_plain_sequential(reqs, cfg, device=cuda)
    a, b = fast[0], plain[0]
    print("served", a, "plain", b, flush=True)
    _check_certificate(reqs[0], b, cfg)
    if a["status"] == b["status"] == "converged":
        assert abs(a["D"] - b["D"]) <= 3e-4 * max(1.0, abs(a["P"]))
