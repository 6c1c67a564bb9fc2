"""Expression-program HIP path vs the FP64 oracle, through the C ABI.

Bounds are identical: the kernel performs the oracle's operations in the
oracle's order, and both use the elementary functions of
include/gearhip_math.h.  Rows with no relaxation fall back to the same
interval, and -inf rows match.  (The bar of the bound is 1e-9 relative; the
test asks for equality and reports the relative error it would have had.)
"""
import numpy as np
import pytest
import torch

from gear_amd.problems import random_bisection_boxes
from gear_amd.program import Program, nlp_n8_m3_problem, uneven_nlp_instances, random_nlp
from oracle import oracle as O

pytestmark = pytest.mark.gpu
RTOL, ATOL = 1e-9, 1e-12


def _dev(a, cuda):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=cuda)


def _check(g, o, what):
    g = np.asarray(g)
    assert g.shape == o.shape, what
    ninf_g, ninf_o = np.isneginf(g), np.isneginf(o)
    assert np.array_equal(ninf_g, ninf_o), f"{what}: -inf rows differ"
    f = ~ninf_o
    assert np.all(np.isfinite(g[f])), f"{what}: non-finite bound"
    err = np.abs(g[f] - o[f])
    tol = RTOL * np.abs(o[f]) + ATOL
    worst = np.max(err / tol) if err.size else 0.0
    exact = np.mean(g[f] == o[f]) if err.size else 1.0
    print(f"{what}: {err.size} bounds, max err/tol {worst:.3g}, equal {exact:.4f}", flush=True)
    assert np.all(err <= tol), what
    assert np.array_equal(g, o), f"{what}: not identical"


def _bound(plans_progs, boxes, cuda, waves=None):
    from gear_amd.dag import DagPlan
    plan = DagPlan(plans_progs, device=cuda)
    lbs = plan.bound([_dev(L, cuda) for L, _ in boxes], [_dev(U, cuda) for _, U in boxes],
                     waves=waves)
    torch.cuda.synchronize()
    return [lb.cpu().numpy() for lb in lbs]


def test_nlp_n8_m3_vs_oracle(cuda):
    print("case nlp n=8 m=3 B=1024", flush=True)
    P = nlp_n8_m3_problem()
    L, U = random_bisection_boxes(P.xl, P.xu, 1024, seed=11)
    lb = _bound([P], [(L, U)], cuda)[0]
    _check(lb, O.dag_crown(P, L, U, threads=8), "nlp_n8_m3")


def test_random_programs_uneven_batches(cuda):
    progs, boxes = [], []
    counts = [0, 1, 63, 64, 65, 200, 7, 130]
    for s, B in enumerate(counts):
        P = random_nlp(10 + s, 3, seed=s)
        progs.append(P)
        boxes.append(random_bisection_boxes(P.xl, P.xu, max(B, 1), seed=s) if B else
                     (np.zeros((0, P.n)), np.zeros((0, P.n))))
    print("case random programs counts", counts, flush=True)
    lbs = _bound(progs, boxes, cuda)
    for s, (P, (L, U)) in enumerate(zip(progs, boxes)):
        assert lbs[s].shape == (len(L), P.R)
        if len(L):
            _check(lbs[s], O.dag_crown(P, L, U, threads=8), f"random program {s}")


def _edge_program():
    P = Program(3)
    P.output(P.mul(P.x(0), P.x(1)))                      # plain bilinear
    P.output(P.x(2))                                     # a bare input row
    P.output(P.const(2.5))                               # a constant row
    P.output(P.recip(P.x(0)))                            # pole inside some boxes
    P.output(P.add(P.mul(0, 0), P.sub(1, 1)))            # same operand twice
    P.output(P.neg(P.pow(P.addc(0, 0.3), 3)))            # odd power
    P.output(P.pow(P.mulc(1, 0.7), 4))                   # even powers
    P.output(P.pow(2, 6))
    P.output(P.sin(P.mulc(2, 5.0)))                      # sin across periods
    P.output(P.log(P.addc(P.sqr(0), 1e-3)))              # log near 0
    P.output(P.mul(P.exp(0), P.sin(P.mul(1, 2))))        # nested relaxed inputs
    P.output(P.div(P.x(1), P.addc(P.sqr(2), 0.5)))       # division
    return P


def test_edge_program(cuda):
    print("case edge program", flush=True)
    P = _edge_program()
    xl, xu = np.array([-2.0, -1.5, -3.0]), np.array([2.0, 1.0, 3.0])
    L, U = random_bisection_boxes(xl, xu, 500, seed=3)
    # a few degenerate boxes
    L[:4] = U[:4]
    lb = _bound([P], [(L, U)], cuda)[0]
    _check(lb, O.dag_crown(P, L, U, threads=8), "edge program")


def _sshape_program():
    """sin / cos over wide and narrow intervals, odd and non-integer powers,
    the operators whose relaxations read tangent tables."""
    P = Program(3)
    P.output(P.sin(P.mulc(0, 7.0)))                      # many periods
    P.output(P.neg(P.sin(P.addc(1, 0.4))))               # upper planes
    P.output(P.cos(P.mulc(2, 2.0)))
    P.output(P.neg(P.cos(P.add(0, 1))))
    P.output(P.pow(P.addc(0, -0.2), 5))                  # odd power 5
    P.output(P.neg(P.pow(P.mulc(1, 3.0), 3)))
    P.output(P.pow(P.addc(P.sqr(2), 0.1), 0.5))          # non-integer powers
    P.output(P.neg(P.pow(P.addc(P.sqr(0), 0.01), 2.5)))
    P.output(P.pow(P.addc(2, 3.5), 1.7))                 # domain cut at x >= 0
    P.output(P.mul(P.sin(0), P.cos(P.mul(1, 2))))        # products of trig terms
    P.output(P.add(P.pow(P.sin(2), 3), P.cos(P.pow(0, 3))))
    return P


@pytest.mark.parametrize("scale", [0.02, 1.0, 6.0])
def test_sshape_program(cuda, scale):
    """Tables and every branch of the sin / cos / odd / non-integer power
    relaxations: the kernel equals the oracle value for value, on boxes from
    1e-2 wide (one convexity piece) to several periods."""
    print(f"case sshape program scale {scale}", flush=True)
    P = _sshape_program()
    rng = np.random.default_rng(int(scale * 100))
    c = rng.uniform(-4, 4, (700, 3))
    w = rng.uniform(0.1, 1.0, (700, 3)) * scale
    L, U = c - w / 2, c + w / 2
    L[:3] = U[:3]                                        # point boxes
    lb = _bound([P], [(L, U)], cuda)[0]
    _check(lb, O.dag_crown(P, L, U, threads=8), f"sshape program x{scale}")


def test_uneven_instances_subset(cuda):
    insts = uneven_nlp_instances()[:24]
    rng = np.random.default_rng(0)
    boxes = []
    for k, P in enumerate(insts):
        B = int(rng.integers(1, 150))
        boxes.append(random_bisection_boxes(P.xl, P.xu, B, seed=k))
    print("case uneven subset", [P.n for P in insts], flush=True)
    lbs = _bound(insts, boxes, cuda)
    for k, (P, (L, U)) in enumerate(zip(insts, boxes)):
        _check(lbs[k], O.dag_crown(P, L, U, threads=8), f"instance {k} n={P.n}")


def test_wave_count_does_not_change_results(cuda):
    P = random_nlp(14, 4, seed=21)
    L, U = random_bisection_boxes(P.xl, P.xu, 300, seed=2)
    a = _bound([P], [(L, U)], cuda, waves=1)[0]
    b = _bound([P], [(L, U)], cuda, waves=37)[0]
    c = _bound([P], [(L, U)], cuda)[0]
    assert np.array_equal(a.view(np.int64), b.view(np.int64))
    assert np.array_equal(a.view(np.int64), c.view(np.int64))


@pytest.mark.parametrize("lds_slots", [1, 8, 64])
def test_lds_slots_do_not_change_results(cuda, lds_slots):
    """The A area kept in LDS (all of it, part of it, one slot) gives the
    same bounds as the global workspace, bit for bit, and the oracle's."""
    from gear_amd.dag import DagPlan
    print(f"case lds slots {lds_slots}", flush=True)
    progs = [random_nlp(14, 4, seed=21), nlp_n8_m3_problem(), _sshape_program()]
    boxes = [random_bisection_boxes(P.xl, P.xu, 200, seed=k) if k < 2 else
             (np.random.default_rng(3).uniform(-3, 1, (200, 3)),
              np.random.default_rng(3).uniform(-3, 1, (200, 3)) + 1.5)
             for k, P in enumerate(progs)]
    plan = DagPlan(progs, device=cuda)
    assert plan.max_aslots > 0
    xL = [_dev(L, cuda) for L, _ in boxes]
    xU = [_dev(U, cuda) for _, U in boxes]
    outs = []
    for ls in (0, lds_slots):
        lbs = [torch.empty(len(L), P.R, dtype=torch.float64, device=cuda)
               for (L, _), P in zip(boxes, progs)]
        cnt = torch.tensor([len(L) for L, _ in boxes], dtype=torch.int64, device=cuda)
        batches = plan.make_batches(xL, xU, lbs, [cnt[k:k + 1] for k in range(len(progs))])
        plan.launch(batches, 96, lds_slots=ls)
        torch.cuda.synchronize()
        outs.append([lb.cpu().numpy() for lb in lbs])
    for k, P in enumerate(progs):
        assert np.array_equal(outs[0][k].view(np.int64), outs[1][k].view(np.int64)), k
        _check(outs[1][k], O.dag_crown(P, *boxes[k], threads=8), f"lds program {k}")


def test_device_counts_follow_batch_size(cuda):
    """Counts are read on the device: shrinking the count leaves the rows
    past it untouched."""
    from gear_amd.dag import DagPlan
    P = random_nlp(9, 2, seed=5)
    L, U = random_bisection_boxes(P.xl, P.xu, 130, seed=1)
    plan = DagPlan([P], device=cuda)
    xL, xU = _dev(L, cuda), _dev(U, cuda)
    lb = torch.full((130, P.R), 7.0, dtype=torch.float64, device=cuda)
    cnt = torch.tensor([70], dtype=torch.int64, device=cuda)
    batches = plan.make_batches([xL], [xU], [lb], [cnt])
    plan.launch(batches, 16)
    torch.cuda.synchronize()
    got = lb.cpu().numpy()
    _check(got[:70], O.dag_crown(P, L[:70], U[:70]), "first 70")
    assert np.all(got[70:] == 7.0)


def test_graph_replay_matches_eager(cuda):
    """The captured bound replays correctly several times, with eager work
    and a changed batch count between replays."""
    from gear_amd.dag import DagPlan
    progs = [random_nlp(8 + k, 2, seed=30 + k) for k in range(3)]
    boxes = [random_bisection_boxes(P.xl, P.xu, 150, seed=k) for k, P in enumerate(progs)]
    plan = DagPlan(progs, device=cuda)
    xL = [_dev(L, cuda) for L, _ in boxes]
    xU = [_dev(U, cuda) for _, U in boxes]
    lb = [torch.zeros(150, P.R, dtype=torch.float64, device=cuda) for P in progs]
    cnt = torch.tensor([150, 150, 150], dtype=torch.int64, device=cuda)
    batches = plan.make_batches(xL, xU, lb, [cnt[k:k + 1] for k in range(3)])
    g = plan.graph(batches, 64)
    for rep, c in enumerate([150, 97, 150, 1]):
        for t in lb:
            t.fill_(-7.0)
        cnt.fill_(c)
        g.replay()
        torch.cuda.synchronize()
        _ = (xL[0] * 2).sum().item()        # eager work between replays
        for k, P in enumerate(progs):
            got = lb[k].cpu().numpy()
            _check(got[:c], O.dag_crown(P, boxes[k][0][:c], boxes[k][1][:c]), f"replay {rep} prog {k}")
            assert np.all(got[c:] == -7.0)


def test_autotuned_waves_give_identical_bounds(cuda):
    from gear_amd.autotune import Autotuner
    from gear_amd.dag import DagPlan
    P = random_nlp(12, 3, seed=40)
    L, U = random_bisection_boxes(P.xl, P.xu, 500, seed=4)
    plan = DagPlan([P], device=cuda)
    xL, xU = _dev(L, cuda), _dev(U, cuda)
    lb = torch.empty(500, P.R, dtype=torch.float64, device=cuda)
    cnt = torch.tensor([500], dtype=torch.int64, device=cuda)
    batches = plan.make_batches([xL], [xU], [lb], [cnt])
    tuner = Autotuner()
    w = plan.tune_waves(batches, 500, tuner=tuner)
    cus = torch.cuda.get_device_properties(cuda).multi_processor_count
    assert w[0] in [k * cus for k in (2, 4, 8, 12)]
    assert w[1] in (0, plan.lds_slots_for(w[0]))
    assert plan.tune_waves(batches, 500, tuner=tuner) == w      # cached
    plan.launch(batches, w)
    torch.cuda.synchronize()
    _check(lb.cpu().numpy(), O.dag_crown(P, L, U), "tuned")


@pytest.mark.parametrize("which", ["nlp", "edge", "sshape", "random"])
def test_planes_match_oracle(cuda, which):
    """With return_A the kernel writes each row's lower plane (lA over the
    program's inputs, lbias before concretization) and the undefined-row
    flags, equal to the oracle's bit for bit; the bounds are unchanged."""
    from gear_amd.dag import DagPlan
    print(f"case planes {which}", flush=True)
    if which == "nlp":
        P = nlp_n8_m3_problem()
        L, U = random_bisection_boxes(P.xl, P.xu, 300, seed=4)
    elif which == "edge":
        P = _edge_program()
        L, U = random_bisection_boxes(np.array([-2.0, -1.5, -3.0]), np.array([2.0, 1.0, 3.0]),
                                      300, seed=5)
    elif which == "sshape":
        P = _sshape_program()
        rng = np.random.default_rng(7)
        c = rng.uniform(-4, 4, (300, 3))
        w = rng.uniform(0.1, 2.0, (300, 3))
        L, U = c - w / 2, c + w / 2
    else:
        P = random_nlp(14, 4, seed=9)
        L, U = random_bisection_boxes(P.xl, P.xu, 200, seed=9)
    plan = DagPlan([P], device=cuda)
    lbs, lAs, lbiases, unds = plan.bound([_dev(L, cuda)], [_dev(U, cuda)], return_A=True)
    torch.cuda.synchronize()
    lb_o, lA_o, lbias_o, und_o = O.dag_crown_planes(P, L, U, threads=8)
    _check(lbs[0].cpu().numpy(), lb_o, f"planes {which} lb")
    assert np.array_equal(unds[0].cpu().numpy(), und_o)
    assert np.array_equal(lAs[0].cpu().numpy(), lA_o)
    assert np.array_equal(lbiases[0].cpu().numpy(), lbias_o)
    print(f"  undefined rows {int(und_o.sum())} of {und_o.size}", flush=True)
