"""HIP path vs the oracle, through the C ABI: FP64 bounds, planes, branch,
split and prune are compared bit for bit."""
import numpy as np
import pytest
import torch

from gear_amd.problems import QCQP, qcqp_dense, random_bisection_boxes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _dev(a, cuda):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=cuda)


def _bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64).view(np.int64)
    b = np.ascontiguousarray(b, dtype=np.float64).view(np.int64)
    return np.array_equal(a, b)


@pytest.mark.parametrize("n,m,B", [(50, 20, 1000), (8, 3, 1024), (1, 0, 33), (7, 2, 65),
                                   (16, 4, 31), (17, 1, 100), (33, 5, 257), (64, 3, 200),
                                   (11, 2, 70), (22, 3, 130), (41, 4, 90), (55, 2, 60),
                                   (60, 1, 50)])
def test_qcqp_crown_f64_bit_exact(cuda, n, m, B):
    from gear_amd.crown import BoundedQCQP
    print(f"case n={n} m={m} B={B}", flush=True)
    p = qcqp_dense(n, m, seed=n + m)
    L, U = random_bisection_boxes(p.xl, p.xu, B, seed=B)
    lb_o, lA_o, lbias_o = O.qcqp_crown(p.Q, p.c, p.d, L, U, return_A=True)
    net = BoundedQCQP(p, cuda)
    lb, lA, lbias = net.bound(_dev(L, cuda), _dev(U, cuda), return_A=True)
    torch.cuda.synchronize()
    assert _bits_equal(lb.cpu().numpy(), lb_o)
    assert _bits_equal(lA.cpu().numpy(), lA_o)
    assert _bits_equal(lbias.cpu().numpy(), lbias_o)
    # the leading-rows form writes the same planes for the rows it keeps
    R = min(2, p.S)
    lb2, lA2, lbias2 = net.bound(_dev(L, cuda), _dev(U, cuda), return_A=True, plane_rows=R)
    assert lA2.shape == (B, R, n) and lbias2.shape == (B, R)
    assert _bits_equal(lb2.cpu().numpy(), lb_o)
    assert _bits_equal(lA2.cpu().numpy(), lA_o[:, :R])
    assert _bits_equal(lbias2.cpu().numpy(), lbias_o[:, :R])


def test_compute_bounds_interface(cuda):
    """compute_bounds returns (lb, None, A_dict) shaped like the reference."""
    from gear_amd.crown import BoundedQCQP, BoundedTensor, PerturbationLpNorm
    p = qcqp_dense(10, 3, seed=1)
    L, U = random_bisection_boxes(p.xl, p.xu, 12, seed=2)
    net = BoundedQCQP(p, cuda)
    xl, xu = _dev(L, cuda), _dev(U, cuda)
    x = BoundedTensor((xl + xu) / 2, PerturbationLpNorm(x_L=xl, x_U=xu))
    lb, ub, A = net.compute_bounds(x=(x,), method="CROWN", bound_upper=False, return_A=True,
                                   needed_A_dict={net.output_name: [net.input_name]})
    assert ub is None and lb.shape == (12, 4)
    rec = A[net.output_name][net.input_name]
    assert rec["lA"].shape == (12, 4, 10) and rec["lbias"].shape == (12, 4)
    assert _bits_equal(lb.cpu().numpy(), O.qcqp_crown(p.Q, p.c, p.d, L, U))


def test_degenerate_and_single_boxes(cuda):
    from gear_amd.crown import BoundedQCQP
    p = qcqp_dense(12, 2, seed=4)
    L, U = random_bisection_boxes(p.xl, p.xu, 40, seed=3)
    L[5] = U[5]            # a point box
    L[6, 3] = U[6, 3]      # one zero-width coordinate
    net = BoundedQCQP(p, cuda)
    lb = net.bound(_dev(L, cuda), _dev(U, cuda)).cpu().numpy()
    assert _bits_equal(lb, O.qcqp_crown(p.Q, p.c, p.d, L, U))
    one = net.bound(_dev(L[:1], cuda), _dev(U[:1], cuda)).cpu().numpy()
    assert _bits_equal(one, lb[:1])
    empty = net.bound(_dev(L[:0], cuda), _dev(U[:0], cuda))
    assert empty.shape == (0, 3)


def test_batch_independence(cuda):
    """A box's bound does not depend on the batch it is bound in."""
    from gear_amd.crown import BoundedQCQP
    p = qcqp_dense(50, 20, seed=3)
    L, U = random_bisection_boxes(p.xl, p.xu, 300, seed=9)
    net = BoundedQCQP(p, cuda)
    full = net.bound(_dev(L, cuda), _dev(U, cuda)).cpu().numpy()
    part = net.bound(_dev(L[117:250], cuda), _dev(U[117:250], cuda)).cpu().numpy()
    assert _bits_equal(full[117:250], part)


def test_branch_split_bit_exact(cuda):
    from gear_amd.crown import input_split_branching, input_split_parallel
    rng = np.random.default_rng(0)
    B, n = 777, 50
    L = rng.standard_normal((B, n))
    U = L + np.abs(rng.standard_normal((B, n)))
    U[::7, 3] = L[::7, 3] + (U[::7, 0] - L[::7, 0])   # ties
    idx = input_split_branching(None, None, _dev(L, cuda), _dev(U, cuda), None, None, "naive")
    idx_o = O.branch_widest(L, U)
    assert np.array_equal(idx[:, 0].cpu().numpy(), idx_o)
    oL, oU, depth = input_split_parallel(_dev(L, cuda), _dev(U, cuda), i_idx=idx)
    eL, eU = O.split(L, U, idx_o)
    assert depth == 1
    assert _bits_equal(oL.cpu().numpy(), eL) and _bits_equal(oU.cpu().numpy(), eU)


@pytest.mark.parametrize("B", [1, 255, 256, 257, 5000, 70000])
def test_filter_append_matches_oracle(cuda, B):
    from gear_amd.crown import DeviceFrontier
    rng = np.random.default_rng(B)
    n, S = 6, 4
    L = rng.standard_normal((B, n))
    U = L + np.abs(rng.standard_normal((B, n)))
    U[::13, 2] = L[::13, 2] - 1.0                    # invalid boxes
    lb = rng.standard_normal((B, S))
    lb[::17, 0] = np.inf
    lb[::19, 1] = np.nan
    thr = np.array([0.8, 1.0, 1.2, np.inf])
    keep, kept, dm = O.prune(lb, thr, L, U)
    f = DeviceFrontier(n, S, cap=B + 10, device=cuda)
    f.push(_dev(np.zeros((3, n)), cuda), _dev(np.ones((3, n)), cuda),
           _dev(np.zeros((3, S)), cuda))
    got = f.filter_append(_dev(lb, cuda), _dev(thr, cuda), _dev(L, cuda), _dev(U, cuda))
    assert got == kept and len(f) == 3 + kept
    assert _bits_equal(f.L[3:3 + kept].cpu().numpy(), L[keep])
    assert _bits_equal(f.U[3:3 + kept].cpu().numpy(), U[keep])
    assert _bits_equal(f.lb[3:3 + kept].cpu().numpy(), lb[keep])
    assert float(f.disc_min.item()) == dm
    glb = float(f.global_lb().item())
    assert glb == min(dm, np.nanmin(np.concatenate([np.zeros(3), lb[keep][:, 0]])))


def test_bab_iterations_match_oracle_loop(cuda):
    """Six iterations of the regular B&B order on the device reproduce the
    oracle-driven loop (sort by the largest margin lb - thr, descending and
    stable; pick from the end, widest branch, bisect, bound, prune,
    append)."""
    from gear_amd.bab import BaBConfig, QCQPBranchAndBound
    p = qcqp_dense(10, 3, seed=11)
    bb = QCQPBranchAndBound(p, BaBConfig(batch=64, cap=4096, reorder=False), device=cuda)
    bb.root(alm=False)
    bb.set_incumbent(-1.0, None)       # an injected incumbent prunes part of the tree
    # oracle mirror
    FL, FU = p.xl[None].copy(), p.xu[None].copy()
    Flb = O.qcqp_crown(p.Q, p.c, p.d, FL, FU)
    thr = np.concatenate([[-1.0], p.rhs[1:] + 1e-6])
    for it in range(6):
        keys = (Flb - thr).max(axis=1)
        order = np.argsort(-keys, kind="stable")
        FL, FU, Flb = FL[order], FU[order], Flb[order]
        nb = min(64, len(FL))
        pL, pU = FL[-nb:], FU[-nb:]
        FL, FU, Flb = FL[:-nb], FU[:-nb], Flb[:-nb]
        idx = O.branch_widest(pL, pU)
        cL, cU = O.split(pL, pU, idx)
        lb = O.qcqp_crown(p.Q, p.c, p.d, cL, cU)
        keep, _, _ = O.prune(lb, thr, cL, cU)
        FL = np.vstack([FL, cL[keep]])
        FU = np.vstack([FU, cU[keep]])
        Flb = np.vstack([Flb, lb[keep]])
        bb.step()
        c = len(bb.frontier)
        assert c == len(FL), it
        assert _bits_equal(bb.frontier.L[:c].cpu().numpy(), FL)
        assert _bits_equal(bb.frontier.lb[:c].cpu().numpy(), Flb)


def test_child_keys_match_oracle(cuda):
    """gh_child_keys_f64 equals the oracle's child keys bit for bit on the
    planes of real parents, including invalid child boxes and an infinite
    plane (which keep the parent's bound)."""
    from gear_amd.crown import child_keys
    p = qcqp_dense(12, 4, seed=5)
    L, U = random_bisection_boxes(p.xl, p.xu, 200, seed=6)
    plb, lA, lbias = O.qcqp_crown(p.Q, p.c, p.d, L, U, return_A=True)
    lbias[7, 0] = np.inf
    cL, cU = O.split(L, U, O.branch_widest(L, U))
    cL[3, 1], cU[3, 1] = 1.0, -1.0
    ref = O.child_keys(plb, lA, lbias, cL, cU)
    got = child_keys(_dev(plb, cuda), _dev(lA, cuda), _dev(lbias, cuda),
                     _dev(cL, cuda), _dev(cU, cuda)).cpu().numpy()
    assert _bits_equal(got, ref)
    assert (got[:, 0] >= np.concatenate([plb[:, 0], plb[:, 0]])).all()
    assert (got[:, 0] > np.concatenate([plb[:, 0], plb[:, 0]])).any()


@pytest.mark.parametrize("batch,cap,slack,iters", [(32, 4096, None, 6), (4, 128, 20, 14)])
def test_reordered_bab_iterations_match_oracle_loop(cuda, batch, cap, slack, iters):
    """Iterations of the reordered B&B order (the default) on the device
    reproduce the oracle-driven loop: sort, pick 2 batch domains from the
    end, bound them and keep the larger of that bound and the stored key,
    widest branch, bisect, child keys from the parents' planes, prune,
    append.  The second case gives the frontier 20 rows of slack, so it
    compacts its live records during the run, which must change nothing."""
    from gear_amd.bab import BaBConfig, QCQPBranchAndBound
    from gear_amd.crown import DeviceFrontier
    p = qcqp_dense(10, 3, seed=11)
    bb = QCQPBranchAndBound(p, BaBConfig(batch=batch, cap=cap), device=cuda)
    if slack is not None:
        bb.frontier = DeviceFrontier(p.n, p.S, cap, cuda, slack=slack)
    assert bb.cfg.reorder
    bb.root(alm=False)
    bb.set_incumbent(-1.0, None)
    FL, FU = p.xl[None].copy(), p.xu[None].copy()
    Flb = O.qcqp_crown(p.Q, p.c, p.d, FL, FU)
    thr = np.concatenate([[-1.0], p.rhs[1:] + 1e-6])
    raised = 0
    compactions = 0
    for it in range(iters):
        keys = (Flb - thr).max(axis=1)
        order = np.argsort(-keys, kind="stable")
        FL, FU, Flb = FL[order], FU[order], Flb[order]
        nb = min(2 * batch, len(FL))
        pL, pU, pk = FL[-nb:], FU[-nb:], Flb[-nb:]
        FL, FU, Flb = FL[:-nb], FU[:-nb], Flb[:-nb]
        lb, lA, lbias = O.qcqp_crown(p.Q, p.c, p.d, pL, pU, return_A=True)
        lb = np.maximum(lb, pk)
        cL, cU = O.split(pL, pU, O.branch_widest(pL, pU))
        ck = O.child_keys(lb, lA, lbias, cL, cU)
        raised += int((ck[:, 0] > np.concatenate([lb[:, 0], lb[:, 0]])).sum())
        keep, _, _ = O.prune(ck, thr, cL, cU)
        FL = np.vstack([FL, cL[keep]])
        FU = np.vstack([FU, cU[keep]])
        Flb = np.vstack([Flb, ck[keep]])
        hw0 = bb.frontier.hw
        assert bb.step() == nb
        compactions += bb.frontier.hw < hw0
        c = len(bb.frontier)
        assert c == len(FL), it
        assert _bits_equal(bb.frontier.L[:c].cpu().numpy(), FL)
        assert _bits_equal(bb.frontier.U[:c].cpu().numpy(), FU)
        assert _bits_equal(bb.frontier.lb[:c].cpu().numpy(), Flb)
    assert raised > 0
    if slack is not None:
        assert compactions > 0


@pytest.mark.parametrize("K,n,m", [(300, 12, 4), (257, 50, 20), (40, 70, 2), (9, 1, 0)])
def test_branch_split_keys_matches_oracle(cuda, K, n, m):
    """The fused branch + split + child-key kernel equals the oracle's
    widest branch, bisection and child keys, bit for bit, including NaN and
    tied widths, invalid boxes and an infinite plane."""
    from gear_amd.crown import branch_split_keys
    print(f"case K={K} n={n} m={m}", flush=True)
    p = qcqp_dense(n, m, seed=n + m) if n <= 64 else None
    rng = np.random.default_rng(K + n)
    L = -rng.random((K, n)) - 0.1
    U = rng.random((K, n)) + 0.1
    if n > 2:
        U[::5, 1] = L[::5, 1] + (U[::5, 0] - L[::5, 0])    # tied widest widths
        U[3::11, 0] = np.nan                                # NaN width at 0
        U[4::13, 2] = np.nan                                # a later NaN width
        U[6::17, 1] = L[6::17, 1] - 1.0                     # invalid box
    S = m + 1
    if p is not None:
        Lc, Uc = np.nan_to_num(L, nan=0.5), np.nan_to_num(U, nan=0.5)
        plb, lA, lbias = O.qcqp_crown(p.Q, p.c, p.d, np.minimum(Lc, Uc), np.maximum(Lc, Uc),
                                      return_A=True)
    else:
        plb = rng.standard_normal((K, S))
        lA = rng.standard_normal((K, S, n))
        lbias = rng.standard_normal((K, S))
    lbias[7 % K, 0] = np.inf
    eL, eU = O.split(L, U, O.branch_widest(L, U))
    ref = O.child_keys(plb, lA, lbias, eL, eU)
    cL, cU, keys = branch_split_keys(_dev(plb, cuda), _dev(L, cuda), _dev(U, cuda),
                                     _dev(lA, cuda), _dev(lbias, cuda))

    def same(a, b):        # bit for bit, any NaN equal to any NaN
        a, b = np.asarray(a), np.asarray(b)
        nan = np.isnan(a) & np.isnan(b)
        return np.array_equal(np.isnan(a), np.isnan(b)) and _bits_equal(np.where(nan, 0, a),
                                                                         np.where(nan, 0, b))
    assert same(cL.cpu().numpy(), eL)
    assert same(cU.cpu().numpy(), eU)
    assert same(keys.cpu().numpy(), ref)
    # without planes the keys are the parents' bounds
    _, _, k0 = branch_split_keys(_dev(plb, cuda), _dev(L, cuda), _dev(U, cuda))
    assert _bits_equal(k0.cpu().numpy(), np.vstack([plb, plb]))


@pytest.mark.parametrize("K,n,m,planes", [(300, 12, 4, True), (257, 50, 20, True),
                                          (40, 70, 2, True), (9, 1, 0, True), (1, 5, 1, True),
                                          (123, 50, 20, False)])
def test_split_append_equals_separate_calls(cuda, K, n, m, planes):
    """gh_split_append_f64 leaves the frontier (physical rows, slot ids with
    the dead record for pruned children, survivor count, dropped-threshold
    minimum) exactly as branch_split_keys + filter_append(sync=False) do,
    on parents with NaN and tied widths, invalid boxes, an infinite plane,
    verified and pruned children, after rows already in the frontier."""
    from gear_amd.crown import DeviceFrontier, branch_split_keys
    print(f"case K={K} n={n} m={m} planes={planes}", flush=True)
    rng = np.random.default_rng(K * 7 + n)
    S = m + 1
    L = -rng.random((K, n)) - 0.1
    U = rng.random((K, n)) + 0.1
    if n > 2:
        U[::5, 1] = L[::5, 1] + (U[::5, 0] - L[::5, 0])
        U[3::11, 0] = np.nan
        U[4::13, 2] = np.nan
        U[6::17, 1] = L[6::17, 1] - 1.0
    plb = rng.standard_normal((K, S))
    lA = rng.standard_normal((K, S, n)) if planes else None
    lbias = rng.standard_normal((K, S)) if planes else None
    if planes:
        lbias[7 % K, 0] = np.inf
    # thresholds that verify some children on the objective row and some on
    # a constraint row, and keep the rest
    thr = np.full(S, np.inf)
    thr[0] = float(np.median(plb[:, 0])) + 0.5
    if S > 1:
        thr[1] = 1.0
    d = lambda a: None if a is None else _dev(a, cuda)
    fronts = []
    for fused in (False, True):
        f = DeviceFrontier(n, S, 4 * K + 64, device=cuda, slack=64)
        pre = 13
        f.push(_dev(np.zeros((pre, n)), cuda), _dev(np.ones((pre, n)), cuda),
               _dev(np.full((pre, S), -5.0), cuda))
        if fused:
            f.split_append(d(plb), d(L), d(U), d(thr), d(lA), d(lbias), sync=False)
        else:
            cL, cU, keys = branch_split_keys(d(plb), d(L), d(U), d(lA), d(lbias))
            f.filter_append(keys, d(thr), cL, cU, sync=False)
        torch.cuda.synchronize()
        fronts.append(f)
    a, b = fronts
    kept = int(a.total.item())
    assert kept == int(b.total.item())
    if K >= 40:
        assert 0 < kept < 2 * K
    assert a.count == b.count and a.hw == b.hw
    assert torch.equal(a.order[:a.count], b.order[:b.count])
    rows = 13 + kept
    assert _bits_equal(a._L[:rows].cpu().numpy(), b._L[:rows].cpu().numpy())
    assert _bits_equal(a._U[:rows].cpu().numpy(), b._U[:rows].cpu().numpy())
    assert _bits_equal(a._lb[:rows].cpu().numpy(), b._lb[:rows].cpu().numpy())
    assert _bits_equal(a.disc_min.cpu().numpy(), b.disc_min.cpu().numpy())
    assert a.compact_order() == b.compact_order() == 13 + kept


@pytest.mark.parametrize("precision", ["fp64", "fp32s"])
def test_bound_rows_equals_gathered(cuda, precision):
    """Bounding boxes in place through a row index (gh_qcqp_crown_rows_*)
    gives bit for bit the bounds and planes of bounding the gathered boxes,
    and split_append from the index the frontier of split_append from the
    gathered parents."""
    from gear_amd.crown import BoundedQCQP, DeviceFrontier
    p = qcqp_dense(50, 20, seed=3)
    L, U = random_bisection_boxes(p.xl, p.xu, 20000, seed=4)
    net = BoundedQCQP(p, cuda, precision=precision)
    rng = np.random.default_rng(5)
    for B in (17000, 300):        # the staged and the unstaged kernel
        idx = torch.as_tensor(rng.choice(len(L), B, replace=False), dtype=torch.int64,
                              device=cuda)
        dL, dU = _dev(L, cuda), _dev(U, cuda)
        a = net.bound(dL, dU, return_A=True, plane_rows=2, rows=idx)
        b = net.bound(dL[idx], dU[idx], return_A=True, plane_rows=2)
        torch.cuda.synchronize()
        for x, y in zip(a, b):
            assert _bits_equal(x.cpu().numpy(), y.cpu().numpy()), (precision, B)
        thr = torch.full((p.S,), float("inf"), dtype=torch.float64, device=cuda)
        thr[0] = float(torch.quantile(a[0][:, 0], 0.3))
        fr = []
        for rows in (idx, None):
            f = DeviceFrontier(p.n, p.S, 2 * B + 8, device=cuda)
            if rows is None:
                f.split_append(a[0], dL[idx], dU[idx], thr, a[1][:, :1].contiguous(),
                               a[2][:, :1].contiguous())
            else:
                f.split_append(a[0], dL, dU, thr, a[1][:, :1].contiguous(),
                               a[2][:, :1].contiguous(), rows=rows)
            fr.append(f.rows())
        for x, y in zip(*fr):
            assert _bits_equal(x.cpu().numpy(), y.cpu().numpy()), (precision, B)


def test_unsynced_steps_equal_synced(cuda):
    """Steps issued without reading the survivor count (dead slots for the
    pruned children) leave, after compact_order, the same frontier and
    global bound as steps that read it, bit for bit."""
    from gear_amd.bab import BaBConfig, QCQPBranchAndBound
    p = qcqp_dense(12, 4, seed=2)
    out = []
    for sync in (True, False):
        bb = QCQPBranchAndBound(p, BaBConfig(batch=64, cap=1 << 14), device=cuda)
        bb.set_incumbent(-0.8, None)      # prunes part of every step's children
        bb.presplit(256, alm=False)       # every step picks a full 128
        for _ in range(8):
            bb.step(sync=sync)
        if not sync:
            assert bb.frontier.unsynced > 0
        c = bb.frontier.compact_order()
        assert bb.frontier.unsynced == 0
        L, U, lb = bb.frontier.rows()
        out.append((c, L.cpu().numpy(), U.cpu().numpy(), lb.cpu().numpy(),
                    float(bb.frontier.global_lb().item()), bb.stats.boxes_bounded))
    (c0, L0, U0, lb0, g0, n0), (c1, L1, U1, lb1, g1, n1) = out
    assert c0 == c1 and n0 == n1 and g0 == g1
    assert _bits_equal(L0, L1) and _bits_equal(U0, U1) and _bits_equal(lb0, lb1)


def test_compaction_with_unsynced_appends(cuda):
    """A compaction that runs while unsynced appends still hold dead slots
    keeps exactly the survivors (the dead record never becomes a row): after
    appends without a count, picks from the end and a compaction, the
    frontier holds the oracle-pruned survivors in order, bit for bit."""
    from gear_amd.crown import DeviceFrontier
    rng = np.random.default_rng(11)
    n, S, B, P = 5, 3, 200, 150
    thr = np.array([0.0, 0.5, np.inf])
    f = DeviceFrontier(n, S, cap=2000, device=cuda, slack=100)
    order = []                       # logical order: a survivor (L, U, lb) or None (dead)
    compacted = False
    for step in range(14):
        L = rng.standard_normal((B, n))
        U = L + np.abs(rng.standard_normal((B, n)))
        lb = rng.standard_normal((B, S))
        keep, kept, _ = O.prune(lb, thr, L, U)
        assert 0 < kept < B
        hw0 = f.hw
        f.filter_append(_dev(lb, cuda), _dev(thr, cuda), _dev(L, cuda), _dev(U, cuda),
                        sync=False)
        compacted |= f.hw < hw0
        order += [(L[k], U[k], lb[k]) for k in np.flatnonzero(keep)] + [None] * (B - kept)
        f.pick_out(P)
        order = order[:-P]
    assert compacted
    c = f.compact_order()
    live = [r for r in order if r is not None]
    gL, gU, glb = (t.cpu().numpy() for t in f.rows())
    assert c == len(live) == len(gL)
    assert _bits_equal(gL, np.array([r[0] for r in live]))
    assert _bits_equal(gU, np.array([r[1] for r in live]))
    assert _bits_equal(glb, np.array([r[2] for r in live]))


@pytest.mark.parametrize("n,m,B", [(50, 20, 16384 + 37), (17, 4, 20000), (55, 3, 16400),
                                   (41, 2, 16500), (10, 3, 17000)])
def test_qcqp_crown_f64_staged_bit_exact(cuda, n, m, B):
    """Batches of 16,384 boxes or more take the LDS-staged kernel (128 boxes
    per workgroup, Q tiles double-buffered in LDS); bounds and planes stay
    bit-identical to the oracle, including the partial last workgroup."""
    import os
    from gear_amd.crown import BoundedQCQP
    print(f"case n={n} m={m} B={B}", flush=True)
    p = qcqp_dense(n, m, seed=n * m)
    L, U = random_bisection_boxes(p.xl, p.xu, B, seed=B)
    lb_o, lA_o, lbias_o = O.qcqp_crown(p.Q, p.c, p.d, L, U, return_A=True,
                                       threads=min(16, os.cpu_count() or 1))
    net = BoundedQCQP(p, cuda)
    lb, lA, lbias = net.bound(_dev(L, cuda), _dev(U, cuda), return_A=True)
    torch.cuda.synchronize()
    assert _bits_equal(lb.cpu().numpy(), lb_o)
    assert _bits_equal(lA.cpu().numpy(), lA_o)
    assert _bits_equal(lbias.cpu().numpy(), lbias_o)


@pytest.mark.parametrize("N,S", [(1, 3), (100, 3), (8191, 21), (8192, 3), (8193, 40),
                                 (20000, 21), (65536, 21), (70000, 3), (200000, 17), (600000, 3)])
def test_frontier_sort_equals_stable_argsort(cuda, N, S):
    """gh_frontier_sort_f64 orders the rows named by a slot list exactly as
    torch's stable descending argsort of their keys max_s (lb - thr) does
    (ties, NaN, +-inf, -0, a repeated dead slot), and writes the keys."""
    from gear_amd._lib import lib, ptr, stream_ptr
    print(f"case N={N} S={S}", flush=True)
    rng = np.random.default_rng(N)
    rows = N + 50
    lb = np.round(rng.standard_normal((rows, S)) * 3) / 2                # many ties
    lb[rng.random(rows) < 0.03, 1 % S] = np.nan
    lb[rng.random(rows) < 0.02, 0] = np.inf
    lb[rng.random(rows) < 0.02, :] = -np.inf
    lb[rng.random(rows) < 0.05, :] = -0.0
    thr = np.round(rng.standard_normal(S), 1)
    thr[S - 1] = np.inf if N % 2 else thr[S - 1]
    order = rng.permutation(rows)[:N]
    order[rng.random(N) < 0.1] = rows - 1          # slots naming one record
    keys = np.max(lb[order] - thr, axis=1)
    keys[np.isnan(lb[order] - thr).any(axis=1)] = np.nan
    kt = torch.as_tensor(keys, device=cuda)
    perm = torch.argsort(kt, descending=True, stable=True)
    o = torch.as_tensor(order, dtype=torch.int64, device=cuda)
    kO = torch.empty(N, dtype=torch.float64, device=cuda)
    vO = torch.empty(N, dtype=torch.int64, device=cuda)
    kT, vT = torch.empty_like(kO), torch.empty_like(vO)
    assert lib().gh_frontier_sort_f64(N, S, ptr(o), ptr(_dev(lb, cuda)), ptr(_dev(thr, cuda)),
                                      ptr(kO), ptr(vO), ptr(kT), ptr(vT), stream_ptr()) == 0
    assert torch.equal(vO, o[perm])
    got, want = kO.cpu().numpy(), keys[perm.cpu().numpy()]
    assert np.array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    assert np.array_equal(got[ok], want[ok])        # -0 == +0


def test_merge_desc_equals_stable_sort(cuda):
    """gh_merge_desc_f64 of a sorted prefix and a sorted tail equals torch's
    stable descending argsort of the concatenation (ties, NaN, +-inf, -0)."""
    from gear_amd._lib import lib, ptr, stream_ptr
    rng = np.random.default_rng(3)
    # the last two at the loop's sizes: many ties (segments past the LDS
    # capacity, searched in global memory) and continuous keys (in LDS)
    for nA, nB, ties in ((1000, 300, 1), (0, 50, 1), (70, 0, 1), (1, 1, 1), (5000, 4096, 1),
                         (400000, 65536, 1), (300000, 65536, 0)):
        print(f"case nA={nA} nB={nB} ties={ties}", flush=True)
        k = rng.standard_normal(nA + nB) * 3
        if ties:
            k = np.round(k) / 2                                      # many ties
        k[rng.random(nA + nB) < 0.03] = np.nan
        k[rng.random(nA + nB) < 0.02] = np.inf
        k[rng.random(nA + nB) < 0.02] = -np.inf
        k[rng.random(nA + nB) < 0.05] = -0.0
        keys = torch.as_tensor(k, device=cuda)
        vals = torch.arange(nA + nB, dtype=torch.int64, device=cuda)
        pa = torch.argsort(keys[:nA], descending=True, stable=True)
        pb = torch.argsort(keys[nA:], descending=True, stable=True)
        kA, vA = keys[:nA][pa].contiguous(), vals[:nA][pa].contiguous()
        kB, vB = keys[nA:][pb].contiguous(), vals[nA:][pb].contiguous()
        kO = torch.empty(nA + nB, dtype=torch.float64, device=cuda)
        vO = torch.empty(nA + nB, dtype=torch.int64, device=cuda)
        # the prefix as the frontier holds it: already in the order of a stable sort
        ref = torch.argsort(torch.cat([kA, kB]), descending=True, stable=True)
        refv = torch.cat([vA, vB])[ref]
        assert lib().gh_merge_desc_f64(nA, ptr(kA) if nA else None, ptr(vA) if nA else None, nB,
                                       ptr(kB) if nB else None, ptr(vB) if nB else None, ptr(kO),
                                       ptr(vO), stream_ptr()) == 0
        assert torch.equal(vO, refv), (nA, nB)


@pytest.mark.parametrize("case", ["plain", "ties", "nonfinite", "guard", "rows"])
def test_branch_sb_matches_oracle(cuda, case):
    """gh_branch_sb_f64 equals the oracle's smart branching: the score's
    argmax with the lowest index on ties, +inf thresholds taken as the box's
    bound, the batch-wide widest fallback on a non-finite input, the
    exhaustive guard, boxes read through a row index."""
    from gear_amd.crown import branch_sb
    print(f"case sb {case}", flush=True)
    rng = np.random.default_rng(hash(case) % 1000)
    B, S, n = 300, 5, 37
    L = rng.uniform(-1, 0, (B, n))
    U = L + rng.uniform(1e-6, 1.0, (B, n))
    lA = rng.standard_normal((B, S, n)) * (rng.random((B, S, n)) < 0.5)
    lb = rng.standard_normal((B, S))
    thr = np.array([np.inf, 0.3, 0.1, 1.0, 0.0])
    root = np.full(n, 2.0)
    rows = None
    if case == "ties":
        U = L + 0.5
        lA = np.sign(lA)
    elif case == "nonfinite":
        lb[7, 2] = -np.inf
    elif case == "guard":
        U[:, :5] = L[:, :5] + 1e-9           # narrow inputs with large coefficients
        lA[:, :, :5] = 1e3
    elif case == "rows":
        perm = rng.permutation(B)
        rows = perm
    d = lambda a: _dev(a, cuda)
    if rows is None:
        got = branch_sb(d(L), d(U), d(lb), d(thr), d(lA), root_width=d(root))
        want = O.branch_sb(L, U, lb, thr, lA, root_w=root)
    else:
        got = branch_sb(d(L), d(U), d(lb), d(thr), d(lA), root_width=d(root),
                        rows=torch.as_tensor(rows, device=cuda))
        want = O.branch_sb(L[rows], U[rows], lb, thr, lA, root_w=root)
    torch.cuda.synchronize()
    assert np.array_equal(got.cpu().numpy(), want)
    if case == "guard":
        assert np.all(want >= 5)


def test_reordered_bab_with_sb_matches_oracle_loop(cuda):
    """The reordered loop with smart branching (the GEAR presets' 'sb')
    reproduces the oracle-driven loop bit for bit: the picked domains'
    planes of every row give the split input, then children, keys, prune."""
    from gear_amd.bab import BaBConfig, QCQPBranchAndBound
    p = qcqp_dense(10, 3, seed=11)
    batch, iters = 16, 8
    bb = QCQPBranchAndBound(p, BaBConfig(batch=batch, cap=4096, branching="sb"), device=cuda)
    bb.root(alm=False)
    bb.set_incumbent(-1.0, None)
    FL, FU = p.xl[None].copy(), p.xu[None].copy()
    Flb = O.qcqp_crown(p.Q, p.c, p.d, FL, FU)
    thr = np.concatenate([[-1.0], p.rhs[1:] + 1e-6])
    root = p.xu - p.xl
    not_widest = 0
    for it in range(iters):
        keys = (Flb - thr).max(axis=1)
        order = np.argsort(-keys, kind="stable")
        FL, FU, Flb = FL[order], FU[order], Flb[order]
        nb = min(2 * batch, len(FL))
        pL, pU, pk = FL[-nb:], FU[-nb:], Flb[-nb:]
        FL, FU, Flb = FL[:-nb], FU[:-nb], Flb[:-nb]
        lb, lA, lbias = O.qcqp_crown(p.Q, p.c, p.d, pL, pU, return_A=True)
        lb = np.maximum(lb, pk)
        idx = O.branch_sb(pL, pU, lb, thr, lA, root_w=root)
        not_widest += int((idx != O.branch_widest(pL, pU)).sum())
        cL, cU = O.split(pL, pU, idx)
        ck = O.child_keys(lb, lA, lbias, cL, cU)
        keep, _, _ = O.prune(ck, thr, cL, cU)
        FL = np.vstack([FL, cL[keep]])
        FU = np.vstack([FU, cU[keep]])
        Flb = np.vstack([Flb, ck[keep]])
        assert bb.step() == nb
        c = len(bb.frontier)
        assert c == len(FL), it
        assert _bits_equal(bb.frontier.L[:c].cpu().numpy(), FL)
        assert _bits_equal(bb.frontier.U[:c].cpu().numpy(), FU)
        assert _bits_equal(bb.frontier.lb[:c].cpu().numpy(), Flb)
    print(f"sb chose a non-widest input {not_widest} times", flush=True)
    assert not_widest > 0
