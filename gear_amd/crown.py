"""Host-side mirror of the reference's bound / split / branch / prune calls,
dispatching to the gfx950 kernels of libgearhip.so.

Call shapes follow the reference so a caller reads the same:
  * BoundedQCQP.compute_bounds(x=(BoundedTensor,), C=None, method='CROWN',
    bound_upper=False, return_A=True, needed_A_dict={out: {in}}) returns
    (lb (B,S), None, A_dict) like BoundedModule.compute_bounds
    (auto_LiRPA/bound_general.py:1275-1286; backward_bound.py:1098-1111);
  * input_split_branching(net, dom_lb, x_L, x_U, lA, thresholds, 'naive')
    (complete_verifier/input_split/branching_heuristics.py:59-86);
  * input_split_parallel(x_L, x_U, shape, split_depth=1, i_idx=...)
    (complete_verifier/input_split/split.py:26-122);
  * filter_verified_domains-style pruning into a DeviceFrontier
    (complete_verifier/input_split/branching_domains.py:238-558).
Every entry point runs the HIP kernels; there is no CPU or torch fallback.
"""
import torch

from . import _lib
from ._lib import check, lib, ptr, stream_ptr
from .problems import DeviceQCQP, QCQP


class PerturbationLpNorm:
    """Box perturbation (norm = inf) given by x_L, x_U
    (auto_LiRPA/perturbations.py:129-500); only the box form is supported."""

    def __init__(self, eps=None, norm=float("inf"), x_L=None, x_U=None):
        if norm != float("inf"):
            raise ValueError("only the L-inf box perturbation is supported")
        if x_L is None or x_U is None:
            raise ValueError("x_L and x_U are required")
        self.norm, self.x_L, self.x_U = norm, x_L, x_U


class BoundedTensor:
    """A tensor with a box perturbation (auto_LiRPA/bounded_tensor.py)."""

    def __init__(self, x, ptb: PerturbationLpNorm):
        self.data, self.ptb = x, ptb


def _check_boxes(x_L, x_U, n, rows):
    """Number of boxes to bound: x_L.shape[0], or len(rows) when the boxes are
    rows of x_L, x_U (int64 indices; the caller keeps them inside x_L, the
    kernels do not check them)."""
    if x_L.dim() != 2 or x_L.shape != x_U.shape or x_L.shape[1] != n:
        raise ValueError(f"boxes must be (B, {n})")
    if rows is None:
        return x_L.shape[0]
    if rows.dtype != torch.int64 or rows.dim() != 1 or not rows.is_contiguous():
        raise ValueError("rows must be a contiguous int64 vector")
    return rows.shape[0]


class BoundedQCQP:
    """A QCQP program bound on one device.  `compute_bounds` is the CROWN
    lower bound of every row over every box (C = eye(S))."""

    output_name = "/out"
    input_name = "/x"

    # workspace cap of the large-n path (boxes are bounded in chunks that fit)
    LARGE_WS_BYTES = 8 << 30

    def __init__(self, problem: QCQP, device="cuda", precision="fp64", large=None):
        """precision: "fp64" (bit-identical to the FP64 oracle) or "fp32s"
        (FP32 with a rigorous outward margin; never above the FP64 bound).
        n > 64 (or large=True) takes the large-n FP32-sound path, whose dense
        products are batched FP32 GEMMs; FP64 covers n <= 64."""
        if precision not in ("fp64", "fp32s"):
            raise ValueError(f"precision {precision!r}")
        self.large = bool(large) or problem.n > 64
        if self.large and precision != "fp32s":
            raise _lib.GearHipError("the FP64 kernels cover n <= 64; use precision='fp32s'")
        self.problem = problem
        self.precision = precision
        self.dev = DeviceQCQP(problem, device)
        self.device = self.dev.device
        self.n, self.S = problem.n, problem.S
        self.last_flagged = 0
        self.packed = self.packed32 = None
        self._ws = None
        if self.large:
            nb = int(lib().gh_qcqp_large_plan_bytes(self.n, self.S))
            self.plan = torch.empty(nb, dtype=torch.uint8, device=self.device)
            check(lib().gh_qcqp_large_prepare(self.n, self.S, ptr(self.dev.Q), ptr(self.plan),
                                              stream_ptr()), "gh_qcqp_large_prepare")
            return
        nbytes = int(lib().gh_qcqp_pack_bytes(self.n, self.S))
        self.packed = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        check(lib().gh_qcqp_pack_f64(self.n, self.S, ptr(self.dev.Q), ptr(self.packed),
                                     stream_ptr()), "gh_qcqp_pack_f64")
        self.packed32 = None
        if precision == "fp32s":
            nb32 = int(lib().gh_qcqp_pack_f32_bytes(self.n, self.S))
            self.packed32 = torch.empty(nb32, dtype=torch.uint8, device=self.device)
            check(lib().gh_qcqp_pack_f32(self.n, self.S, ptr(self.dev.Q), ptr(self.packed32),
                                         stream_ptr()), "gh_qcqp_pack_f32")
        self.last_flagged = 0

    def bound(self, x_L, x_U, return_A=False, out=None, stream=None, plane_rows=None,
              rows=None):
        """lb (B, S) [, lA (B, R, n), lbias (B, R)] for boxes x_L, x_U (B, n),
        with the planes of the leading R = plane_rows rows (default S).  In
        the FP32-sound form the planes are lA x + lbias <= f_s(x) on the box
        with the planes' rigorous error folded into lbias (n <= 64).  rows
        (B,) int64: bound the boxes rows[b] of x_L, x_U (any number of rows)
        where they are stored, without gathering them."""
        if self.precision == "fp32s":
            return self.bound_fp32s(x_L, x_U, out=out, stream=stream, return_A=return_A,
                                    plane_rows=plane_rows, rows=rows)
        x_L = x_L.contiguous()
        x_U = x_U.contiguous()
        if x_L.dtype != torch.float64 or x_U.dtype != torch.float64:
            raise TypeError("the FP64 path takes float64 boxes")
        B = _check_boxes(x_L, x_U, self.n, rows)
        lb = out if out is not None else torch.empty(B, self.S, dtype=torch.float64,
                                                     device=self.device)
        lA = lbias = None
        R = self.S if plane_rows is None else int(plane_rows)
        if return_A:
            lA = torch.empty(B, R, self.n, dtype=torch.float64, device=self.device)
            lbias = torch.empty(B, R, dtype=torch.float64, device=self.device)
        d = self.dev
        if rows is None:
            check(lib().gh_qcqp_crown_f64(B, self.n, self.S, ptr(self.packed), ptr(d.c),
                                          ptr(d.d), ptr(x_L), ptr(x_U), ptr(lb), ptr(lA),
                                          ptr(lbias), R, stream_ptr(stream)), "gh_qcqp_crown_f64")
        else:
            check(lib().gh_qcqp_crown_rows_f64(B, self.n, self.S, ptr(self.packed), ptr(d.c),
                                               ptr(d.d), ptr(x_L), ptr(x_U), ptr(rows), ptr(lb),
                                               ptr(lA), ptr(lbias), R, stream_ptr(stream)),
                  "gh_qcqp_crown_rows_f64")
        return (lb, lA, lbias) if return_A else lb

    def bound_fp32s(self, x_L, x_U, out=None, stream=None, return_A=False, plane_rows=None,
                    rows=None):
        """FP32-sound lower bounds (B, S) as float64 values, each <= the FP64
        bound of the same box.  `last_flag` marks boxes where a near-tie
        McCormick choice was settled in FP64 (diagnostic).  rows: as bound()
        (n <= 64)."""
        x_L = x_L.contiguous()
        x_U = x_U.contiguous()
        if rows is not None and self.large:
            raise _lib.GearHipError("the large-n FP32-sound path bounds gathered boxes only")
        B = _check_boxes(x_L, x_U, self.n, rows)
        lb = out if out is not None else torch.empty(B, self.S, dtype=torch.float64,
                                                     device=self.device)
        flag = torch.empty(B, dtype=torch.int32, device=self.device)
        d = self.dev
        if self.large:
            per_box = int(lib().gh_qcqp_large_workspace_bytes(self.n, self.S, 1))
            chunk = max(1, min(B, self.LARGE_WS_BYTES // per_box))
            nb = int(lib().gh_qcqp_large_workspace_bytes(self.n, self.S, chunk))
            if self._ws is None or self._ws.numel() < nb:
                self._ws = torch.empty(nb, dtype=torch.uint8, device=self.device)
            check(lib().gh_qcqp_crown_f32s_large(B, self.n, self.S, ptr(self.plan), ptr(d.Q),
                                                 ptr(d.c), ptr(d.d), ptr(x_L), ptr(x_U), ptr(lb),
                                                 ptr(flag), ptr(self._ws), self._ws.numel(),
                                                 stream_ptr(stream)), "gh_qcqp_crown_f32s_large")
            self.last_flag = flag
            return lb
        if return_A and self.large:
            raise _lib.GearHipError("the large-n FP32-sound path has no plane output")
        if self.packed32 is None:
            raise _lib.GearHipError("constructed without precision='fp32s'")
        lA = lbias = None
        R = self.S if plane_rows is None else int(plane_rows)
        if return_A:
            lA = torch.empty(B, R, self.n, dtype=torch.float64, device=self.device)
            lbias = torch.empty(B, R, dtype=torch.float64, device=self.device)
        if rows is None:
            check(lib().gh_qcqp_crown_f32s(B, self.n, self.S, ptr(self.packed32), ptr(d.Q),
                                           ptr(d.c), ptr(d.d), ptr(x_L), ptr(x_U), ptr(lb),
                                           ptr(flag), ptr(lA), ptr(lbias), R, stream_ptr(stream)),
                  "gh_qcqp_crown_f32s")
        else:
            check(lib().gh_qcqp_crown_rows_f32s(B, self.n, self.S, ptr(self.packed32), ptr(d.Q),
                                                ptr(d.c), ptr(d.d), ptr(x_L), ptr(x_U), ptr(rows),
                                                ptr(lb), ptr(flag), ptr(lA), ptr(lbias), R,
                                                stream_ptr(stream)), "gh_qcqp_crown_rows_f32s")
        self.last_flag = flag
        return (lb, lA, lbias) if return_A else lb

    def compute_bounds(self, x=None, C=None, method="CROWN", bound_lower=True,
                       bound_upper=False, return_A=False, needed_A_dict=None, **kwargs):
        if method.lower() not in ("crown", "backward"):
            raise NotImplementedError(f"method {method!r}: only CROWN is on this path")
        if bound_upper or not bound_lower:
            raise NotImplementedError("only lower bounds are on this path")
        if C is not None:
            raise NotImplementedError("C must be the default eye(S)")
        bt = x[0] if isinstance(x, (tuple, list)) else x
        res = self.bound(bt.ptb.x_L, bt.ptb.x_U, return_A=return_A)
        if not return_A:
            return res, None
        lb, lA, lbias = res
        A_dict = {self.output_name: {self.input_name: {
            "lA": lA, "uA": None, "lbias": lbias, "ubias": None, "unstable_idx": None}}}
        return lb, None, A_dict


def branch_sb(x_L, x_U, dom_lb, thresholds, lA, coeff_thresh=1e-3, root_width=None,
              min_width_ratio=1e-3, rows=None, stream=None):
    """(B,) int32 smart-branching choices (gh_branch_sb_f64): the GEAR
    presets' sb heuristic (branching_heuristics.py:25-125, 176-200).
    thresholds (S,); lA (B, S, n); rows: boxes rows[b] of x_L, x_U."""
    B, S, n = lA.shape
    x_L, x_U = x_L.contiguous(), x_U.contiguous()
    dom_lb, thresholds, lA = dom_lb.contiguous(), thresholds.contiguous(), lA.contiguous()
    if dom_lb.shape != (B, S) or thresholds.shape != (S,) or x_L.shape[1] != n:
        raise ValueError("sb branching takes lb (B, S), thresholds (S,), lA (B, S, n)")
    if rows is None and x_L.shape[0] != B:
        raise ValueError("boxes (B, n) or a row index of B")
    rw = None if root_width is None else root_width.contiguous()
    idx = torch.empty(B, dtype=torch.int32, device=lA.device)
    flag = torch.empty(1, dtype=torch.int32, device=lA.device)
    check(lib().gh_branch_sb_f64(B, n, S, ptr(x_L), ptr(x_U), ptr(rows), ptr(dom_lb),
                                 ptr(thresholds), ptr(lA), float(coeff_thresh), ptr(rw),
                                 float(min_width_ratio), ptr(idx), ptr(flag), stream_ptr(stream)),
          "gh_branch_sb_f64")
    return idx


def input_split_branching(net, dom_lb, x_L, x_U, lA, thresholds, branching_method="naive",
                          split_depth=1, **kwargs):
    """(B, 1) int64 indices of the input to split per box: 'naive' the
    widest (ties: lowest index), 'sb' the smart-branching score (branch_sb;
    thresholds (S,), lA (B, S, n), root_width= optional)."""
    if split_depth != 1:
        raise NotImplementedError("depth-1 splits only")
    if branching_method == "sb":
        return branch_sb(x_L, x_U, dom_lb, thresholds, lA,
                         root_width=kwargs.get("root_width")).long().unsqueeze(1)
    if branching_method != "naive":
        raise NotImplementedError(f"branching {branching_method!r}")
    B, n = x_L.shape
    x_L, x_U = x_L.contiguous(), x_U.contiguous()
    idx = torch.empty(B, dtype=torch.int32, device=x_L.device)
    check(lib().gh_branch_widest_f64(B, n, ptr(x_L), ptr(x_U),
                                     ptr(idx), stream_ptr()), "gh_branch_widest_f64")
    return idx.long().unsqueeze(1)


def input_split_parallel(x_L, x_U, shape=None, split_depth=1, i_idx=None, split_partitions=2,
                         split_hint=None):
    """Bisect every box on i_idx[:, 0]: returns (new_x_L, new_x_U, 1) with the
    upper halves first, then the lower halves."""
    if split_depth != 1 or split_partitions != 2 or split_hint is not None:
        raise NotImplementedError("depth-1 bisection only")
    B, n = x_L.shape
    x_L, x_U = x_L.contiguous(), x_U.contiguous()
    idx = i_idx[:, 0].to(torch.int32).contiguous()
    oL = torch.empty(2 * B, n, dtype=x_L.dtype, device=x_L.device)
    oU = torch.empty_like(oL)
    check(lib().gh_split_f64(B, n, ptr(x_L), ptr(x_U), ptr(idx),
                             ptr(oL), ptr(oU), stream_ptr()), "gh_split_f64")
    return oL, oU, 1


def child_keys(plb, lA, lbias, cL, cU, stream=None):
    """Keys (2K, S) of the 2K children of K parents from the parents'
    bounds plb (K, S) and planes lA (K, R, n), lbias (K, R) of R >= 1 rows
    (gh_child_keys_f64, _raise_child_keys of batch_branch_and_bound.py:
    211-225)."""
    K, S = plb.shape
    n = cL.shape[1]
    R = lA.shape[1]
    if lA.shape != (K, R, n) or lbias.shape != (K, R) or not lA.is_contiguous() \
            or not lbias.is_contiguous():
        raise ValueError("planes must be contiguous (K, R, n) and (K, R)")
    plb, cL, cU = plb.contiguous(), cL.contiguous(), cU.contiguous()
    keys = torch.empty(2 * K, S, dtype=torch.float64, device=plb.device)
    check(lib().gh_child_keys_f64(K, n, S, R, ptr(plb), ptr(lA), ptr(lbias),
                                  ptr(cL), ptr(cU), ptr(keys),
                                  stream_ptr(stream)), "gh_child_keys_f64")
    return keys


def branch_split_keys(plb, pL, pU, lA=None, lbias=None, stream=None):
    """Widest branching, bisection and child keys of K parents in one kernel
    (gh_branch_split_keys_f64): (cL, cU (2K, n), keys (2K, S)), equal to
    input_split_branching + input_split_parallel + child_keys."""
    K, S = plb.shape
    n = pL.shape[1]
    R = 1
    if lA is not None:
        R = lA.shape[1]
        if lA.shape != (K, R, n) or lbias.shape != (K, R) or not lA.is_contiguous() \
                or not lbias.is_contiguous():
            raise ValueError("planes must be contiguous (K, R, n) and (K, R)")
    cL = torch.empty(2 * K, n, dtype=torch.float64, device=pL.device)
    cU = torch.empty_like(cL)
    keys = torch.empty(2 * K, S, dtype=torch.float64, device=pL.device)
    pL, pU, plb = pL.contiguous(), pU.contiguous(), plb.contiguous()
    check(lib().gh_branch_split_keys_f64(K, n, S, R, ptr(pL), ptr(pU),
                                         ptr(plb), ptr(lA), ptr(lbias), ptr(cL),
                                         ptr(cU), ptr(keys), stream_ptr(stream)),
          "gh_branch_split_keys_f64")
    return cL, cU, keys


class DeviceFrontier:
    """Device-resident domain storage with stack order (pop from the end),
    as StackTensorStorage (complete_verifier/tensor_storage.py:351-406).

    A domain's record -- x_L, x_U (n), lb (S), and `payload` more doubles
    when asked for (the cut rows that clip_type 'complete' stores with each
    domain) -- stays in one physical row from its append until it is picked
    or pruned.  The stack order is a separate logical array of int64 slot
    ids, so re-sorting the frontier moves 8 bytes per domain, not the
    record.  Appends go to a physical high-water mark; when the slack above
    it runs out, the live records are compacted in logical order.  `L`,
    `U`, `lb`, `pay` and `rows()` give logical-order copies.

    An append without a host sync (filter_append(sync=False)) gives all B
    children a logical slot: survivors their physical rows, the others the
    dead record (lb = +inf, an empty box), so the next step can be issued
    before the survivor count is read.  Dead entries sort to the front,
    never lower the global bound and are pruned if picked; `count` is then
    an upper bound on the live rows until compact_order() (a sync) drops
    them."""

    def __init__(self, n, S, cap, device="cuda", payload=0, slack=None):
        self.n, self.S, self.cap = n, S, int(cap)
        self.device = torch.device(device)
        self.W = int(payload)
        # physical rows: cap plus the slack that lets appends wait for a
        # compaction (children are appended before their parents' rows are
        # reclaimed)
        self.phys = self.cap + int(slack if slack is not None else max(self.cap // 2, 4096))
        self.dead = self.phys           # the dead record's physical row
        self._alloc(self.phys)
        self.order = torch.empty(self.cap, dtype=torch.int64, device=self.device)
        self._order2 = torch.empty_like(self.order)    # the sort's other buffer
        # sort keys aligned with the order; order[:_sorted] is sorted under
        # the thresholds named by _tag (appends and picks keep that prefix)
        self._okeys = torch.empty(self.cap, dtype=torch.float64, device=self.device)
        self._okeys2 = torch.empty_like(self._okeys)
        self._sorted = 0
        self._tag = None
        self._iota = torch.arange(self.cap, dtype=torch.int64, device=self.device)
        self.count = 0          # logical rows
        self.hw = 0             # physical rows in use (live or picked)
        self.total = torch.zeros(1, dtype=torch.int64, device=self.device)
        self.disc_min = torch.full((1,), float("inf"), dtype=torch.float64, device=self.device)
        self._ws = None
        self._sk = None         # sort buffers, allocated by the first sort
        self._pending = None

    def _storage(self, rows):
        """(L, U, lb, payload) arrays of `rows` physical rows plus the dead
        record (an empty box whose bound is +inf on every row) at `rows`."""
        d = dict(dtype=torch.float64, device=self.device)
        L = torch.empty(rows + 1, self.n, **d)
        U = torch.empty(rows + 1, self.n, **d)
        lb = torch.empty(rows + 1, self.S, **d)
        pay = torch.empty(rows + 1, self.W, **d) if self.W else None
        L[rows] = 1.0
        U[rows] = 0.0
        lb[rows] = float("inf")
        if pay is not None:
            pay[rows] = 0.0
            pay[rows, self.S * self.n:] = -1.0
        return L, U, lb, pay

    def _alloc(self, rows):
        self._L, self._U, self._lb, self._pay = self._storage(rows)
        # the compaction's target, made up front so that no step allocates
        self._spare = self._storage(rows)
        # run the compaction's library kernels once (a masked select and row
        # gathers), so their first use inside a step does not load code
        one = torch.zeros(1, dtype=torch.int64, device=self.device)
        one[one != 1]
        for src, dst in zip((self._L, self._U, self._lb, self._pay), self._spare):
            if src is not None:
                torch.index_select(src, 0, one, out=dst[:1])
        self.unsynced = 0         # appends since the last exact count

    def __len__(self):
        return self.count

    # -- logical views (copies), for callers that read rows by position
    def rows(self, start=0, end=None):
        """(x_L, x_U, lb[, payload]) of logical rows start..end, copies."""
        end = self.count if end is None else min(end, self.count)
        idx = self.order[start:end]
        out = (self._L.index_select(0, idx), self._U.index_select(0, idx),
               self._lb.index_select(0, idx))
        if self._pay is not None:
            out = out + (self._pay.index_select(0, idx),)
        return out

    @property
    def L(self):
        return self.rows()[0]

    @property
    def U(self):
        return self.rows()[1]

    @property
    def lb(self):
        return self.rows()[2]

    @property
    def pay(self):
        return self.rows()[3] if self._pay is not None else None

    def push(self, x_L, x_U, lb):
        """Unconditional append (root)."""
        B = x_L.shape[0]
        if self.count + B > self.cap:
            raise RuntimeError("frontier capacity exceeded")
        self._room(B)
        self._L[self.hw:self.hw + B] = x_L
        self._U[self.hw:self.hw + B] = x_U
        self._lb[self.hw:self.hw + B] = lb
        torch.add(self._iota[:B], self.hw, out=self.order[self.count:self.count + B])
        self.count += B
        self.hw += B

    def _room(self, B):
        """Make B physical rows free above the high-water mark, compacting
        the live records in logical order when the slack has run out."""
        if self.hw + B <= self.phys:
            return
        # unsynced appends leave dead slots in the order: drop them first, or
        # the copy below would turn the dead record into ordinary rows
        self.compact_order()
        live = self.order[:self.count]
        # the live rows move, in logical order, into the second storage (kept
        # between compactions, so none allocates after the first)
        old = (self._L, self._U, self._lb, self._pay)
        for src, dst in zip(old, self._spare):
            if src is not None:
                torch.index_select(src, 0, live, out=dst[:self.count])
        self._L, self._U, self._lb, self._pay = self._spare
        self._spare = old
        self.order[:self.count].copy_(self._iota[:self.count])
        self.hw = self.count
        if self.hw + B > self.phys:
            raise RuntimeError("frontier capacity exceeded")

    def sort(self, thr, tag=None):
        """Order the rows by their largest margin max_s (lb_s - thr_s),
        descending, so picking from the end takes the least verified domains
        first: UnsortedInputDomainList.sort with sort_heuristic 'lb_rhs',
        sort_index None and sort_descending, as the GEAR presets set it
        (complete_verifier/input_split/branching_domains.py:1171-1241,
        976-981; gear/presets/gear-f0.yaml:409, 550-552).  Ties keep their
        current order (a stable sort; the reference's argsort leaves tie
        order unspecified).  Only the slot ids move.  tag names thr (e.g.
        the incumbent it holds): when it matches the last sort's, only the
        rows appended since are sorted and merged in."""
        c = self.count
        thr = thr.contiguous()
        p = self._sorted if (tag is not None and tag == self._tag) else 0
        if p >= c:
            self._sorted, self._tag = c, tag
            return
        if self._sk is None:
            # sorted tail and the sort's scratch
            self._sk = torch.empty(self.cap, dtype=torch.float64, device=self.device)
            self._sv = torch.empty(self.cap, dtype=torch.int64, device=self.device)
            self._tk = torch.empty_like(self._sk)
            self._tv = torch.empty_like(self._sv)
        n = c - p
        if p == 0:
            # every row: keys and the stable sort straight into the other buffers
            check(lib().gh_frontier_sort_f64(n, self.S, ptr(self.order), ptr(self._lb),
                                             ptr(thr), ptr(self._okeys2), ptr(self._order2),
                                             ptr(self._tk), ptr(self._tv), stream_ptr()),
                  "gh_frontier_sort_f64")
        else:
            # sort the tail, then merge it behind the sorted prefix (ties:
            # prefix first) -- the stable sort of the whole
            check(lib().gh_frontier_sort_f64(n, self.S, ptr(self.order[p:]), ptr(self._lb),
                                             ptr(thr), ptr(self._sk), ptr(self._sv),
                                             ptr(self._tk), ptr(self._tv), stream_ptr()),
                  "gh_frontier_sort_f64")
            check(lib().gh_merge_desc_f64(p, ptr(self._okeys), ptr(self.order), n, ptr(self._sk),
                                          ptr(self._sv), ptr(self._okeys2), ptr(self._order2),
                                          stream_ptr()), "gh_merge_desc_f64")
        self.order, self._order2 = self._order2, self.order
        self._okeys, self._okeys2 = self._okeys2, self._okeys
        self._sorted, self._tag = c, tag

    def pick_out(self, batch):
        """(x_L, x_U, lb) of the last `batch` logical rows (the parents of the
        next split), gathered; the rows leave the frontier."""
        nb = min(batch, self.count)
        s = self.count - nb
        self.count = s
        self._sorted = min(self._sorted, s)
        idx = self.order[s:s + nb]
        self._picked = idx
        d = dict(dtype=torch.float64, device=self.device)
        pL = torch.empty(nb, self.n, **d)
        pU = torch.empty(nb, self.n, **d)
        plb = torch.empty(nb, self.S, **d)
        check(lib().gh_gather_rows_f64(nb, ptr(idx), self.n, ptr(self._L), ptr(pL), self.n,
                                       ptr(self._U), ptr(pU), self.S, ptr(self._lb), ptr(plb),
                                       stream_ptr()), "gh_gather_rows_f64")
        return pL, pU, plb

    def pick_rows(self, batch):
        """Slot ids and bound rows (idx, lb) of the last `batch` logical rows,
        which leave the frontier as in pick_out; their boxes are read in place
        (rows idx of the frontier's storage, self.storage()) by the bound and
        split_append, which avoids gathering them.  idx is a view of the
        order: the next append overwrites it (split_append reads each entry
        before it writes the same slot)."""
        nb = min(batch, self.count)
        s = self.count - nb
        self.count = s
        self._sorted = min(self._sorted, s)
        idx = self.order[s:s + nb]
        self._picked = idx
        plb = torch.empty(nb, self.S, dtype=torch.float64, device=self.device)
        check(lib().gh_gather_rows_f64(nb, ptr(idx), self.S, ptr(self._lb), ptr(plb), 0, None,
                                       None, 0, None, None, stream_ptr()), "gh_gather_rows_f64")
        return idx, plb

    def storage(self):
        """The physical box arrays (x_L, x_U) that pick_rows' slot ids index."""
        return self._L, self._U

    def in_place_room(self, batch):
        """True when picking `batch` rows with pick_rows and appending their
        2 batch children leaves the picked rows in place (no compaction)."""
        return self.hw + 2 * min(batch, self.count) <= self.phys

    def pick_payload(self, nb):
        """The payload rows of the domains the last pick_out(nb) returned."""
        return self._pay.index_select(0, self._picked[:nb])

    def workspace(self, B, nbytes=None):
        need = int(lib().gh_filter_workspace_bytes(B)) if nbytes is None else int(nbytes)
        if self._ws is None or self._ws.numel() < need:
            self._ws = torch.empty(need, dtype=torch.uint8, device=self.device)
        return self._ws

    def split_append(self, plb, pL, pU, thr, lA=None, lbias=None, sync=True, rows=None,
                     idx=None):
        """Widest branching, bisection, child keys from the parents'
        objective planes, prune and append for K parents in one call
        (gh_split_append_f64): the frontier branch_split_keys followed by
        filter_append of the 2K children would give, without materializing
        the children.  lA (K, R, n) / lbias (K, R) may be None (the keys are
        the parents' bounds).  rows (K,) int64: parent p's box is row rows[p]
        of pL, pU (e.g. this frontier's own storage after pick_rows; the
        append must then not compact it, see in_place_room).  Returns the
        survivor count when sync is set, else None (as filter_append)."""
        if self._pay is not None:
            raise ValueError("a frontier with a payload appends through filter_append")
        K = plb.shape[0] if rows is not None else pL.shape[0]
        B = 2 * K
        R = 1
        if lA is not None:
            R = lA.shape[1]
            if lA.shape != (K, R, self.n) or lbias.shape != (K, R) or not lA.is_contiguous() \
                    or not lbias.is_contiguous():
                raise ValueError("planes must be contiguous (K, R, n) and (K, R)")
        if rows is not None:
            if rows.shape != (K,) or rows.dtype != torch.int64 or pL.shape[1] != self.n:
                raise ValueError("rows must be (K,) int64 over (N, n) boxes")
            if self.hw + B > self.phys and pL.data_ptr() == self._L.data_ptr():
                raise RuntimeError("an append from the frontier's own rows would compact them")
        elif pL.shape != (K, self.n) or pU.shape != (K, self.n):
            raise ValueError("parents must be (K, n) boxes")
        if plb.shape != (K, self.S):
            raise ValueError("parent bounds must be (K, S)")
        if self.count + B > self.cap:
            raise RuntimeError("frontier capacity exceeded")
        self._room(B)
        self._last_B = B
        ws = self.workspace(B, lib().gh_split_append_workspace_bytes(K))
        pL, pU, plb, thr = pL.contiguous(), pU.contiguous(), plb.contiguous(), thr.contiguous()
        tail = (ptr(self._L), ptr(self._U), ptr(self._lb), self.hw, self.phys,
                ptr(self.order[self.count:self.count + B]), self.dead, ptr(self.total),
                ptr(self.disc_min), ptr(ws), stream_ptr())
        if idx is None:
            check(lib().gh_split_append_f64(K, self.n, self.S, R, ptr(pL), ptr(pU), ptr(rows),
                                            ptr(plb), ptr(lA), ptr(lbias), ptr(thr), *tail),
                  "gh_split_append_f64")
        else:
            if idx.shape != (K,) or idx.dtype != torch.int32:
                raise ValueError("idx must be (K,) int32")
            idx = idx.contiguous()
            check(lib().gh_split_append_idx_f64(K, self.n, self.S, R, ptr(pL), ptr(pU),
                                                ptr(rows), ptr(plb), ptr(lA), ptr(lbias),
                                                ptr(thr), ptr(idx), *tail),
                  "gh_split_append_idx_f64")
        self._pending = B
        if sync:
            return self.finish_append()
        self.count += B
        self.hw += B
        self.unsynced += 1
        self._pending = None
        return None

    def filter_append(self, lb, thr, x_L, x_U, sync=True, payload=None):
        """Prune boxes verified against thresholds thr (S,) and append the
        survivors in order.  payload (P, W): box k carries row k % P (needed
        when the frontier holds a payload).  Returns the survivor count (a
        host sync) when sync is set, else None (call finish_append later)."""
        B = x_L.shape[0]
        if self.count + B > self.cap:
            raise RuntimeError("frontier capacity exceeded")
        self._room(B)
        self._last_B = B
        ws = self.workspace(B)
        check(lib().gh_filter_append_f64(B, self.S, self.n, ptr(lb), ptr(thr), ptr(x_L),
                                         ptr(x_U), ptr(self._L), ptr(self._U), ptr(self._lb),
                                         self.hw, self.phys, ptr(self.total),
                                         ptr(self.disc_min), ptr(ws), stream_ptr()),
              "gh_filter_append_f64")
        if self._pay is not None:
            if payload is None or payload.shape[1] != self.W:
                raise ValueError(f"this frontier stores a payload of {self.W} per domain")
            payload = payload.contiguous()
            check(lib().gh_filter_payload_f64(B, self.W, payload.shape[0], ptr(payload), ptr(ws),
                                              ptr(self._pay), self.hw, self.phys, stream_ptr()),
                  "gh_filter_payload_f64")
        # survivor r sits in physical row hw + r; the other children's slots
        # name the dead record
        check(lib().gh_fill_order_i64(B, ptr(self.total), self.hw, self.dead,
                                      ptr(self.order[self.count:self.count + B]), stream_ptr()),
              "gh_fill_order_i64")
        self._pending = B
        if sync:
            return self.finish_append()
        # without the count: every child keeps its slot (dead ones included)
        self.count += B
        self.hw += B
        self.unsynced += 1
        self._pending = None
        return None

    def commit(self, kept):
        """Account the survivors of the last filter_append(sync=False),
        read from `total` later, when the caller wants an exact count:
        drops that append's dead slots (the last B - kept)."""
        B = self._last_B
        self.count -= B - kept
        self.hw -= B - kept
        self._sorted = min(self._sorted, self.count)
        self.unsynced -= 1

    def finish_append(self):
        kept = int(self.total.item())
        self.count += kept
        self.hw += kept
        self._pending = None
        return kept

    def compact_order(self):
        """Drop the dead entries of unsynced appends from the logical order
        (one host sync); `count` is exact afterwards."""
        if self.unsynced == 0:
            return self.count
        o = self.order[:self.count]
        keep = o != self.dead
        live = o[keep]
        c = live.numel()
        self.order[:c] = live
        self.count = c
        self.unsynced = 0
        self._sorted = 0            # keys are recomputed by the next sort
        return c

    def global_lb(self, out=None):
        """min(lb[:, 0] over the frontier, discarded-threshold minimum) as a
        device scalar."""
        out = out if out is not None else torch.empty(1, dtype=torch.float64, device=self.device)
        out.copy_(self.disc_min)
        check(lib().gh_min_rows_f64(self.count, self.S, ptr(self.order), ptr(self._lb), ptr(out),
                                    stream_ptr()), "gh_min_rows_f64")
        return out
