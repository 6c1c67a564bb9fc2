"""CPU oracle -- test infrastructure only.

Python face of the C restatements in oracle/crown_qcqp.c (FP64 CROWN of a
QCQP with the reference's rounding enclosures, widest branching, bisection,
prune) and oracle/crown_dag.c (FP64 CROWN of expression programs: per-node
intermediate bounds, the reference's relaxations of mul / sqr / pow / exp /
log / reciprocal / sin / cos -- the S-shaped ones with their tangent tables in
oracle/sshaped.c -- interval fallback for rows with no relaxation).
Only tests/, the smoke check and the benchmark's CPU-baseline leg may import
this module; the product path never does.

Every function is pinned in tests/test_oracle.py against what fixes it
independently: sampled soundness (lb <= f(x) at vertices and random points),
closed forms (a pure linear row gives the exact box minimum within the
enclosure margin; a single bilinear term gives the exact McCormick minimum),
brute force on tiny inputs, and the reference's own explicit expectations
for split / branch / prune (CV/tests/test_input_split_split.py and
branching_domains.py semantics).  The program oracle is pinned in
tests/test_oracle_dag.py: the reference's golden CROWN bounds of x^4, x^6,
non-integer powers and sin (solver/tests/data_64/pow_test_data and
s_shape_test_data, extracted to tests/golden/), the tangent tables against
the tangency conditions they solve, closed forms of the exp / log /
reciprocal / square / cubic / cos planes, cancellation invariants of the
intermediate bounds, and sampled soundness of every operator and of random
programs.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None
_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_I32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_U8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        h = ctypes.CDLL(LIB_PATH)
        h.gho_qcqp_crown.argtypes = [ctypes.c_int] * 3 + [_D] * 5 + [ctypes.c_void_p] * 3 + [ctypes.c_int]
        h.gho_qcqp_crown.restype = ctypes.c_int
        h.gho_branch_widest.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D, _I32]
        h.gho_branch_widest.restype = None
        h.gho_split.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D, _I32, _D, _D]
        h.gho_split.restype = None
        h.gho_prune.argtypes = [ctypes.c_int] * 3 + [_D] * 4 + [_U8, ctypes.POINTER(ctypes.c_double)]
        h.gho_prune.restype = ctypes.c_long
        _C = ctypes.c_int
        _V = ctypes.c_void_p
        h.gho_constrained_concretize.argtypes = [_C] * 4 + [_D] * 5 + [_C, _V, _I32, _I32, _V, _V]
        h.gho_constrained_concretize.restype = ctypes.c_int
        h.gho_constraint_rows.argtypes = [_C, _C, _D, _D, _D, _D, _C, _I32, _D, _I32]
        h.gho_constraint_rows.restype = ctypes.c_int
        h.gho_clip_domains.argtypes = [_C] * 4 + [_D] * 7 + [_I32]
        h.gho_clip_domains.restype = ctypes.c_int
        _lib = h
    return _lib


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def qcqp_crown(Q, c, d, xL, xU, return_A=False, threads=1):
    """lb (B, S) [, lA (B, S, n), lbias (B, S)] -- FP64 oracle bound."""
    Q, c, d, xL, xU = map(_c, (Q, c, d, xL, xU))
    S, n = Q.shape[0], Q.shape[1]
    B = xL.shape[0]
    lb = np.empty((B, S))
    lA = np.empty((B, S, n)) if return_A else None
    lbias = np.empty((B, S)) if return_A else None
    p = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)
    lib().gho_qcqp_crown(B, n, S, Q, c, d, xL, xU, p(lb), p(lA), p(lbias), threads)
    if return_A:
        return lb, lA, lbias
    return lb


def branch_widest(xL, xU):
    xL, xU = _c(xL), _c(xU)
    idx = np.empty(xL.shape[0], dtype=np.int32)
    lib().gho_branch_widest(xL.shape[0], xL.shape[1], xL, xU, idx)
    return idx


def split(xL, xU, idx):
    xL, xU = _c(xL), _c(xU)
    B, n = xL.shape
    oL = np.empty((2 * B, n))
    oU = np.empty((2 * B, n))
    lib().gho_split(B, n, xL, xU, np.ascontiguousarray(idx, dtype=np.int32), oL, oU)
    return oL, oU


def prune(lb, thr_rows, xL, xU, disc_min=np.inf):
    """keep mask (B,), kept count, updated discarded-threshold minimum.
    thr_rows (S,) is shared by every box."""
    lb, xL, xU = _c(lb), _c(xL), _c(xU)
    B, S = lb.shape
    thr = np.ascontiguousarray(np.broadcast_to(np.asarray(thr_rows, dtype=np.float64), (B, S)))
    keep = np.empty(B, dtype=np.uint8)
    dm = ctypes.c_double(disc_min)
    kept = lib().gho_prune(B, S, xL.shape[1], lb, thr, xL, xU, keep, ctypes.byref(dm))
    return keep.astype(bool), int(kept), dm.value


# ------------------------------------------------------------ expression DAGs
class _GhProg(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("V", ctypes.c_int32), ("R", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("nodes", ctypes.c_void_p), ("out", ctypes.c_void_p)]


def _prog_struct(prog):
    nodes, out = prog.arrays()
    st = _GhProg(prog.n, prog.V, prog.R, 0, nodes.ctypes.data_as(ctypes.c_void_p),
                 out.ctypes.data_as(ctypes.c_void_p))
    return st, (nodes, out)   # keep the arrays alive with the struct


def branch_sb(L, U, lb, thr, lA, coeff_thresh=1e-3, root_w=None, min_width_ratio=1e-3):
    """(B,) int32 smart-branching choices (gho_branch_sb)."""
    L, U, lb, lA = _c(L), _c(U), _c(lb), _c(lA)
    thr = _c(np.asarray(thr, dtype=np.float64))
    B, S, n = lA.shape
    idx = np.empty(B, dtype=np.int32)
    h = lib()
    h.gho_branch_sb.argtypes = [ctypes.c_int] * 3 + [_D] * 5 + [ctypes.c_double, ctypes.c_void_p,
                                                               ctypes.c_double, _I32]
    h.gho_branch_sb.restype = None
    rw = None if root_w is None else _c(np.asarray(root_w, dtype=np.float64))
    h.gho_branch_sb(B, n, S, L, U, lb, thr, lA, float(coeff_thresh),
                    None if rw is None else rw.ctypes.data_as(ctypes.c_void_p),
                    float(min_width_ratio), idx)
    return idx


def _dag_lib():
    h = lib()
    if not hasattr(h, "_dag_ready"):
        h.gho_dag_crown.argtypes = [ctypes.POINTER(_GhProg), ctypes.c_int, _D, _D, _D, ctypes.c_int]
        h.gho_dag_crown.restype = ctypes.c_int
        h.gho_dag_ibp.argtypes = [ctypes.POINTER(_GhProg), ctypes.c_int, _D, _D, _D, _D]
        h.gho_dag_ibp.restype = ctypes.c_int
        h.gho_dag_crown_planes.argtypes = [ctypes.POINTER(_GhProg), ctypes.c_int, _D, _D, _D, _D,
                                           _D, _U8, ctypes.c_int]
        h.gho_dag_crown_planes.restype = ctypes.c_int
        h.gho_dag_set_flags.argtypes = [ctypes.c_int]
        h.gho_dag_set_flags.restype = None
        h.gho_relax_unary.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, _D, _D, _D,
                                      ctypes.c_int]
        h.gho_relax_unary.restype = ctypes.c_int
        h.gho_sshape_tables.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        h.gho_sshape_tables.restype = ctypes.c_int
        h._dag_ready = True
    return h


SIN_LEGACY_INDEX = 1   # sshaped.h GHO_SIN_LEGACY_INDEX


def dag_crown(prog, xL, xU, threads=1, flags=0):
    """lb (B, R): FP64 oracle CROWN lower bound of every row of an
    expression program (gear_amd.program.Program) over boxes xL, xU (B, n).
    flags=SIN_LEGACY_INDEX rounds sin's right-endpoint table index down, as
    the reference did when it wrote its golden sin bounds."""
    xL, xU = _c(np.atleast_2d(xL)), _c(np.atleast_2d(xU))
    st, keep = _prog_struct(prog)
    lb = np.empty((xL.shape[0], prog.R))
    h = _dag_lib()
    h.gho_dag_set_flags(int(flags))
    try:
        h.gho_dag_crown(ctypes.byref(st), xL.shape[0], xL, xU, lb, threads)
    finally:
        h.gho_dag_set_flags(0)
    return lb


def dag_crown_planes(prog, xL, xU, threads=1):
    """(lb (B, R), lA (B, R, n), lbias (B, R), undefined (B, R) bool): the
    bounds and each row's lower plane lA x + lbias over the inputs, before
    concretization (backward_bound.py:1093-1111); a row whose relaxation
    was undefined has A = 0 and bias -inf (:445-463)."""
    xL, xU = _c(np.atleast_2d(xL)), _c(np.atleast_2d(xU))
    st, keep = _prog_struct(prog)
    B = xL.shape[0]
    lb = np.empty((B, prog.R))
    lA = np.empty((B, prog.R, prog.n))
    lbias = np.empty((B, prog.R))
    und = np.empty((B, prog.R), dtype=np.uint8)
    _dag_lib().gho_dag_crown_planes(ctypes.byref(st), B, xL, xU, lb, lA, lbias, und, threads)
    return lb, lA, lbias, und.astype(bool)


def relax_unary(op, c, l, u, flags=0):
    """(N, 4) relaxation planes (lw, lb, uw, ub) of sin / cos / an odd or
    non-integer power over intervals [l, u], and the undefined sides."""
    l, u = _c(np.atleast_1d(l)), _c(np.atleast_1d(u))
    R = np.zeros((l.size, 4))
    ud = _dag_lib().gho_relax_unary(int(op), float(c), l.size, l, u, R, int(flags))
    return R, ud


def sshape_tables(which, p=0):
    """The tangent tables: which 0 -> sin (d_lower, d_upper) (2, 162);
    1 -> sin endpoint tables (9, 1001): x, then (d, cycles) of xl_lower,
    xl_upper, xu_lower, xu_upper; 2 -> odd power p (d_lower, d_upper)
    (2, 50005)."""
    h = _dag_lib()
    n = h.gho_sshape_tables(which, p, None)
    out = np.empty(n)
    h.gho_sshape_tables(which, p, out.ctypes.data_as(ctypes.c_void_p))
    return out.reshape(9 if which == 1 else 2, -1)


def dag_ibp(prog, xL, xU):
    xL, xU = _c(np.atleast_2d(xL)), _c(np.atleast_2d(xU))
    st, keep = _prog_struct(prog)
    lo = np.empty((xL.shape[0], prog.R))
    hi = np.empty((xL.shape[0], prog.R))
    _dag_lib().gho_dag_ibp(ctypes.byref(st), xL.shape[0], xL, xU, lo, hi)
    return lo, hi


def math_fn(name, x):
    """The shared elementary functions of include/gearhip_math.h on an
    array (name: exp, expm1, log, sin, cos, asin, pow2 .. pow6)."""
    codes = {"exp": 0, "expm1": 1, "log": 2, "sin": 3, "cos": 4, "asin": 5}
    code = codes[name] if name in codes else 10 + int(name[3:])
    x = _c(x)
    y = np.empty_like(x)
    h = lib()
    h.gho_math.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D]
    h.gho_math.restype = ctypes.c_int
    h.gho_math(code, x.size, x, y)
    return y


# ------------------------------------------------------------- LP certificate
U64 = 2.0 ** -53
ETA64 = 2.0 ** -1074


def _gamma(k):
    return k * U64 / (1.0 - k * U64)


def lp_dual_value(c, c0, G, h, l, u, y):
    """Rigorous lower bound of the weak-duality value
        D(y) = c0 - h^T y + sum_j min(r_j l_j, r_j u_j),  r = c + G^T y,
    of  min c^T x + c0  s.t.  G x <= h,  l <= x <= u  (finite boxes), from
    FP64 operations: the single-pass error bound of
    complete_verifier/nlp_solver/rigorous_dual.py:1-54, 118-168 (Higham,
    Lemma 3.1).  c, l, u (n,), G (m, n), h, y (m,); returns lb <= D(y)."""
    c, G, h, l, u = (np.asarray(a, dtype=np.float64) for a in (c, G, h, l, u))
    y = np.maximum(np.asarray(y, dtype=np.float64), 0.0)
    m, n = G.shape
    infl = 1.0 + _gamma(2 * (m + n) + 16)
    act = (y[:, None] != 0) & (G != 0)
    P = np.where(act, G * y[:, None], 0.0)
    r = P.sum(0) + c
    rho = np.abs(P).sum(0) + np.abs(c)
    k = act.sum(0).astype(np.float64)
    kk = np.where(k > 0, _gamma(k + 1.0), 0.0)
    eps = kk * rho + 4 * ETA64 * (k + 2)
    t = np.minimum(r * l, r * u)
    M = np.maximum(np.abs(l), np.abs(u))
    hy = np.where(y > 0, h * y, 0.0)
    val = c0 - hy.sum() + t.sum()
    N = 1.0 + (y > 0).sum() + (t != 0).sum()
    E = _gamma(N + 2) * (abs(c0) + np.abs(hy).sum() + np.abs(t).sum()) + (M * eps).sum() \
        + (m + n + 3) * ETA64
    E = E * infl
    E = E + 4 * U64 * (abs(val) + E)
    lb = val - E
    return lb if np.isfinite(lb) else -np.inf


def constrained_concretize(obj, A, b, xL, xU, max_rows=16):
    """Lower bounds root (B, H) of min obj[b, h] . x over the box [xL, xU]
    intersected with the cuts A[b, k] . x + b[b, k] <= 0 (constraints_solving
    with sort_out_constr_batches, drop_redundant_constraints and the
    sequential dual sweep, concretize_func.py:269-850; oracle/concretize.c
    states the order of evaluation).  Returns (root, flags, kept): flags bit
    0 = some cut misses the box (infeasible), bit 1 = every cut holds on the
    whole box; kept = rows the sweep visited."""
    obj, A, b, xL, xU = (_c(v) for v in (obj, A, b, xL, xU))
    B, H, X = obj.shape
    K = A.shape[1]
    root = np.empty((B, H))
    flags = np.empty(B, dtype=np.int32)
    kept = np.empty(B, dtype=np.int32)
    rc = lib().gho_constrained_concretize(B, H, K, X, obj, A.reshape(B, K, X) if K else
                                          np.zeros((B, 1, X)), b.reshape(B, K) if K else
                                          np.zeros((B, 1)), xL, xU, int(max_rows),
                                          root.ctypes.data, flags, kept, None, None)
    if rc:
        raise ValueError("unsupported size (X <= 64)")
    return root, flags, kept


def constrained_bound(lA, lbias, A, b, xL, xU, lb, max_rows=16):
    """Bounds (B, S) of the planes lA x + lbias over the boxes and cuts:
    max(lb, lbias widened + the constrained root), +inf where a cut misses
    the box (gh_constrained_bound_f64's contract).  Returns (bound, flags,
    kept)."""
    lA, lbias, A, b, xL, xU, lb = (_c(v) for v in (lA, lbias, A, b, xL, xU, lb))
    B, H, X = lA.shape
    K = A.shape[1]
    out = lb.copy()
    flags = np.empty(B, dtype=np.int32)
    kept = np.empty(B, dtype=np.int32)
    rc = lib().gho_constrained_concretize(B, H, K, X, lA, A, b, xL, xU, int(max_rows), None,
                                          flags, kept, lbias.ctypes.data, out.ctypes.data)
    if rc:
        raise ValueError("unsupported size (X <= 64)")
    return out, flags, kept


def constraint_rows(A, b, xL, xU, max_rows=0):
    """Rows (indices, centred offsets d, flags) the sweep of one box visits
    (drop_redundant_constraints after sort_out_constr_batches)."""
    A, b, xL, xU = (_c(v) for v in (A, b, xL, xU))
    K, X = A.shape
    keep = np.empty(max(K, 1), dtype=np.int32)
    d = np.empty(max(K, 1))
    flags = np.empty(1, dtype=np.int32)
    nk = lib().gho_constraint_rows(K, X, A, b, xL, xU, int(max_rows), keep, d, flags)
    return keep[:nk].copy(), d[:nk].copy(), int(flags[0])


def clip_domains(lA, lbias, thr, xL, xU):
    """Clipped boxes (oL, oU) and state (B) (0 skipped: non-finite data,
    1 clipped, 2 certified and emptied, 3 clipped to empty) of boxes
    xL, xU (B, n) by the planes lA (P, S, n), lbias (P, S) of set b % P
    against thresholds thr (S) (clip_domains, complete_verifier/input_split/
    clip.py:147-503, one iteration; oracle/concretize.c)."""
    lA, lbias, thr, xL, xU = (_c(v) for v in (lA, lbias, thr, xL, xU))
    P, S, n = lA.shape
    B = xL.shape[0]
    oL = np.empty_like(xL)
    oU = np.empty_like(xU)
    state = np.empty(B, dtype=np.int32)
    if lib().gho_clip_domains(B, P, n, S, lA, lbias, thr, xL, xU, oL, oU, state):
        raise ValueError("unsupported size (n <= 64)")
    return oL, oU, state


def _red64s(T):
    """Row sums of T (B, n) in the red64s order (see plane_minimum_down)."""
    B, n = T.shape
    P = np.zeros((B, 64))
    for j0 in range(0, n, 64):
        w = min(64, n - j0)
        P[:, :w] = P[:, :w] + T[:, j0:j0 + w]
    lanes = np.arange(64)
    for o in (32, 16, 8, 4, 2, 1):
        P = P + P[:, lanes ^ o]
    return P[:, 0]


def plane_minimum_down(a, b, xl, xu):
    """Minimum of the plane a x + b over the box, never above the exact value:
    sum(a c - |a| r) + b lowered by (n + 4) u (sum |a| (|c| + r) + |b|) and
    one step toward -inf (complete_verifier/input_split/
    batch_branch_and_bound.py:189-208).  Both sums run in the "red64s"
    order: slot l sums the terms i = l, l + 64, ... in order from +0.0, then
    p[l] + p[l ^ o] for o = 32, 16, 8, 4, 2, 1 (the margin covers any order;
    this one is the kernels')."""
    one = np.ndim(xl) == 1
    a, xl, xu = (np.atleast_2d(np.asarray(v, dtype=np.float64)) for v in (a, xl, xu))
    c, r = (xl + xu) / 2, (xu - xl) / 2
    pm = _red64s(a * c - np.abs(a) * r)
    mg = _red64s(np.abs(a) * (np.abs(c) + r))
    pm = pm + b
    margin = (xl.shape[1] + 4) * U64 * (mg + np.abs(b))
    out = np.nextafter(pm - margin, -np.inf)
    return out[0] if one else out


def child_keys(plb, lA, lbias, cL, cU):
    """Keys (2K, S) of the 2K children ([upper; lower] halves) of K parents
    (batch_branch_and_bound.py:211-225): row 0 max(parent bound, the
    parent's objective plane minimised over the child box), rows >= 1 the
    parent's bound."""
    K = plb.shape[0]
    keys = np.vstack([plb, plb]).copy()
    par = np.arange(2 * K) % K
    pm = plane_minimum_down(lA[par, 0], lbias[par, 0], cL, cU)
    ok = np.all(cL <= cU, axis=1) & np.isfinite(pm)
    keys[ok, 0] = np.maximum(keys[ok, 0], pm[ok])
    return keys


# ------------------------------------------------------------- ALM schedule
def _rpow(rho, a):
    return np.sqrt(rho) if a == 0.5 else rho ** a


def alm_schedule_check(cfg, trace, S, rtol_pow=2e-15):
    """Replays the outer schedule of ALM-PGD restarts from their trace and
    lists every disagreement (empty when the kernel followed it).

    trace (N, outer_max_iter + 1, 8 + 2 (S - 1)): record 0 the start, record
    k + 1 after outer iteration k: rho, eps_k, eta_k after the update, the
    violation, the objective, inner steps, flag (0 updated, 1 stopped before
    the update, 2 not run), the last inner gradient norm, mu after the
    update, g = f - rhs before it.  From the recorded objective and g alone
    it restates, per restart as the reference runs one restart of one
    domain (complete_verifier/nlp_solver/alm.py):
      * start: rho0, eps = rho^-alpha, eta = rho^-beta, mu = 0; the start
        point is the first best point (:1380-1390, 1505-1522);
      * the violation max(0, max_s g_s) (:1633-1645);
      * best point: feasible first, else least violated (:1650-1700);
      * exits: objective patience, then first-feasibility patience
        (:1800-1849);
      * BaB mode (update_multipliers_always false): where the violation is
        within eps_k, mu <- max(mu + rho g, 0) and eps_k, eta_k divided by
        rho; else rho <- min(kappa rho, max_rho) when the violation exceeds
        1e-3 and eps_k, eta_k reset to rho^-alpha, rho^-beta (:1884-1907);
        root mode: mu updated every iteration, rho raised when the
        violation fell by less than 1.25x and exceeds feas_tol (:1857-1883);
        both floored at eps_final, eta_final (:1913-1914);
      * an inner loop that ends early ended at its stationarity test
        (gradient norm <= eta_k of that iteration, :764-776) or at a plateau
        check (:678-699).
    rho and mu must match exactly; eps_k, eta_k within rtol_pow (x^0.8 is a
    library power)."""
    tr = np.asarray(trace, dtype=np.float64)
    m = S - 1
    probs = []
    tol = cfg.feas_tol
    for r in range(tr.shape[0]):
        R = tr[r]
        rho = cfg.rho0
        eps, eta = 1.0 / _rpow(rho, cfg.alpha), 1.0 / _rpow(rho, cfg.beta)
        mu = np.zeros(m)
        rec = R[0]
        g = rec[8 + m:8 + 2 * m]
        viol = max(0.0, float(g.max())) if m else 0.0
        if rec[3] != viol:
            probs.append((r, 0, "start violation", rec[3], viol))
        best, best_viol, best_feas = rec[4], viol, viol <= tol
        prev_best, glob, prev_viol, no_imp, plat = np.inf, np.inf, np.inf, 0, -1
        for name, got, want in (("rho", rec[0], rho), ("eps", rec[1], eps), ("eta", rec[2], eta)):
            if abs(got - want) > rtol_pow * abs(want):
                probs.append((r, 0, name, got, want))
        stopped = False
        for k in range(cfg.outer_max_iter):
            rec = R[k + 1]
            flag = int(rec[6])
            if stopped:
                if flag != 2:
                    probs.append((r, k, "ran after its stop", flag, 2))
                continue
            if flag == 2:
                probs.append((r, k, "not run before a stop", flag, 0))
                break
            g = rec[8 + m:8 + 2 * m]
            phi = rec[4]
            viol = max(0.0, float(g.max())) if m else 0.0
            if rec[3] != viol:
                probs.append((r, k, "violation", rec[3], viol))
            steps, gnorm = int(rec[5]), rec[7]
            if steps < cfg.inner_steps:
                plateau_ok = (cfg.inner_plateau_check > 0 and steps % cfg.inner_plateau_check == 0
                              and steps >= cfg.inner_plateau_check * (cfg.inner_plateau_patience + 1))
                if not (gnorm <= eta or plateau_ok):
                    probs.append((r, k, "early inner exit", steps, gnorm))
            feas = viol <= tol
            improved = feas and (not best_feas or phi < best)
            better = (not feas) and (not best_feas) and viol < best_viol
            if improved or better:
                best, best_viol = phi, viol
                best_feas = best_feas or improved
            stop = False
            if cfg.obj_patience >= 0:
                no_imp = 0 if best < prev_best - cfg.obj_improvement_tol else no_imp + 1
                prev_best = min(prev_best, best)
                stop = no_imp >= cfg.obj_patience and best_feas
            if not stop and cfg.first_feas_patience >= 0 and best_feas:
                if plat < 0 or best < glob - cfg.obj_improvement_tol:
                    plat, glob = k, best
                elif k - plat >= cfg.first_feas_patience:
                    stop = True
            if stop != (flag == 1):
                probs.append((r, k, "stop", flag, stop))
            if not stop:
                if cfg.update_multipliers_always:
                    mu = np.maximum(mu + rho * g, 0.0)
                    if k >= 1:
                        ratio = prev_viol / max(viol, 1e-12)
                        if ratio < 1.25 and viol > tol:
                            rho = min(cfg.kappa * rho, cfg.max_rho)
                    eps, eta = 1.0 / _rpow(rho, cfg.alpha), 1.0 / _rpow(rho, cfg.beta)
                    prev_viol = viol
                elif viol <= eps:
                    mu = np.maximum(mu + rho * g, 0.0)
                    eps, eta = eps / rho, eta / rho
                else:
                    if viol > 1e-3:
                        rho = min(cfg.kappa * rho, cfg.max_rho)
                    eps, eta = 1.0 / _rpow(rho, cfg.alpha), 1.0 / _rpow(rho, cfg.beta)
                eps, eta = max(eps, cfg.eps_final), max(eta, cfg.eta_final)
            else:
                stopped = True
            if rec[0] != rho:
                probs.append((r, k, "rho", rec[0], rho))
            for name, got, want in (("eps", rec[1], eps), ("eta", rec[2], eta)):
                if abs(got - want) > rtol_pow * abs(want):
                    probs.append((r, k, name, got, want))
            if m and not np.array_equal(rec[8:8 + m], mu):
                probs.append((r, k, "mu", rec[8:8 + m], mu))
    return probs
