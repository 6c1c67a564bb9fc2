"""Constrained concretization on the device (gh_constrained_concretize_f64):
lower bounds of min obj[b, h] . x over a box intersected with linear cuts
A[b, k] . x + bc[b, k] <= 0 -- constraints_solving after
sort_out_constr_batches and drop_redundant_constraints
(auto_LiRPA/concretize_func.py:269-850), the concretization of
clip_type 'complete' (complete_verifier/input_split/batch_branch_and_bound.py:
420-428)."""
import torch

from ._lib import check, lib, ptr, stream_ptr

MAX_X = 64


def constrained_concretize(obj, A, bc, x_L, x_U, max_rows=16, stream=None):
    """(root (B, H), flags (B) int32, kept (B) int32) for obj (B, H, X),
    cuts A (B, K, X), bc (B, K), boxes x_L, x_U (B, X), all float64 on one
    device.  flags bit 0: a cut misses the box (infeasible domain); bit 1:
    every cut holds on the whole box."""
    B, H, X = obj.shape
    K = A.shape[1] if A is not None else 0
    if X > MAX_X:
        raise ValueError(f"constrained concretization supports X <= {MAX_X}")
    dev = obj.device
    obj = obj.contiguous()
    x_L = x_L.contiguous()
    x_U = x_U.contiguous()
    if K:
        A = A.contiguous()
        bc = bc.contiguous()
        if A.shape != (B, K, X) or bc.shape != (B, K):
            raise ValueError("cuts must be (B, K, X) and (B, K)")
    root = torch.empty(B, H, dtype=torch.float64, device=dev)
    flags = torch.empty(B, dtype=torch.int32, device=dev)
    kept = torch.empty(B, dtype=torch.int32, device=dev)
    check(lib().gh_constrained_concretize_f64(B, H, K, X, ptr(obj), ptr(A if K else None),
                                              ptr(bc if K else None), ptr(x_L), ptr(x_U),
                                              int(max_rows), ptr(root), ptr(flags), ptr(kept),
                                              stream_ptr(stream)),
          "gh_constrained_concretize_f64")
    return root, flags, kept


def constrained_bound(lA, lbias, A, bc, x_L, x_U, lb, max_rows=16, stream=None):
    """In place on lb (B, H): max(lb, lbias + the constrained root of lA),
    +inf on every row of a box a cut misses (gh_constrained_bound_f64).
    A (B, K, X) and bc (B, K) need contiguous rows only (a box stride of
    their own, e.g. views into the frontier's per-domain records).  Returns
    (lb, flags, kept)."""
    B, H, X = lA.shape
    K = A.shape[1]
    if X > MAX_X:
        raise ValueError(f"constrained concretization supports X <= {MAX_X}")
    for t in (lA, lbias, x_L, x_U, lb):
        if not t.is_contiguous():
            raise ValueError("constrained_bound takes contiguous planes, boxes and bounds")
    if A.stride(2) != 1 or A.stride(1) != X or bc.stride(1) != 1:
        raise ValueError("cut rows must be contiguous inside each box")
    flags = torch.empty(B, dtype=torch.int32, device=lA.device)
    kept = torch.empty(B, dtype=torch.int32, device=lA.device)
    check(lib().gh_constrained_bound_f64(B, H, K, X, ptr(lA), ptr(lbias), ptr(A), A.stride(0),
                                         ptr(bc), bc.stride(0), ptr(x_L), ptr(x_U),
                                         int(max_rows), ptr(lb), ptr(flags), ptr(kept),
                                         stream_ptr(stream)),
          "gh_constrained_bound_f64")
    return lb, flags, kept


def clip_domains(lA, lbias, thr, x_L, x_U, out=None, stream=None):
    """Clipped boxes (oL, oU) and state (B) int32 of boxes x_L, x_U (B, n) by
    the planes lA (P, S, n), lbias (P, S) of set b % P against thr (S)
    (gh_clip_domains_f64).  out=(oL, oU) may be (x_L, x_U) for in place."""
    P, S, n = lA.shape
    B = x_L.shape[0]
    x_L = x_L.contiguous()
    x_U = x_U.contiguous()
    oL, oU = out if out is not None else (torch.empty_like(x_L), torch.empty_like(x_U))
    state = torch.empty(B, dtype=torch.int32, device=x_L.device)
    lA, lbias, thr = lA.contiguous(), lbias.contiguous(), thr.contiguous()
    check(lib().gh_clip_domains_f64(B, P, n, S, ptr(lA), ptr(lbias),
                                    ptr(thr), ptr(x_L), ptr(x_U), ptr(oL), ptr(oU),
                                    ptr(state), stream_ptr(stream)), "gh_clip_domains_f64")
    return oL, oU, state
