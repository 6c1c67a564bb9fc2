"""FP64 LP dual certificate over boxes (include/gearhip.h,
gh_lp_dual_bound_f64): the rigorous lower bound of the objective plane over
the box cut by the constraint planes, as the reference's lp_bound
(complete_verifier/nlp_solver/dual_ascent.py:459-561) computes it per
domain from the CROWN planes (complete_verifier/input_split/
bounding.py:489-520)."""
import torch

from ._lib import check, lib, ptr, stream_ptr


def qcqp_lp_bound(lA, lbias, thr, x_L, x_U, iters=500, return_y=False, stream=None):
    """lb (B,) [, y (B, S - 1)]: certified lower bound of
    min lA[b,0] x + lbias[b,0]  s.t.  lA[b,s] x + lbias[b,s] <= thr[s] (s >= 1),
    x in [x_L[b], x_U[b]]."""
    B, S, n = lA.shape
    # contiguous copies stay bound to locals until the launch is queued: a
    # temporary freed before the call could be handed to the next copy
    lA, lbias, thr = lA.contiguous(), lbias.contiguous(), thr.contiguous()
    x_L, x_U = x_L.contiguous(), x_U.contiguous()
    lb = torch.empty(B, dtype=torch.float64, device=lA.device)
    y = torch.empty(B, max(S - 1, 1), dtype=torch.float64, device=lA.device) if return_y else None
    check(lib().gh_lp_dual_bound_f64(B, n, S, ptr(lA), ptr(lbias), ptr(thr), ptr(x_L),
                                     ptr(x_U), int(iters), ptr(lb), ptr(y),
                                     stream_ptr(stream)), "gh_lp_dual_bound_f64")
    return (lb, y) if return_y else lb
