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
    lb = torch.empty(B, dtype=torch.float64, device=lA.device)
    y = torch.empty(B, max(S - 1, 1), dtype=torch.float64, device=lA.device) if return_y else None
    check(lib().gh_lp_dual_bound_f64(B, n, S, ptr(lA.contiguous()), ptr(lbias.contiguous()),
                                     ptr(thr.contiguous()), ptr(x_L.contiguous()),
                                     ptr(x_U.contiguous()), int(iters), ptr(lb), ptr(y),
                                     stream_ptr(stream)), "gh_lp_dual_bound_f64")
    return (lb, y) if return_y else lb
