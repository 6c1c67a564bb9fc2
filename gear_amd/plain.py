"""The plain FP64 GPU path: the same CROWN algorithm for a QCQP executed op
by op with PyTorch-ROCm tensors (no fused kernels, no graphs), as the
reference's tensor code runs it (auto_LiRPA/backward_bound.py:880-1206 with
the Linear / Mul / ReduceSum / Add operators).  It is the baseline the fused
kernels are measured against, not a fallback: nothing in the product path
calls it.  Matrix products use the library's own summation order, so its
results agree with the fused path to rounding (about 1e-13 relative), not
bit for bit.
"""
import torch

EPS = 2.220446049250313080847e-16


def _k(n):
    return (n + 16) * EPS * 4


def _h(k):
    return (k + 1) * EPS / 2 * (1 + 1e-6)


def _widen(b, e, sign):
    w = torch.where(torch.isfinite(b), b + sign * e, b)
    return torch.where(torch.isposinf(e), sign * torch.inf, w)


def _add(a, b):
    r = a + b
    return _widen(r, torch.where(torch.isfinite(r), r.abs(), 0) * _k(1), -1)


def plain_qcqp_crown(dev, L, U, return_A=False, plane_rows=None, rows=None, **kw):
    """lb (B, S) for boxes L, U (B, n) on the device problem `dev`; with
    return_A also the planes lA (B, R, n) and lbias (B, R) of the leading
    R = plane_rows rows (the bias before concretization, as the fused path
    returns it).  rows: bound the boxes rows[b] of L, U."""
    if rows is not None:
        L, U = L.index_select(0, rows), U.index_select(0, rows)
    S, n = dev.S, dev.n
    B = L.shape[0]
    W = dev.Q.reshape(S * n, n)
    Wa = W.abs()
    x0 = (U + L) / 2
    ep = (U - L) / 2
    mx = torch.maximum(-L, U)
    c = (x0 @ W.t()).view(B, S, n)
    dv = (ep @ Wa.t()).view(B, S, n)
    mg = (mx @ Wa.t()).view(B, S, n)
    e = mg * _k(n)
    yl = _widen(c - dv, e, -1)
    yu = _widen(c + dv, e, 1)
    xl = L[:, None, :].expand(B, S, n)
    xu = U[:, None, :].expand(B, S, n)
    corners = torch.stack([xl * yl, xl * yu, xu * yl, xu * yu])
    ql = corners.amin(0).sum(-1)
    qu = corners.amax(0).sum(-1)
    Ca = dev.c.abs()
    lc, ldv, lmg = x0 @ dev.c.t(), ep @ Ca.t(), mx @ Ca.t()
    le = (lmg + dev.d.abs()) * _k(n)
    ll = _widen(lc - ldv + dev.d, le, -1)
    lu = _widen(lc + ldv + dev.d, le, 1)
    add_err = (torch.maximum(ql.abs(), qu.abs()) + torch.maximum(ll.abs(), lu.abs())) * _k(2)
    # coeff-aware McCormick lower plane
    m1 = yl * torch.where(yl >= 0, xl, xu) + xl * torch.where(xl >= 0, yl, yu) + (-xl) * yl
    m2 = yu * torch.where(yu >= 0, xl, xu) + xu * torch.where(xu >= 0, yl, yu) + (-xu) * yu
    r = (m1 >= m2).to(L.dtype)
    al = (yl - yu) * r + yu
    be = (xl - xu) * r + xu
    ga = (yu * xu - yl * xl) * r - yu * xu
    acc = torch.zeros(B, S, dtype=L.dtype, device=L.device)
    acc = _add(acc, -add_err)
    acc = _add(acc, 0.0)
    acc = _add(acc, dev.d - (_h(S) * lmg + _h(S) * dev.d.abs()))
    acc = _add(acc, ga.sum(-1))
    Acm = dev.c.unsqueeze(0) + al
    acc = _add(acc, -_h(1) * (Acm.abs() * mx[:, None, :]).sum(-1))
    acc = _add(acc, 0.0)
    A = Acm + torch.einsum("bsj,sji->bsi", be, dev.Q)
    own = (be.abs() * (_h(S * n) * mg)).sum(-1)
    acc = _add(acc, -(own + _h(1) * (A.abs() * mx[:, None, :]).sum(-1)))
    bias = _widen(acc, acc.abs() * _k(n), -1)
    base = (A * x0[:, None, :]).sum(-1)
    et = (A.abs() * ep[:, None, :]).sum(-1)
    ma = (A.abs() * mx[:, None, :]).sum(-1)
    root = _widen(base - et, ma * _k(n), -1)
    root = torch.where(torch.isfinite(root), root, -torch.inf)
    lb = _add(bias, root)
    if not return_A:
        return lb
    R = S if plane_rows is None else int(plane_rows)
    return lb, A[:, :R].contiguous(), acc[:, :R].contiguous()
