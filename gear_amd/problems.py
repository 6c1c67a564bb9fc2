"""Problem containers and seeded synthetic generators.

A QCQP has S = m + 1 rows,  f_s(x) = x^T Q_s x + c_s^T x + d_s,  row 0 the
objective (minimised) and rows 1..m constraints f_s(x) <= rhs_s, over the box
[xl, xu].  This is the form the fused bound kernels take; its lowering into
the reference's operator vocabulary is Linear -> Mul -> ReduceSum plus a
Linear term (the same program the reference's BoundedModule would trace
from `QCQPModule`).
"""
from dataclasses import dataclass, field

import numpy as np
import torch


@dataclass
class QCQP:
    Q: np.ndarray          # (S, n, n) float64
    c: np.ndarray          # (S, n)
    d: np.ndarray          # (S,)
    rhs: np.ndarray        # (S,), rhs[0] unused (objective)
    xl: np.ndarray         # (n,)
    xu: np.ndarray         # (n,)
    name: str = "qcqp"
    meta: dict = field(default_factory=dict)

    @property
    def n(self):
        return self.Q.shape[1]

    @property
    def S(self):
        return self.Q.shape[0]

    @property
    def m(self):
        return self.S - 1

    def values(self, x):
        """f_s(x) for x (N, n) -> (N, S), float64 numpy."""
        x = np.asarray(x, dtype=np.float64)
        quad = np.einsum("nj,sji,ni->ns", x, self.Q, x)
        return quad + x @ self.c.T + self.d

    def to_device(self, device, dtype=torch.float64):
        return DeviceQCQP(self, device, dtype)


class DeviceQCQP:
    """Problem data resident on one device (contiguous, row-major)."""

    def __init__(self, p: QCQP, device, dtype=torch.float64):
        self.host = p
        self.n, self.S = p.n, p.S
        self.device = torch.device(device)
        t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=self.device)
        self.Q, self.c, self.d = t(p.Q), t(p.c), t(p.d)
        self.rhs, self.xl, self.xu = t(p.rhs), t(p.xl), t(p.xu)


class QCQPModule(torch.nn.Module):
    """The QCQP as a torch module over x (B, n) -> (B, S): the graph the
    bound kernels relax (y = W x; z = x * y; q = sum(z); l = Cm x + d;
    f = q + l).  Used by the plain op-by-op path and by ALM-PGD."""

    def __init__(self, p: QCQP, dtype=torch.float64):
        super().__init__()
        S, n = p.S, p.n
        self.S, self.n = S, n
        self.register_buffer("W", torch.as_tensor(p.Q.reshape(S * n, n), dtype=dtype))
        self.register_buffer("Cm", torch.as_tensor(p.c, dtype=dtype))
        self.register_buffer("dv", torch.as_tensor(p.d, dtype=dtype))

    def forward(self, x):
        y = x @ self.W.t()
        z = x.unsqueeze(1) * y.view(-1, self.S, self.n)
        return z.sum(-1) + (x @ self.Cm.t() + self.dv)


def random_bisection_boxes(xl, xu, leaves, seed):
    """Seeded random bisection of the box until `leaves` leaves: repeat
    {pick a leaf uniformly, pick a dimension uniformly, bisect at
    (xL + xU) / 2}."""
    rng = np.random.default_rng(seed)
    L = [np.array(xl, dtype=np.float64)]
    U = [np.array(xu, dtype=np.float64)]
    n = len(xl)
    while len(L) < leaves:
        k = int(rng.integers(len(L)))
        i = int(rng.integers(n))
        lo, hi = L[k], U[k]
        mid = (lo[i] + hi[i]) / 2.0
        a_hi = hi.copy(); a_hi[i] = mid
        b_lo = lo.copy(); b_lo[i] = mid
        L[k], U[k] = lo, a_hi
        L.append(b_lo); U.append(hi.copy())
    return np.stack(L), np.stack(U)


def qcqp_dense(n, m, seed, name=None, box=1.0):
    """Seeded nonconvex QCQP: every coefficient (Q_s, c_s, d_s) N(0,1)/sqrt(n),
    constraints f_k(x) <= 0, x in [-box, box]^n."""
    rng = np.random.default_rng(seed)
    S = m + 1
    scale = 1.0 / np.sqrt(n)
    Q = rng.standard_normal((S, n, n)) * scale
    c = rng.standard_normal((S, n)) * scale
    d = rng.standard_normal(S) * scale
    rhs = np.zeros(S)
    xl = -box * np.ones(n)
    xu = box * np.ones(n)
    return QCQP(Q, c, d, rhs, xl, xu, name=name or f"qcqp_n{n}_m{m}_s{seed}",
                meta={"seed": seed})


def qcqp_n50_m20_problem():
    """Full batched B&B workload: n = 50, m = 20, coefficients N(0,1)/sqrt(n),
    x in [-1, 1]^50, seed 3."""
    return qcqp_dense(50, 20, seed=3, name="qcqp_n50_m20")


def qcqp_n500_m300_problem():
    """Precision workload: dense QCQP n = 500, m = 300, seed 5."""
    return qcqp_dense(500, 300, seed=5, name="qcqp_n500_m300")


def serving_sizes(count, seed=6):
    """Instance sizes of the serving stream: n = clip(round(lognormal(ln 60,
    1.0)), 5, 1000), m = 20."""
    rng = np.random.default_rng(seed)
    n = np.clip(np.round(rng.lognormal(np.log(60.0), 1.0, size=count)), 5, 1000).astype(int)
    return n


def flops_per_box(n, S):
    """Algorithmic FP64 work of one subdomain bound (all S rows), by the
    counting rule of the measurement plan: the IBP centre and radius
    products 4 S n^2, the backward product 2 S n^2, and about 40 S n
    elementwise / reduction operations.  (The kernels also form the
    magnitude product |Q| max(|l|, |u|) for the rounding enclosure, 2 S n^2
    more; it is executed work, not counted here.)"""
    return 6 * S * n * n + 40 * S * n


def bytes_per_box(n, S, with_A=False):
    """Algorithmic HBM bytes of one subdomain bound: the box in (2 n doubles),
    lb out (S doubles), plus lA (S n) and lbias (S) when returned."""
    b = 8 * (2 * n + S)
    if with_A:
        b += 8 * (S * n + S)
    return b
