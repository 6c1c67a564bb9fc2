"""FP32-sound form: every bound is <= the FP64 oracle bound (zero
violations), sound on sampled points, and its looseness is small."""
import numpy as np
import pytest
import torch

from gear_amd.problems import qcqp_dense, random_bisection_boxes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,m,B", [(50, 20, 8192), (8, 3, 2048), (17, 4, 1000), (33, 6, 1500),
                                   (64, 3, 1000), (1, 1, 64)])
def test_fp32s_below_fp64_oracle(cuda, n, m, B):
    from gear_amd.crown import BoundedQCQP
    print(f"case n={n} m={m} B={B}", flush=True)
    p = qcqp_dense(n, m, seed=100 + n)
    L, U = random_bisection_boxes(p.xl, p.xu, B, seed=B + n)
    ref = O.qcqp_crown(p.Q, p.c, p.d, L, U, threads=8)
    net = BoundedQCQP(p, cuda, precision="fp32s")
    lb = net.bound(torch.as_tensor(L, device=cuda), torch.as_tensor(U, device=cuda)).cpu().numpy()
    assert np.all(np.isfinite(lb))
    viol = lb > ref
    assert not viol.any(), f"{viol.sum()} bounds above the FP64 oracle"
    scale = np.abs(ref) + 1.0
    loose = (ref - lb) / scale
    flagged = int(net.last_flag.sum())
    print(f"looseness max {loose.max():.3e} mean {loose.mean():.3e}, "
          f"near-tie boxes settled in FP64 {flagged}/{B}", flush=True)
    assert loose.max() < 1e-3


def test_fp32s_sound_on_samples(cuda):
    from gear_amd.crown import BoundedQCQP
    p = qcqp_dense(12, 4, seed=7)
    L, U = random_bisection_boxes(p.xl, p.xu, 256, seed=3)
    net = BoundedQCQP(p, cuda, precision="fp32s")
    lb = net.bound(torch.as_tensor(L, device=cuda), torch.as_tensor(U, device=cuda)).cpu().numpy()
    rng = np.random.default_rng(0)
    for b in range(len(L)):
        x = L[b] + rng.random((200, 12)) * (U[b] - L[b])
        assert np.all(p.values(x) >= lb[b])


def test_fp32s_exact_ties(cuda):
    """Symmetric boxes put McCormick choices exactly on ties; they are
    settled with the FP64 choice, so the bound stays below FP64."""
    from gear_amd.crown import BoundedQCQP
    p = qcqp_dense(6, 2, seed=5)
    # symmetric boxes around 0 make x-ranges and y-ranges sign-symmetric,
    # which puts many McCormick choices exactly on a tie
    B = 512
    rng = np.random.default_rng(1)
    w = rng.random((B, 6)) + 0.1
    L, U = -w, w
    net = BoundedQCQP(p, cuda, precision="fp32s")
    tl, tu = torch.as_tensor(L, device=cuda), torch.as_tensor(U, device=cuda)
    lb = net.bound(tl, tu)
    flag = net.last_flag.bool().cpu().numpy()
    ref = O.qcqp_crown(p.Q, p.c, p.d, L, U)
    got = lb.cpu().numpy()
    assert flag.any()
    assert np.all(got <= ref)
    assert np.all((ref - got) / (np.abs(ref) + 1) < 1e-3)


def test_fp32s_planes_are_valid_lower_planes(cuda):
    """The FP32-sound planes lA x + lbias (lbias lowered by the planes'
    error) are below every row at sampled points and vertices of the box."""
    from gear_amd.crown import BoundedQCQP
    from gear_amd.problems import qcqp_dense, random_bisection_boxes
    p = qcqp_dense(24, 6, seed=3)
    L, U = random_bisection_boxes(p.xl, p.xu, 64, seed=9)
    net = BoundedQCQP(p, cuda, precision="fp32s")
    lb, lA, lbias = net.bound(torch.as_tensor(L, device=cuda), torch.as_tensor(U, device=cuda),
                              return_A=True)
    lA, lbias, lb = lA.cpu().numpy(), lbias.cpu().numpy(), lb.cpu().numpy()
    rng = np.random.default_rng(1)
    for b in range(len(L)):
        X = L[b] + rng.random((500, p.n)) * (U[b] - L[b])
        X = np.vstack([X, np.where(rng.random((100, p.n)) < 0.5, L[b], U[b])])
        F = p.values(X)
        P = X @ lA[b].T + lbias[b]
        assert np.all(P <= F + 1e-12 * (1 + np.abs(F))), b
        # the plane's box minimum is at least the concretized bound's value
        pm = np.minimum(lA[b] * L[b], lA[b] * U[b]).sum(1) + lbias[b]
        assert np.all(pm >= lb[b] - 1e-6 * (1 + np.abs(lb[b])))
    # the objective-row form writes row 0 of the same planes
    _, lA1, lbias1 = net.bound(torch.as_tensor(L, device=cuda), torch.as_tensor(U, device=cuda),
                               return_A=True, plane_rows=1)
    assert np.array_equal(lA1.cpu().numpy(), lA[:, :1])
    assert np.array_equal(lbias1.cpu().numpy(), lbias[:, :1])


def test_fp32s_staged_equals_unstaged(cuda):
    """The LDS-staged form (batches >= 16,384 boxes) computes the same bits
    as the one-wavefront form on the same boxes (bounds, planes, flags)."""
    from gear_amd.crown import BoundedQCQP
    from gear_amd.problems import qcqp_dense, random_bisection_boxes
    p = qcqp_dense(50, 20, seed=7)
    L, U = random_bisection_boxes(p.xl, p.xu, 20000, seed=4)
    tL, tU = torch.as_tensor(L, device=cuda), torch.as_tensor(U, device=cuda)
    net = BoundedQCQP(p, cuda, precision="fp32s")
    lb, lA, lbias = net.bound(tL, tU, return_A=True)
    f = net.last_flag.clone()
    parts = [net.bound(tL[a:a + 10000], tU[a:a + 10000], return_A=True) + (net.last_flag.clone(),)
             for a in (0, 10000)]
    for got, k in ((lb, 0), (lA, 1), (lbias, 2), (f, 3)):
        ref = torch.cat([parts[0][k], parts[1][k]])
        assert torch.equal(got.view(torch.int64) if got.dtype == torch.float64 else got,
                           ref.view(torch.int64) if ref.dtype == torch.float64 else ref)
