"""Large-n FP32-sound QCQP bound (fused FP32 MFMA kernels + FP64 relaxation) vs
the FP64 oracle: never above it (zero violations), valid at sampled points,
and looseness reported."""
import numpy as np
import pytest
import torch

from gear_amd.problems import qcqp_n500_m300_problem, qcqp_dense, random_bisection_boxes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _dev(a, cuda):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=cuda)


def _run(p, L, U, cuda, **kw):
    from gear_amd.crown import BoundedQCQP
    net = BoundedQCQP(p, cuda, precision="fp32s", large=True)
    for k, v in kw.items():
        setattr(net, k, v)
    lb = net.bound(_dev(L, cuda), _dev(U, cuda))
    torch.cuda.synchronize()
    return lb.cpu().numpy(), net.last_flag.cpu().numpy()


def _compare(lb32, lb64, what, loose_max):
    assert np.all(np.isfinite(lb32)), what
    viol = int(np.sum(lb32 > lb64))
    loose = (lb64 - lb32) / (np.abs(lb64) + 1.0)
    print(f"{what}: {lb32.size} bounds, violations {viol}, looseness max {loose.max():.3e} "
          f"mean {loose.mean():.3e}", flush=True)
    assert viol == 0, what
    assert loose.max() < loose_max, what


@pytest.mark.parametrize("n,m,B", [(100, 10, 300), (20, 5, 500), (257, 3, 64)])
def test_large_below_fp64_oracle(cuda, n, m, B):
    print(f"case n={n} m={m} B={B}", flush=True)
    p = qcqp_dense(n, m, seed=n)
    L, U = random_bisection_boxes(p.xl, p.xu, B, seed=B)
    lb32, flag = _run(p, L, U, cuda)
    lb64 = O.qcqp_crown(p.Q, p.c, p.d, L, U, threads=8)
    _compare(lb32, lb64, f"n={n} m={m}", 1e-2)
    print(f"near-tie boxes settled in FP64: {int(flag.sum())}", flush=True)


def test_qcqp_n500_sample_below_fp64_oracle(cuda):
    """n = 500, m = 300: a 16-box sample against the oracle."""
    print("case qcqp n=500 m=300 B=16", flush=True)
    p = qcqp_n500_m300_problem()
    L, U = random_bisection_boxes(p.xl, p.xu, 16, seed=7)
    lb32, _ = _run(p, L, U, cuda)
    lb64 = O.qcqp_crown(p.Q, p.c, p.d, L, U, threads=16)
    _compare(lb32, lb64, "qcqp_n500_m300", 5e-2)


def test_large_sampled_soundness(cuda):
    p = qcqp_dense(80, 6, seed=2)
    L, U = random_bisection_boxes(p.xl, p.xu, 40, seed=3)
    lb32, _ = _run(p, L, U, cuda)
    rng = np.random.default_rng(0)
    for b in range(len(L)):
        X = L[b] + rng.random((200, p.n)) * (U[b] - L[b])
        F = p.values(X)
        assert np.all(F >= lb32[b] - 1e-9 * (1 + np.abs(F))), b


def test_chunked_workspace_and_degenerate_boxes(cuda):
    p = qcqp_dense(90, 4, seed=9)
    L, U = random_bisection_boxes(p.xl, p.xu, 50, seed=1)
    L[:3] = U[:3]
    from gear_amd import _lib
    per_box = int(_lib.lib().gh_qcqp_large_workspace_bytes(p.n, p.S, 1))
    lb_small, _ = _run(p, L, U, cuda, LARGE_WS_BYTES=7 * per_box)     # 7-box chunks
    lb_full, _ = _run(p, L, U, cuda)
    lb64 = O.qcqp_crown(p.Q, p.c, p.d, L, U, threads=8)
    _compare(lb_small, lb64, "chunked", 1e-2)
    _compare(lb_full, lb64, "one chunk", 1e-2)
    # a box's bound does not depend on the chunk it is bound in
    assert np.array_equal(lb_small, lb_full)


def test_empty_batch(cuda):
    p = qcqp_dense(70, 2, seed=1)
    lb32, _ = _run(p, np.zeros((0, 70)), np.zeros((0, 70)), cuda)
    assert lb32.shape == (0, 3)
