"""Constrained concretization kernel against the oracle, bit for bit."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _case(rng, B, H, K, X, special=False):
    xL = -rng.random((B, X)) - 0.05
    xU = rng.random((B, X)) + 0.05
    x0 = (xL + xU) / 2
    A = rng.standard_normal((B, K, X))
    b = -(A @ x0[..., None])[..., 0] + 0.4 * rng.standard_normal((B, K))
    obj = rng.standard_normal((B, H, X))
    if special and K:
        A[::7, 0, 0] = np.nan                       # a non-finite cut
        A[1::5, -1] = 0.0                           # an all-zero cut row
        b[2::9, 0] = 1e3                            # misses the box: infeasible
        A[3::4, :, X // 2] = 0.0                    # zero coefficients: q = +-inf / nan
        obj[4::6, :, 0] = 0.0
        xU[5::8, 0] = xL[5::8, 0]                   # a flat coordinate
        b[6::10] = -1e3                             # every cut redundant
    return obj, A, b, xL, xU


@pytest.mark.parametrize("B,H,K,X,max_rows,special", [
    (300, 3, 1, 9, 16, False), (300, 4, 6, 12, 16, True), (257, 21, 21, 50, 16, True),
    (64, 5, 21, 50, 0, True), (100, 2, 40, 64, 16, True), (50, 1, 3, 1, 0, False),
    (80, 3, 0, 7, 16, False), (128, 7, 12, 33, 4, True)])
def test_matches_oracle(cuda, B, H, K, X, max_rows, special):
    from gear_amd.concretize import constrained_concretize
    print(f"case B={B} H={H} K={K} X={X} cap={max_rows}", flush=True)
    rng = np.random.default_rng(B + K + X)
    obj, A, b, xL, xU = _case(rng, B, H, K, X, special)
    r0, f0, k0 = O.constrained_concretize(obj, A, b, xL, xU, max_rows=max_rows)
    d = lambda a: torch.as_tensor(a, device=cuda)
    r, f, k = constrained_concretize(d(obj), d(A), d(b), d(xL), d(xU), max_rows=max_rows)
    torch.cuda.synchronize()
    assert np.array_equal(f.cpu().numpy(), f0)
    assert np.array_equal(k.cpu().numpy(), k0)
    got = r.cpu().numpy()
    assert np.array_equal(got.view(np.int64), r0.view(np.int64)), \
        f"max diff {np.nanmax(np.abs(got - r0))}"
    if special:
        assert (f0 & 1).any() and (f0 & 2).any() and (k0 > 0).any()
