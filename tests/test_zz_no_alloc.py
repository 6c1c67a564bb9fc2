"""The allocation-free split-append (gh_split_append_f64_noalloc) against the
workspace form (gh_split_append_f64), through the C ABI: equal outputs bit
for bit on two calls in a row, the second with fewer parents (the shared
device scratch keeps the first call's contents), for the 16-lane (n, S <=
64) and the wave-per-parent kernels; K above the scratch's maximum is
refused before anything is launched."""
import numpy as np
import pytest
import torch

from gear_amd._lib import GH_EUNSUPPORTED, check, lib, ptr, stream_ptr

pytestmark = pytest.mark.gpu


def _inputs(K, n, S, R, seed, cuda):
    rng = np.random.default_rng(seed)
    d = dict(dtype=torch.float64, device=cuda)
    pL = rng.uniform(-1, 0.5, (K, n))
    pU = pL + rng.uniform(0.01, 1.5, (K, n))
    plb = rng.standard_normal((K, S))
    lA = rng.standard_normal((K, R, n)) / np.sqrt(n)
    lbias = rng.standard_normal((K, R))
    thr = np.concatenate([[0.4], rng.uniform(2.0, 3.0, S - 1)])
    return [torch.as_tensor(a, **d) for a in (pL, pU, plb, lA, lbias, thr)]


def _run(noalloc, K, n, S, R, ins, cuda):
    pL, pU, plb, lA, lbias, thr = ins
    cap = 2 * K + 8
    d = dict(dtype=torch.float64, device=cuda)
    FL = torch.full((cap + 1, n), -7.0, **d)
    FU = torch.full((cap + 1, n), -7.0, **d)
    Flb = torch.full((cap + 1, S), -7.0, **d)
    order = torch.full((2 * K,), -5, dtype=torch.int64, device=cuda)
    total = torch.zeros(1, dtype=torch.int64, device=cuda)
    disc = torch.full((1,), float("inf"), **d)
    base, dead = 3, cap
    args = [K, n, S, R, ptr(pL), ptr(pU), None, ptr(plb), ptr(lA), ptr(lbias), ptr(thr),
            ptr(FL), ptr(FU), ptr(Flb), base, cap, ptr(order), dead, ptr(total), ptr(disc)]
    if noalloc:
        check(lib().gh_split_append_f64_noalloc(*args, stream_ptr()), "noalloc")
    else:
        ws = torch.empty(int(lib().gh_split_append_workspace_bytes(K)), dtype=torch.uint8,
                         device=cuda)
        check(lib().gh_split_append_f64(*args, ptr(ws), stream_ptr()), "split_append")
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in (FL, FU, Flb, order, total, disc)]


@pytest.mark.parametrize("n,S", [(50, 21), (70, 5)])
def test_noalloc_equals_workspace_form(cuda, n, S):
    for K, seed in ((5000, 1), (1200, 2)):        # the second call is smaller
        print(f"case noalloc n={n} S={S} K={K}", flush=True)
        ins = _inputs(K, n, S, 1, seed, cuda)
        a = _run(False, K, n, S, 1, ins, cuda)
        b = _run(True, K, n, S, 1, ins, cuda)
        assert 0 < a[4][0] < 2 * K                 # some children pruned, some kept
        for x, y in zip(a, b):
            assert np.array_equal(x.view(np.uint8), y.view(np.uint8))


def test_noalloc_refuses_oversized_batch(cuda):
    kmax = int(lib().gh_split_append_noalloc_max_parents())
    assert kmax == 1 << 17
    rc = lib().gh_split_append_f64_noalloc(kmax + 1, 4, 2, 1, *([None] * 10), 0, 1, None, 0, None,
                                           None, stream_ptr())
    assert rc == GH_EUNSUPPORTED
