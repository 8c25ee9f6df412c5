"""dsht_inv_noalloc_d (the inverse transform with a static device scratch, no allocation)
against dsht_inv_d, through the C ABI: two calls in a row, the second on a smaller input, in
a fresh plan and after other work; multi-chunk calls; the refused cases."""
import ctypes

import numpy as np
import pytest

from conftest import red_spectrum, rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-13


def _call(p, kind, wt, xy, noalloc):
    import torch
    import dsht
    lib = dsht.load()
    f = lib.dsht_inv_noalloc_d if noalloc else lib.dsht_inv_d
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = f(p.h, kind, dsht._tptr(wt), wt.stride(0), 1, dsht._tptr(xy), xy.stride(0), xy.shape[0], None, s)
    return rc


@pytest.mark.parametrize("N,I,J,ncols", [(42, 128, 64, [40, 7]), (170, 512, 256, [200, 33])])
def test_noalloc_matches_inv_d(gpu, N, I, J, ncols):
    import torch
    import dsht
    p = dsht.Plan(N, I, J, radius=6.371e6)
    for call, nc in enumerate(ncols):  # the second call is smaller than the first
        w = torch.tensor(red_spectrum(N, nc, 500 + 10 * N + call).T.copy(), device=gpu)
        for kind in (dsht.VAL, dsht.GRADLON, dsht.GRADLAT, dsht.GRADMU):
            ref = torch.full((nc, J, I), np.nan, dtype=torch.float64, device=gpu)
            got = torch.full((nc, J, I), np.nan, dtype=torch.float64, device=gpu)
            assert _call(p, kind, w, got, True) == 0, dsht.load().dsht_last_error()
            assert _call(p, kind, w, ref, False) == 0, dsht.load().dsht_last_error()
            torch.cuda.synchronize()
            err = rel_l2(got.cpu().numpy(), ref.cpu().numpy())
            print(f"N={N} call={call} ncols={nc} kind={kind} rel_l2={err:.2e}", flush=True)
            assert err < TOL, (N, call, nc, kind, err)


def test_noalloc_against_oracle_and_refusals(gpu):
    import torch
    import dsht
    import oracle
    N, I, J, nc = 21, 64, 32, 5
    p = dsht.Plan(N, I, J, radius=6.371e6)
    o = oracle.Oracle(N, I, J, 6.371e6)
    w = red_spectrum(N, nc, 9)
    wt = torch.tensor(w.T.copy(), device=gpu)
    xy = torch.empty((nc, J, I), dtype=torch.float64, device=gpu)
    p.inv_noalloc_d(dsht.GRADLAT, wt, xy)
    torch.cuda.synchronize()
    assert rel_l2(xy.cpu().numpy(), o.inv(w, dsht.GRADLAT)) < 1e-12
    p.set_precision(dsht.PREC_BF16X6)
    with pytest.raises(dsht.DshtError, match="FP64 form only"):
        p.inv_noalloc_d(dsht.VAL, wt, xy)
    p.set_precision(dsht.PREC_FP64)
    with pytest.raises(dsht.DshtError, match="not an inverse kind"):
        p.inv_noalloc_d(dsht.DIVMU, wt, xy)
    # a plan whose 16-column Fourier block (22 * 12288 * 16 * 16 bytes = 69 MB) exceeds the
    # 64 MiB scratch is refused before any launch
    big = dsht.Plan(21, 64, 12288, radius=6.371e6)
    xb = torch.empty((1, 12288, 64), dtype=torch.float64, device=gpu)
    with pytest.raises(dsht.DshtError, match="cannot hold 16 columns"):
        big.inv_noalloc_d(dsht.VAL, wt[:, :1].contiguous(), xb)
