"""FP32 grid I/O of the split precision forms (dsht_fwd_f32_d, dsht_inv_f32_d,
dsht_inv_vgrad_f32_d) against the FP64 oracle and against the same form with FP64 grids.

Bounds: each form's own bound of test_gpu_mixed.py against the oracle (the FP32 rounding of
the grid values adds about 6e-8 relative), and 2e-7 relative L2 between FP32 and FP64 grid
I/O of the same form (FP32 rounding of the outputs, or of the forward inputs)."""
import numpy as np
import pytest

from conftest import red_spectrum, rel_l2

pytestmark = pytest.mark.gpu

BOUND = {"bf16x6": 1e-6, "fp16x3": 1e-6, "bf16x3": 6e-5, "fp16": 4e-3}


def _prec(name):
    import dsht
    return {v: k for k, v in dsht.PREC_NAMES.items()}[name]


@pytest.mark.parametrize("form", list(BOUND))
@pytest.mark.parametrize("N,I,J", [(42, 128, 64), (170, 512, 256), (85, 270, 136)])
def test_f32_grids_against_oracle_and_f64_grids(gpu, form, N, I, J):
    import torch
    import dsht
    import oracle
    R = 6.371e6
    p = dsht.Plan(N, I, J, radius=R)
    o = oracle.Oracle(N, I, J, R)
    p.set_precision(_prec(form))
    nc = 24
    w = red_spectrum(N, nc, 7 + N)
    wt = torch.tensor(w.T.copy(), device=gpu)
    # inverse: every kind, and the fused value/gradient call
    f32 = [torch.full((nc, J, I), np.nan, dtype=torch.float32, device=gpu) for _ in range(3)]
    f64 = [torch.empty((nc, J, I), dtype=torch.float64, device=gpu) for _ in range(3)]
    p.inv_vgrad_f32_d(wt, *f32)
    p.inv_vgrad_d(wt, *f64)
    torch.cuda.synchronize()
    cols = [0, 11, nc - 1]
    for k, kind in enumerate(("VAL", "GRADLON", "GRADLAT")):
        got = f32[k].double().cpu().numpy()
        assert rel_l2(got, f64[k].cpu().numpy()) < 2e-7, (form, kind)
        assert rel_l2(got[cols], o.inv(w[cols], getattr(dsht, kind))) < BOUND[form], (form, kind)
    for kind in ("VAL", "GRADLAMBDA", "GRADMU", "GRADLAT"):
        k = getattr(dsht, kind)
        g = torch.empty((nc, J, I), dtype=torch.float32, device=gpu)
        p.inv_f32_d(k, wt, g)
        torch.cuda.synchronize()
        assert rel_l2(g[cols].double().cpu().numpy(), o.inv(w[cols], k)) < BOUND[form], (form, kind)
    # forward of FP32 grids: the oracle transforms the same (FP32-rounded) values in FP64
    rng = np.random.default_rng(N)
    # a large zonal mean at T42 (the recipe of test_gpu_mixed.py); zero-mean fields at T85
    # and T170, where the split forms' DIVMU of (20 + noise) cos(phi) itself exceeds 1e-6
    # with FP64 grids (1.3e-6 in bf16x6 at T170: the polar 1/(1 - mu^2) weights)
    offset = 20.0 if N <= 42 else 0.0
    xy = ((rng.standard_normal((nc, J, I)) + offset) * o.coslat[None, :, None]).astype(np.float32)
    xt = torch.tensor(xy, device=gpu)
    xd = xt.double()
    for kind in ("VAL", "DIVLAMBDA", "DIVMU"):
        k = getattr(dsht, kind)
        a = torch.zeros((p.lm, nc), dtype=torch.float64, device=gpu)
        b = torch.zeros_like(a)
        p.fwd_f32_d(k, xt, a)
        p.fwd_d(k, xd, b)
        torch.cuda.synchronize()
        an = a.cpu().numpy().T
        assert rel_l2(an, b.cpu().numpy().T) < 2e-7, (form, kind)
        assert rel_l2(an[cols], o.fwd(xy[cols].astype(np.float64), k)) < BOUND[form], (form, kind)


def test_f32_grids_refused_in_fp64(gpu):
    import torch
    import dsht
    p = dsht.Plan(21, 64, 32, radius=6.371e6)
    w = torch.zeros((p.lm, 2), dtype=torch.float64, device=gpu)
    g = torch.empty((2, 32, 64), dtype=torch.float32, device=gpu)
    with pytest.raises(dsht.DshtError, match="split precision"):
        p.inv_f32_d(dsht.VAL, w, g)
    with pytest.raises(dsht.DshtError, match="split precision"):
        p.fwd_f32_d(dsht.VAL, g, w)
