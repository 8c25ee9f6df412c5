"""Fused value + horizontal gradient inverse (dsht_inv_vgrad_d / dsht_xy_vgrad_w).

One call returns xy_w(w), xy_GradLon_w(w) and xy_GradLat_w(w) (SpmlUtil_mod.f90:414-428).
Checks:
  * against the FP64 oracle at <= 1e-12 (FP64 form) and at each split form's error bound;
  * against the three separate inverse calls of the same plan: <= 1e-13 relative L2 for
    FP64, and <= 5e-7 for the split forms (1.4e-7 measured at T170), whose fused GradLon
    multiplies the value's FP32 Fourier coefficients by i m inside the FP32 inverse FFT
    (one more FP32 rounding than the separate call, which applies i m in the FP64
    contraction epilogue);
  * every kernel family (DSHT_TUNE_LEG), NULL outputs, both column blocks and the chunked
    T170 path.
"""
import numpy as np
import pytest

from conftest import red_spectrum, rel_l2

pytestmark = pytest.mark.gpu

BOUND = {"bf16x6": 1e-6, "fp16x3": 1e-6, "bf16x3": 6e-5, "fp16": 4e-3}
SPLIT_SEP = 5e-7  # fused vs separate calls, split forms (see above)


def _prec(name):
    import dsht
    return {v: k for k, v in dsht.PREC_NAMES.items()}[name]


def _dev(plan, w, gpu):
    import torch
    nc = w.shape[0]
    wt = torch.tensor(w.T.copy(), device=gpu)  # LC layout (L, ncols)
    outs = [torch.full((nc, plan.jm, plan.im), np.nan, dtype=torch.float64, device=gpu) for _ in range(3)]
    return wt, outs


def _separate(plan, wt, like):
    import torch
    import dsht
    res = []
    for k in (dsht.VAL, dsht.GRADLON, dsht.GRADLAT):
        o = torch.empty_like(like)
        plan.inv_d(k, wt, o)
        res.append(o)
    torch.cuda.synchronize()
    return [r.cpu().numpy() for r in res]


@pytest.mark.parametrize("N,I,J", [(5, 16, 8), (21, 64, 32), (42, 128, 64), (31, 96, 48)])
@pytest.mark.parametrize("leg", [0, 2])
@pytest.mark.parametrize("cb", [16, 32])
def test_vgrad_fp64(gpu, N, I, J, leg, cb):
    import torch
    import dsht
    import oracle
    p = dsht.Plan(N, I, J, radius=6.371e6)
    o = oracle.Oracle(N, I, J, 6.371e6)
    p.set_tuning(dsht.TUNE_LEG, leg)
    p.set_tuning(dsht.TUNE_LEG_CB, cb)
    w = red_spectrum(N, 37, seed=N + 5)
    wt, outs = _dev(p, w, gpu)
    p.inv_vgrad_d(wt, *outs)
    torch.cuda.synchronize()
    got = [t.cpu().numpy() for t in outs]
    refs = [o.inv(w, k) for k in (dsht.VAL, dsht.GRADLON, dsht.GRADLAT)]
    sep = _separate(p, wt, outs[0])
    for name, g, r, s in zip(("VAL", "GRADLON", "GRADLAT"), got, refs, sep):
        assert np.isfinite(g).all(), name
        assert rel_l2(g, r) < 1e-12, (name, rel_l2(g, r))
        assert rel_l2(g, s) < 1e-13, (name, rel_l2(g, s))


@pytest.mark.parametrize("leg", [0, 1, 2, 3])
@pytest.mark.parametrize("form", list(BOUND))
@pytest.mark.parametrize("N,I,J", [(21, 64, 32), (42, 128, 64)])
def test_vgrad_split_forms(gpu, form, N, I, J, leg):
    import torch
    import dsht
    import oracle
    p = dsht.Plan(N, I, J, radius=6.371e6)
    o = oracle.Oracle(N, I, J, 6.371e6)
    p.set_precision(_prec(form))
    p.set_tuning(dsht.TUNE_LEG, leg)
    w = red_spectrum(N, 21, 3) * 1e-5
    wt, outs = _dev(p, w, gpu)
    if leg == 1:  # a retired split-form design
        with pytest.raises(dsht.DshtError, match="retired"):
            p.inv_vgrad_d(wt, *outs)
        return
    p.inv_vgrad_d(wt, *outs)
    torch.cuda.synchronize()
    got = [t.cpu().numpy() for t in outs]
    sep = _separate(p, wt, outs[0])
    for name, k, g, s in zip(("VAL", "GRADLON", "GRADLAT"), (dsht.VAL, dsht.GRADLON, dsht.GRADLAT), got, sep):
        r = o.inv(w, k)
        assert rel_l2(g, r) < BOUND[form], (form, name, rel_l2(g, r))
        assert rel_l2(g, s) < SPLIT_SEP, (form, name, rel_l2(g, s))


@pytest.mark.parametrize("prec", ["fp64", "bf16x6"])
def test_vgrad_null_outputs(gpu, prec):
    import torch
    import dsht
    N, I, J = 42, 128, 64
    p = dsht.Plan(N, I, J, radius=6.371e6)
    if prec != "fp64":
        p.set_precision(_prec(prec))
    w = red_spectrum(N, 9, 11)
    wt, outs = _dev(p, w, gpu)
    sep = _separate(p, wt, outs[0])
    for keep in ((1, 2), (0,), (2,), (0, 1)):
        o = [outs[i] if i in keep else None for i in range(3)]
        for t in outs:
            t.fill_(np.nan)
        p.inv_vgrad_d(wt, *o)
        torch.cuda.synchronize()
        for i in range(3):
            h = outs[i].cpu().numpy()
            if i in keep:
                assert rel_l2(h, sep[i]) < (1e-13 if prec == "fp64" else SPLIT_SEP), (keep, i)
            else:
                assert np.isnan(h).all(), (keep, i)  # untouched
    # every output NULL is an argument error of the C entry point
    rc = dsht.load().dsht_inv_vgrad_d(p.h, dsht._tptr(wt), wt.stride(0), 1, None, None, None, I * J, 9, None, None)
    assert rc != 0 and "NULL" in dsht.load().dsht_last_error().decode()
    with pytest.raises(ValueError):
        p.inv_vgrad_d(wt, None, None, None)


@pytest.mark.parametrize("prec", ["fp64", "bf16x6", "fp16x3"])
def test_vgrad_t170_chunked(gpu, prec):
    """T170 over several column chunks (64 columns per chunk forced)."""
    import torch
    import dsht
    N, I, J, nc = 170, 512, 256, 150
    p = dsht.Plan(N, I, J, radius=6.371e6)
    if prec != "fp64":
        p.set_precision(_prec(prec))
    p.set_tuning(dsht.TUNE_CHUNK_COLS, 64)
    w = red_spectrum(N, nc, 23)
    wt, outs = _dev(p, w, gpu)
    p.inv_vgrad_d(wt, *outs)
    torch.cuda.synchronize()
    got = [t.cpu().numpy() for t in outs]
    sep = _separate(p, wt, outs[0])
    tol = 1e-13 if prec == "fp64" else SPLIT_SEP
    for i in range(3):
        assert rel_l2(got[i], sep[i]) < tol, (prec, i, rel_l2(got[i], sep[i]))
    if prec == "fp64":
        return
    # the split form against the FP64 form of the same plan
    p.set_precision(dsht.PREC_FP64)
    ref = _separate(p, wt, outs[0])
    bound = 1e-6 if prec == "bf16x6" else 3e-6  # fp16x3's polar-latitude flush, as in the bench errors
    for i in range(3):
        assert rel_l2(got[i], ref[i]) < bound, (prec, i, rel_l2(got[i], ref[i]))


def test_vgrad_host_api(gpu):
    import dsht
    import oracle
    N, I, J = 21, 64, 32
    p = dsht.Plan(N, I, J, radius=6.371e6)
    o = oracle.Oracle(N, I, J, 6.371e6)
    w = red_spectrum(N, 4, 31)
    xy, lon, lat = p.xy_GradLonLat_w(w)
    assert rel_l2(xy, o.inv(w, dsht.VAL)) < 1e-12
    assert rel_l2(lon, o.inv(w, dsht.GRADLON)) < 1e-12
    assert rel_l2(lat, o.inv(w, dsht.GRADLAT)) < 1e-12
    none, lon1, lat1 = p.xy_GradLonLat_w(w[0], value=False)
    assert none is None and lon1.shape == (J, I)
    assert rel_l2(lon1, lon[0]) < 1e-14 and rel_l2(lat1, lat[0]) < 1e-14
