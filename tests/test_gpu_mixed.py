"""Split-precision Legendre forms against the FP64 oracle.

Error bounds (relative L2 over the whole output), from the split width of each form:
  bf16x6, fp16x3   <= 1e-6   (the bar the T170 metric is quoted at)
  bf16x3           <= 6e-5   (2 x 8-bit terms: ~2^-16 per product; DIVMU's polar
                              1/(1-mu^2) weights add a factor ~2)
  fp16             <= 4e-3   (11-bit single term)
At T170 the fp16x3 forward transform reaches ~2e-6: with one power-of-two scale per
column, the quadrature-weighted values of the most polar latitudes fall below FP16's
normal range and are flushed.  bf16x6 keeps FP32's exponent range and meets 1e-6.
Every form also reports its T170 error in bench.py.
"""
import numpy as np
import pytest

from conftest import red_spectrum, rel_l2

pytestmark = pytest.mark.gpu

BOUND = {"bf16x6": 1e-6, "fp16x3": 1e-6, "bf16x3": 6e-5, "fp16": 4e-3}


def _prec(name):
    import dsht
    return {v: k for k, v in dsht.PREC_NAMES.items()}[name]


# contraction kernel families (dsht_set_tuning DSHT_TUNE_LEG): per form, retired (refused),
# packed data panels, whole-K LDS staging
@pytest.mark.parametrize("leg", [0, 1, 2, 3])
@pytest.mark.parametrize("form", list(BOUND))
@pytest.mark.parametrize("N,I,J", [(21, 64, 32), (42, 128, 64), (31, 96, 48)])
def test_mixed_transforms(gpu, form, N, I, J, leg):
    import dsht
    import oracle
    R = 6.371e6
    p = dsht.Plan(N, I, J, radius=R)
    o = oracle.Oracle(N, I, J, R)
    p.set_precision(_prec(form))
    p.set_tuning(dsht.TUNE_LEG, leg)
    tol = BOUND[form]
    w = red_spectrum(N, 6, 3) * 1e-5  # magnitudes of a vorticity field
    if leg == 1:  # a retired split-form design
        with pytest.raises(dsht.DshtError, match="retired"):
            p._inv(w, dsht.VAL)
        with pytest.raises(dsht.DshtError, match="retired"):
            p._fwd(np.ones((1, J, I)), dsht.VAL)
        return
    for kind in ("VAL", "GRADLAMBDA", "GRADMU", "GRADLAT"):
        k = getattr(dsht, kind)
        assert rel_l2(p._inv(w, k), o.inv(w, k)) < tol, (form, kind)
    rng = np.random.default_rng(4)
    xy = (rng.standard_normal((5, J, I)) + 20.0) * o.coslat[None, :, None]
    for kind in ("VAL", "DIVLAMBDA", "DIVMU"):
        k = getattr(dsht, kind)
        assert rel_l2(p._fwd(xy, k), o.fwd(xy, k)) < tol, (form, kind)
    A, B = xy[:3], xy[2:] * 1e-3
    assert rel_l2(p.w_AlphaOptr_xy(A, B), o.alpha_fwd(A, B)) < tol
    vor, div = p.calc_UVCosLat2VorDiv(A, B)
    rv, rd = o.uv2vordiv(A, B)
    assert rel_l2(vor, rv) < tol and rel_l2(div, rd) < tol
    wa, wb = red_spectrum(N, 3, 7), red_spectrum(N, 3, 8) * 1e-2
    assert rel_l2(p.xy_AlphaOptr_w(wa, wb), o.alpha_inv(wa, wb)) < tol
    u, v = p.calc_VorDiv2UV(wa * 1e-5, wb * 1e-6)
    ru, rv2 = o.vordiv2uv(wa * 1e-5, wb * 1e-6)
    assert rel_l2(u, ru) < tol and rel_l2(v, rv2) < tol


@pytest.mark.parametrize("leg", [0, 3])
@pytest.mark.parametrize("form", ["bf16x6", "fp16x3"])
def test_mixed_t170_roundtrip(gpu, form, leg):
    """T170 round trip and inverse accuracy at the metric's tolerance (1e-6)."""
    import torch
    import dsht
    N, I, J, nc = 170, 512, 256, 40
    p = dsht.Plan(N, I, J, radius=6.371e6)
    p.set_tuning(dsht.TUNE_LEG, leg)
    w = torch.tensor(red_spectrum(N, nc, 21).T.copy(), device=gpu)
    ref = torch.empty((nc, J, I), dtype=torch.float64, device=gpu)
    p.inv_d(dsht.VAL, w, ref)
    p.set_precision(_prec(form))
    got = torch.empty_like(ref)
    p.inv_d(dsht.VAL, w, got)
    w2 = torch.zeros_like(w)
    p.fwd_d(dsht.VAL, ref, w2)
    torch.cuda.synchronize()
    assert rel_l2(got.cpu().numpy(), ref.cpu().numpy()) < 1e-6
    fwd_tol = 1e-6 if form == "bf16x6" else 3e-6
    assert rel_l2(w2.cpu().numpy(), w.cpu().numpy()) < fwd_tol


@pytest.mark.parametrize("form", ["bf16x6", "fp16x3"])
def test_mixed_forward_large_zonal_mean(gpu, form):
    """A field with a large background (35 + 0.05 x unit-variance perturbation, like
    salinity): the split forms' FP32 forward FFT transforms each row relative to its first
    sample, so the m != 0 coefficients keep FP32 relative precision.  Checked against the
    FP64 form on the m != 0 coefficients (the m = 0 column carries the background through
    the FP32-accumulated contraction, and is bounded relative to the background)."""
    import torch
    import dsht
    N, I, J, nc = 170, 512, 256, 24
    p = dsht.Plan(N, I, J, radius=6.371e6)
    w = torch.tensor(red_spectrum(N, nc, 41).T.copy(), device=gpu)
    pert = torch.empty((nc, J, I), dtype=torch.float64, device=gpu)
    p.inv_d(dsht.VAL, w, pert)
    pert /= pert.std()
    xy = 35.0 + 0.05 * pert
    ref = torch.zeros_like(w)
    p.fwd_d(dsht.VAL, xy, ref)
    p.set_precision(_prec(form))
    got = torch.zeros_like(w)
    p.fwd_d(dsht.VAL, xy, got)
    torch.cuda.synchronize()
    r, q = ref.cpu().numpy(), got.cpu().numpy()
    n = np.floor(np.sqrt(np.arange(p.lm))).astype(int)
    zonal = np.arange(p.lm) == n * (n + 1)  # m = 0
    assert rel_l2(q[zonal], r[zonal]) < 1e-6
    bound = 1e-6 if form == "bf16x6" else 3e-6
    assert rel_l2(q[~zonal], r[~zonal]) < bound, rel_l2(q[~zonal], r[~zonal])


@pytest.mark.parametrize("cb", [16, 32])
@pytest.mark.parametrize("form", list(BOUND))
def test_mixed_forward_column_tiles(gpu, form, cb):
    """The LDS-shared forward with one (16) and two (32) column tiles per block, over a
    column count that leaves the last block half empty."""
    import dsht
    import oracle
    N, I, J, R = 42, 128, 64, 6.371e6
    p = dsht.Plan(N, I, J, radius=R)
    o = oracle.Oracle(N, I, J, R)
    p.set_precision(_prec(form))
    p.set_tuning(dsht.TUNE_LEG_CB, cb)
    tol = BOUND[form]
    rng = np.random.default_rng(9)
    xy = (rng.standard_normal((21, J, I)) + 20.0) * o.coslat[None, :, None]
    for kind in ("VAL", "DIVLAMBDA", "DIVMU"):
        k = getattr(dsht, kind)
        assert rel_l2(p._fwd(xy, k), o.fwd(xy, k)) < tol, (form, cb, kind)
    A, B = xy[:11], xy[10:] * 1e-3
    assert rel_l2(p.w_AlphaOptr_xy(A, B), o.alpha_fwd(A, B)) < tol
    vor, div = p.calc_UVCosLat2VorDiv(A, B)
    rv, rd = o.uv2vordiv(A, B)
    assert rel_l2(vor, rv) < tol and rel_l2(div, rd) < tol


@pytest.mark.parametrize("cb", [16, 32])
@pytest.mark.parametrize("form", list(BOUND))
def test_mixed_forward_four_k_steps(gpu, form, cb):
    """The forward at J = 256 (four 32-latitude K steps: the unrolled kernel whose one-slot
    rounds request the next step's table fragments early) with one (16) and two (32) column
    tiles per block, over 21 columns (the last block partly empty), every forward kind against
    the oracle (rows of both parities, rounds of one and of two slots per wave)."""
    import dsht
    import oracle
    N, I, J, R = 170, 512, 256, 6.371e6
    p = dsht.Plan(N, I, J, radius=R)
    o = oracle.Oracle(N, I, J, R)
    p.set_precision(_prec(form))
    p.set_tuning(dsht.TUNE_LEG_CB, cb)
    tol = {"bf16x6": 1e-6, "fp16x3": 3e-6, "bf16x3": 6e-5, "fp16": 4e-3}[form]
    rng = np.random.default_rng(23)
    xy = rng.standard_normal((21, J, I)) * o.coslat[None, :, None]
    for kind in ("VAL", "DIVLAMBDA", "DIVMU"):
        k = getattr(dsht, kind)
        assert rel_l2(p._fwd(xy, k), o.fwd(xy, k)) < tol, (form, cb, kind)


@pytest.mark.parametrize("leg", [0, 2, 3])
@pytest.mark.parametrize("cb", [16, 32])
@pytest.mark.parametrize("form", list(BOUND))
def test_mixed_inverse_column_tiles(gpu, form, cb, leg):
    """The inverse with one (16) and two (32) column tiles per block, in every family (per
    form, packed, whole-K staged), over 21 columns (the last block partly empty) at an odd
    truncation (the last wavenumber has no odd-parity row): each kind alone against the oracle,
    and the fused value/gradient pass against the separate kinds of the same form."""
    import torch
    import dsht
    import oracle
    N, I, J, R = 31, 96, 48, 6.371e6
    p = dsht.Plan(N, I, J, radius=R)
    o = oracle.Oracle(N, I, J, R)
    p.set_precision(_prec(form))
    p.set_tuning(dsht.TUNE_LEG_CB, cb)
    p.set_tuning(dsht.TUNE_LEG, leg)
    tol = BOUND[form]
    w = red_spectrum(N, 21, 17) * 1e-5
    sep = {}
    for kind in ("VAL", "GRADLON", "GRADLAT"):
        got = p._inv(w, getattr(dsht, kind))
        assert rel_l2(got, o.inv(w, getattr(dsht, kind))) < tol, (form, cb, leg, kind)
        sep[kind] = got
    wt = torch.tensor(w.T.copy(), device=gpu)
    g = [torch.empty((21, J, I), dtype=torch.float64, device=gpu) for _ in range(3)]
    p.inv_vgrad_d(wt, *g)
    torch.cuda.synchronize()
    for kind, t in zip(("VAL", "GRADLON", "GRADLAT"), g):
        assert rel_l2(t.cpu().numpy(), sep[kind].reshape(21, J, I)) < tol, (form, cb, leg, kind)


@pytest.mark.parametrize("form", list(BOUND))
def test_every_form_at_t170(gpu, form):
    """The T170L60 step's grid: every split form at T170 (I=512, J=256) on 40 columns of red-spectrum data
    with a large zonal-mean background (like temperature), against the FP64 form within its
    bound: the fused value/gradient inverse (three output groups) and the forward of grid
    products."""
    import torch
    import dsht
    N, I, J, nc = 170, 512, 256, 40
    p = dsht.Plan(N, I, J, radius=6.371e6)
    w = red_spectrum(N, nc, 61)
    w[:, 0] += 15.0
    wt = torch.tensor(w.T.copy(), device=gpu)
    ref = [torch.empty((nc, J, I), dtype=torch.float64, device=gpu) for _ in range(3)]
    p.inv_vgrad_d(wt, *ref)
    prod = (ref[0] * ref[0] * 1e-2).contiguous()
    sref = torch.zeros((p.lm, nc), dtype=torch.float64, device=gpu)
    p.fwd_d(dsht.VAL, prod, sref)
    p.set_precision(_prec(form))
    got = [torch.empty_like(ref[0]) for _ in range(3)]
    p.inv_vgrad_d(wt, *got)
    sgot = torch.zeros_like(sref)
    p.fwd_d(dsht.VAL, prod, sgot)
    torch.cuda.synchronize()
    errs = [rel_l2(a.cpu().numpy(), b.cpu().numpy()) for a, b in zip(got, ref)]
    errs.append(rel_l2(sgot.cpu().numpy(), sref.cpu().numpy()))
    print(form, ["%.2e" % e for e in errs])
    assert max(errs) < BOUND[form], (form, errs)
