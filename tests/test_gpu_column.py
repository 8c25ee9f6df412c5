"""Vertical column diagnostics on the GPU (dsht_col_omega_d, dsht_col_hydpres_d,
dsht_col_uvbarot_d, dsht_col_quick_d) against the FP64 oracle: HBEDiagnose_OMG, _HydPres and
_UVBarot (model/dogcm/src/dynamics/hspm_vfvm/HBEDiagnose_hspm_vfvm_mod.f90:90-143, 290-325,
221-255) and calc_ADVFlxZ_QUICK (HBEBaroc_hspm_vfvm_mod.f90:233-317).

Both sides evaluate the reference loops in the reference's order with no fused
multiply-adds and correctly rounded division, so the results must agree bit for bit.
Cases: the T42L20, T85L60 and T170L60 grids; layer counts that do and do not fill the
kernels' 8-layer load groups (1, 7, 9); padded column pitches on input and output; the
argument errors of the C entry points.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GRAV = 9.8


def _inputs(km, J, I, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.2, 1.0, km)
    return {
        "cdk": c / c.sum(),
        "div": rng.standard_normal((km, J, I)) * 1e-6,
        "h": 1.0 + 0.05 * rng.standard_normal((km, J, I)),
        "dens": rng.standard_normal((km, J, I)) * 0.5,
        "u": rng.standard_normal((km, J, I)) * 0.1,
        "v": rng.standard_normal((km, J, I)) * 0.1,
    }


def _dev(a, gpu, pad=0):
    """float64 device copy; with pad > 0 the layers sit at a pitch of I*J + pad doubles."""
    import torch
    a = np.ascontiguousarray(a)
    if pad == 0:
        return torch.tensor(a, device=gpu)
    km, J, I = a.shape
    base = torch.full((km, J * I + pad), np.nan, dtype=torch.float64, device=gpu)
    base[:, :J * I] = torch.tensor(a.reshape(km, -1), device=gpu)
    return base[:, :J * I].view(km, J, I)


def _padded_out(nlev, J, I, pad, gpu):
    import torch
    base = torch.full((nlev, J * I + pad), np.nan, dtype=torch.float64, device=gpu)
    return base[:, :J * I].view(nlev, J, I)


@pytest.mark.parametrize("N,I,J,km", [(42, 128, 64, 20), (85, 256, 128, 60), (170, 512, 256, 60),
                                      (21, 64, 32, 1), (21, 64, 32, 7), (42, 128, 64, 9)])
def test_column_diagnostics_bitwise(gpu, N, I, J, km):
    import torch
    import dsht
    import oracle
    p = dsht.Plan(N, I, J, radius=6.371e6)
    x = _inputs(km, J, I, seed=N + km)
    d = {k: _dev(v, gpu) for k, v in x.items()}
    omg = p.col_omega_d(d["div"], d["h"], d["cdk"])
    hp = p.col_hydpres_d(d["dens"], d["h"], d["cdk"], GRAV)
    ub, vb = p.col_uvbarot_d(d["u"], d["v"], d["h"], d["cdk"])
    torch.cuda.synchronize()
    r_omg = oracle.col_omega(x["div"], x["h"], x["cdk"])
    r_hp = oracle.col_hydpres(x["dens"], x["h"], x["cdk"], GRAV)
    r_ub, r_vb = oracle.col_uvbarot(x["u"], x["v"], x["h"], x["cdk"])
    for name, g, r in (("omega", omg, r_omg), ("hydpres", hp, r_hp), ("ubarot", ub, r_ub), ("vbarot", vb, r_vb)):
        g = g.cpu().numpy()
        assert g.shape == r.shape, name
        nbad = int((g != r).sum())
        assert nbad == 0, (name, nbad, float(np.abs(g - r).max()))


def test_column_padded_pitch(gpu):
    """Inputs and outputs at padded layer pitches; the padding is never written."""
    import torch
    import dsht
    import oracle
    N, I, J, km = 42, 128, 64, 20
    p = dsht.Plan(N, I, J, radius=6.371e6)
    x = _inputs(km, J, I, seed=5)
    div, h, dens = _dev(x["div"], gpu, 96), _dev(x["h"], gpu, 96), _dev(x["dens"], gpu, 96)
    cdk = torch.tensor(x["cdk"], device=gpu)
    omg = _padded_out(km + 1, J, I, 40, gpu)
    hp = _padded_out(km, J, I, 40, gpu)
    p.col_omega_d(div, h, cdk, out=omg)
    p.col_hydpres_d(dens, h, cdk, GRAV, out=hp)
    torch.cuda.synchronize()
    assert np.array_equal(omg.cpu().numpy(), oracle.col_omega(x["div"], x["h"], x["cdk"]))
    assert np.array_equal(hp.cpu().numpy(), oracle.col_hydpres(x["dens"], x["h"], x["cdk"], GRAV))
    # the pitch padding after each level is untouched
    pad = omg.as_strided((km + 1, 40), (omg.stride(0), 1), J * I).cpu().numpy()
    assert np.isnan(pad).all()


def test_column_one_output_and_errors(gpu):
    import torch
    import dsht
    import oracle
    N, I, J, km = 21, 64, 32, 5
    p = dsht.Plan(N, I, J, radius=6.371e6)
    x = _inputs(km, J, I, seed=9)
    d = {k: _dev(v, gpu) for k, v in x.items()}
    lib = dsht.load()
    tp = dsht._tptr
    vb = torch.full((J, I), np.nan, dtype=torch.float64, device=gpu)
    rc = lib.dsht_col_uvbarot_d(p.h, tp(d["u"]), tp(d["v"]), tp(d["h"]), I * J, km, tp(d["cdk"]), None,
                                tp(vb), None)
    assert rc == 0
    torch.cuda.synchronize()
    assert np.array_equal(vb.cpu().numpy(), oracle.col_uvbarot(x["u"], x["v"], x["h"], x["cdk"])[1])
    # argument errors are reported, nothing is launched
    out = torch.empty((km + 1, J, I), dtype=torch.float64, device=gpu)
    cases = [
        lib.dsht_col_uvbarot_d(p.h, tp(d["u"]), tp(d["v"]), tp(d["h"]), I * J, km, tp(d["cdk"]), None, None, None),
        lib.dsht_col_omega_d(p.h, tp(d["div"]), tp(d["h"]), I * J, 0, tp(d["cdk"]), tp(out), I * J, None),
        lib.dsht_col_omega_d(p.h, tp(d["div"]), tp(d["h"]), I * J, km, None, tp(out), I * J, None),
        lib.dsht_col_omega_d(p.h, tp(d["div"]), tp(d["h"]), I * J - 1, km, tp(d["cdk"]), tp(out), I * J, None),
        lib.dsht_col_hydpres_d(p.h, tp(d["dens"]), None, I * J, km, tp(d["cdk"]), ctypes.c_double(GRAV),
                               tp(out), I * J, None),
        lib.dsht_col_hydpres_d(p.h, tp(d["dens"]), tp(d["h"]), I * J, km, tp(d["cdk"]), ctypes.c_double(GRAV),
                               tp(out), I * J - 2, None),
    ]
    assert all(rc != 0 for rc in cases), cases
    with pytest.raises(ValueError):
        p.col_omega_d(d["div"], d["h"][:, :, :I // 2], d["cdk"])
    with pytest.raises(TypeError):
        p.col_omega_d(d["div"].float(), d["h"].float(), d["cdk"])


@pytest.mark.parametrize("N,I,J,km", [(42, 128, 64, 20), (85, 256, 128, 60), (170, 512, 256, 60),
                                      (21, 64, 32, 2), (21, 64, 32, 3), (21, 64, 32, 4), (21, 64, 32, 5)])
@pytest.mark.parametrize("linfree", [False, True])
def test_quick_flux_bitwise(gpu, N, I, J, km, linfree):
    """The QUICK flux of the omega the omega kernel wrote, bit for bit against the oracle."""
    import torch
    import dsht
    import oracle
    p = dsht.Plan(N, I, J, radius=6.371e6)
    x = _inputs(km, J, I, seed=3 * N + km)
    rng = np.random.default_rng(km)
    rfdk = rng.uniform(0.5 * km, 2.0 * km, km + 1)
    trc = 10.0 + rng.standard_normal((km, J, I))
    d = {k: _dev(v, gpu) for k, v in x.items()}
    omg = p.col_omega_d(d["div"], d["h"], d["cdk"])
    flx = p.col_quick_d(omg, _dev(trc, gpu), d["cdk"], _dev(rfdk, gpu), linfree=linfree)
    torch.cuda.synchronize()
    r_omg = oracle.col_omega(x["div"], x["h"], x["cdk"])
    ref = oracle.col_quick(r_omg, trc, x["cdk"], rfdk, linfree=linfree)
    got = flx.cpu().numpy()
    nbad = int((got != ref).sum())
    assert nbad == 0, (km, linfree, nbad, float(np.abs(got - ref).max()))


def test_quick_padded_and_errors(gpu):
    import torch
    import dsht
    import oracle
    N, I, J, km = 42, 128, 64, 20
    p = dsht.Plan(N, I, J, radius=6.371e6)
    rng = np.random.default_rng(1)
    omg = rng.standard_normal((km + 1, J, I))
    trc = rng.standard_normal((km, J, I))
    c = rng.uniform(0.2, 1.0, km)
    cdk, rfdk = c / c.sum(), rng.uniform(1.0, 40.0, km + 1)
    out = _padded_out(km + 1, J, I, 24, gpu)
    dc, dr = _dev(cdk, gpu), _dev(rfdk, gpu)
    p.col_quick_d(_dev(omg, gpu, 8), _dev(trc, gpu, 56), dc, dr, out=out)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), oracle.col_quick(omg, trc, cdk, rfdk))
    pad = out.as_strided((km + 1, 24), (out.stride(0), 1), J * I).cpu().numpy()
    assert np.isnan(pad).all()
    lib = dsht.load()
    tp = dsht._tptr
    dom, dtr, dfl = _dev(omg, gpu), _dev(trc, gpu), torch.empty((km + 1, J, I), dtype=torch.float64, device=gpu)
    cases = [
        lib.dsht_col_quick_d(p.h, tp(dom), I * J, tp(dtr), I * J, 1, tp(dc), tp(dr), 0, tp(dfl), I * J, None),
        lib.dsht_col_quick_d(p.h, tp(dom), I * J, tp(dtr), I * J, km, tp(dc), None, 0, tp(dfl), I * J, None),
        lib.dsht_col_quick_d(p.h, tp(dom), I * J - 1, tp(dtr), I * J, km, tp(dc), tp(dr), 0, tp(dfl), I * J, None),
        lib.dsht_col_quick_d(p.h, None, I * J, tp(dtr), I * J, km, tp(dc), tp(dr), 0, tp(dfl), I * J, None),
    ]
    assert all(rc != 0 for rc in cases), cases
    with pytest.raises(ValueError):
        p.col_quick_d(dom[:km], dtr, dc, dr)


def _vdiff_case(km, J, I, seed, pivot):
    """Physical inputs; with pivot, thin layers under thick ones and a large diffusivity make
    DGTSV interchange rows."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.2, 1.0, km)
    if pivot:
        c[1::2] *= 1e-3
    cdk = c / c.sum()
    fdk = np.concatenate([[0.5 * cdk[0]], 0.5 * (cdk[:-1] + cdk[1:]), [0.5 * cdk[-1]]])
    kv = rng.uniform(1e-4, 1e-2, (km, J, I)) * (30.0 if pivot else 1.0)
    e3 = 4000.0 * (1.0 + 0.05 * rng.standard_normal((km, J, I)))
    q0 = 10.0 + rng.standard_normal((km, J, I))
    rhs = rng.standard_normal((km, J, I)) * 1e-5
    return q0, rhs, kv, e3, cdk, fdk


@pytest.mark.parametrize("N,I,J,km", [(42, 128, 64, 20), (85, 256, 128, 60), (170, 512, 256, 60),
                                      (21, 64, 32, 2), (21, 64, 32, 3), (21, 64, 32, 4), (21, 64, 32, 33),
                                      (21, 64, 32, 34), (21, 64, 32, 36), (21, 64, 32, 40), (21, 64, 32, 44),
                                      (21, 64, 32, 64), (21, 64, 32, 70), (21, 64, 32, 100)])
@pytest.mark.parametrize("pivot", [False, True])
def test_vdiff_bitwise(gpu, N, I, J, km, pivot):
    """calc_VDiffTRCEq with DGTSV's elimination, bit for bit against the oracle, with and
    without row interchanges; up to 64 layers the interchanges are kept in a bit mask, above
    (70, 100) every row stores its second superdiagonal.  The layer counts cover where the
    finished rows wait: all in the workspace (km < 34), registers and part of the LDS rows
    only (34-43), registers, all LDS rows and the workspace (44 and up; 24 register rows
    above 64 layers)."""
    import torch
    import dsht
    import oracle
    p = dsht.Plan(N, I, J, radius=6.371e6)
    q0, rhs, kv, e3, cdk, fdk = _vdiff_case(km, J, I, N + km, pivot)
    dt, alpha = 3600.0, 1.0
    out = p.col_vdiff_d(_dev(q0, gpu), _dev(rhs, gpu), _dev(kv, gpu), _dev(e3, gpu), _dev(cdk, gpu), _dev(fdk, gpu),
                        dt, alpha, "D", "N")
    torch.cuda.synchronize()
    ref, swaps = oracle.col_vdiff(q0, rhs, kv, e3, cdk, fdk, dt, alpha, "D", "N")
    if pivot and km >= 20:  # enough thin layers for row interchanges (short columns may have none)
        assert swaps > 0
    got = out.cpu().numpy()
    nbad = int((got != ref).sum())
    assert nbad == 0, (km, pivot, nbad, float(np.abs(got - ref).max()))


@pytest.mark.parametrize("bcs", [("N", "N"), ("D", "D"), ("N", "D")])
def test_vdiff_boundaries_pitch_and_errors(gpu, bcs):
    import torch
    import dsht
    import oracle
    N, I, J, km = 21, 64, 32, 9
    p = dsht.Plan(N, I, J, radius=6.371e6)
    q0, rhs, kv, e3, cdk, fdk = _vdiff_case(km, J, I, 77, True)
    dq0, drhs, dkv, de3 = (_dev(a, gpu, 16) for a in (q0, rhs, kv, e3))
    dc, df = _dev(cdk, gpu), _dev(fdk, gpu)
    out = _padded_out(km, J, I, 32, gpu)
    work = torch.full((3 * km * J * I + 5,), np.nan, dtype=torch.float64, device=gpu)
    p.col_vdiff_d(dq0, drhs, dkv, de3, dc, df, 600.0, 0.5, *bcs, out=out, work=work)
    torch.cuda.synchronize()
    ref, _ = oracle.col_vdiff(q0, rhs, kv, e3, cdk, fdk, 600.0, 0.5, *bcs)
    assert np.array_equal(out.cpu().numpy(), ref)
    pad = out.as_strided((km, 32), (out.stride(0), 1), J * I).cpu().numpy()
    assert np.isnan(pad).all()
    lib = dsht.load()
    tp = dsht._tptr
    d = ctypes.c_double
    small = lib.dsht_col_vdiff_d(p.h, tp(dq0), tp(drhs), tp(dkv), tp(de3), dq0.stride(0), km, tp(dc), tp(df), d(1.0),
                                 d(1.0), ord("D"), ord("N"), tp(out), out.stride(0), tp(work), 3 * km * J * I - 1,
                                 None)
    badbc = lib.dsht_col_vdiff_d(p.h, tp(dq0), tp(drhs), tp(dkv), tp(de3), dq0.stride(0), km, tp(dc), tp(df), d(1.0),
                                 d(1.0), ord("X"), ord("N"), tp(out), out.stride(0), tp(work), work.numel(), None)
    short = lib.dsht_col_vdiff_d(p.h, tp(dq0), tp(drhs), tp(dkv), tp(de3), dq0.stride(0), 1, tp(dc), tp(df), d(1.0),
                                 d(1.0), ord("D"), ord("N"), tp(out), out.stride(0), tp(work), work.numel(), None)
    assert small != 0 and badbc != 0 and short != 0


@pytest.mark.parametrize("N,I,J,km", [(85, 256, 128, 60), (21, 64, 32, 2), (21, 64, 32, 7)])
@pytest.mark.parametrize("btm", ["D", "N"])
def test_vdiff_momentum_bitwise(gpu, N, I, J, km, btm):
    """calc_VDiffMOMEq (bottom friction in the Dirichlet bottom row), bit for bit."""
    import torch
    import dsht
    import oracle
    p = dsht.Plan(N, I, J, radius=6.371e6)
    q0, rhs, kv, e3, cdk, fdk = _vdiff_case(km, J, I, 5 * N + km, True)
    bf = np.random.default_rng(km).uniform(1e-4, 1e-2, (J, I))
    out = p.col_vdiff_d(_dev(q0, gpu), _dev(rhs, gpu), _dev(kv, gpu), _dev(e3, gpu), _dev(cdk, gpu), _dev(fdk, gpu),
                        1800.0, 1.0, "D", btm, btm_frict=torch.tensor(bf, device=gpu))
    torch.cuda.synchronize()
    ref, _ = oracle.col_vdiff(q0, rhs, kv, e3, cdk, fdk, 1800.0, 1.0, "D", btm, btm_frict=bf)
    got = out.cpu().numpy()
    nbad = int((got != ref).sum())
    assert nbad == 0, (km, btm, nbad, float(np.abs(got - ref).max()))


def test_column_host_entries(gpu):
    """The host-array forms (dsht_col_omega, _hydpres, _uvbarot, _quick, _vdiff), the calling
    form of the reference's Fortran routines, bit for bit against the oracle."""
    import dsht
    import oracle
    N, I, J, km = 42, 128, 64, 20
    p = dsht.Plan(N, I, J, radius=6.371e6)
    x = _inputs(km, J, I, seed=21)
    omg = p.xyr_OMG(x["div"], x["h"], x["cdk"])
    assert np.array_equal(omg, oracle.col_omega(x["div"], x["h"], x["cdk"]))
    assert np.array_equal(p.xyz_HydPres(x["dens"], x["h"], x["cdk"], GRAV),
                          oracle.col_hydpres(x["dens"], x["h"], x["cdk"], GRAV))
    ub, vb = p.xy_UVBarot(x["u"], x["v"], x["h"], x["cdk"])
    rub, rvb = oracle.col_uvbarot(x["u"], x["v"], x["h"], x["cdk"])
    assert np.array_equal(ub, rub) and np.array_equal(vb, rvb)
    rfdk = np.random.default_rng(2).uniform(10.0, 40.0, km + 1)
    trc = x["u"] * 10.0
    for lin in (False, True):
        assert np.array_equal(p.xyr_ADVFlxZ_QUICK(omg, trc, x["cdk"], rfdk, linfree=lin),
                              oracle.col_quick(omg, trc, x["cdk"], rfdk, linfree=lin))
    q0, rhs, kv, e3, cdk, fdk = _vdiff_case(km, J, I, 9, True)
    got = p.xyz_VDiff(q0, rhs, kv, e3, cdk, fdk, 3600.0, 1.0, "D", "D")
    ref, sw = oracle.col_vdiff(q0, rhs, kv, e3, cdk, fdk, 3600.0, 1.0, "D", "D")
    assert sw > 0 and np.array_equal(got, ref)
    bf = np.full((J, I), 2e-3)
    got = p.xyz_VDiff(q0, rhs, kv, e3, cdk, fdk, 3600.0, 1.0, "D", "D", xy_BtmFrictCoef=bf)
    ref, _ = oracle.col_vdiff(q0, rhs, kv, e3, cdk, fdk, 3600.0, 1.0, "D", "D", btm_frict=bf)
    assert np.array_equal(got, ref)
    with pytest.raises(dsht.DshtError):
        p.xyr_ADVFlxZ_QUICK(omg[:2], trc[:1], x["cdk"][:1], rfdk[:2])  # km = 1


def test_column_in_place_calls_are_refused(gpu):
    """The column kernels write their outputs while their inputs are still being read, so an
    output that overlaps an input (out=q0 in the diffusion solve, omega over div, a flux over
    omega) is an argument error, not a silently wrong result.  Disjoint calls still work."""
    import torch
    import dsht
    import oracle
    N, I, J, km = 21, 64, 32, 6
    p = dsht.Plan(N, I, J, radius=6.371e6)
    q0, rhs, kv, e3, cdk, fdk = _vdiff_case(km, J, I, 3, True)
    dq0, drhs, dkv, de3 = (_dev(a, gpu) for a in (q0, rhs, kv, e3))
    dc, df = _dev(cdk, gpu), _dev(fdk, gpu)
    for bad in (dq0, drhs, dkv, de3):
        with pytest.raises(dsht.DshtError, match="overlaps"):
            p.col_vdiff_d(dq0, drhs, dkv, de3, dc, df, 600.0, 1.0, "D", "N", out=bad)
    # a workspace view that starts inside q0 (shared storage)
    big = torch.zeros(4 * km * J * I, dtype=torch.float64, device=gpu)
    qv = big[: km * J * I].view(km, J, I)
    qv.copy_(dq0)
    with pytest.raises(dsht.DshtError, match="overlaps"):
        p.col_vdiff_d(qv, drhs, dkv, de3, dc, df, 600.0, 1.0, "D", "N", work=big[km * J * I // 2:])
    out = p.col_vdiff_d(dq0, drhs, dkv, de3, dc, df, 600.0, 1.0, "D", "N")
    torch.cuda.synchronize()
    ref, _ = oracle.col_vdiff(q0, rhs, kv, e3, cdk, fdk, 600.0, 1.0, "D", "N")
    assert np.array_equal(out.cpu().numpy(), ref)
    x = _inputs(km, J, I, seed=4)
    ddiv, dh = _dev(x["div"], gpu), _dev(x["h"], gpu)
    dcdk = _dev(x["cdk"], gpu)
    alias = torch.zeros((km + 1, J, I), dtype=torch.float64, device=gpu)
    alias[:km].copy_(ddiv)
    with pytest.raises(dsht.DshtError, match="overlaps"):
        p.col_omega_d(alias[:km], dh, dcdk, out=alias)
    with pytest.raises(dsht.DshtError, match="overlaps"):
        p.col_hydpres_d(ddiv, dh, dcdk, GRAV, out=dh)
    omg = p.col_omega_d(ddiv, dh, dcdk)
    rfdk = _dev(np.full(km + 1, 12.0), gpu)
    with pytest.raises(dsht.DshtError, match="overlaps"):
        p.col_quick_d(omg, ddiv, dcdk, rfdk, out=omg)
    with pytest.raises(dsht.DshtError, match="overlaps"):
        p.col_uvbarot_d(ddiv, dh, dh, dcdk, ub=ddiv[0])
