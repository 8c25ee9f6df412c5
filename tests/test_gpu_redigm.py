"""Redi/GM isopycnal mixing on the GPU (dsht_col_redigm_d, dsht_redigm_d, dsht_redigm) against
the oracle's restatement of LPhys_RediGM_hspm_vfvm_AddMixingTerm
(model/dogcm/src/physics/lphys/LPhys_RediGM_hspm_vfvm_mod.f90:218-372).

* The grid-space part, given the same gradient grids, evaluates the reference's expressions in
  its order with no fused multiply-adds on both sides; only tanh and sin (device library
  against the host's) may differ in the last bit, so the bound is 1e-13 relative to each
  output's largest magnitude, and most values agree bit for bit.  Cases: DM95 and GKW91
  interior tapering; slopes below and above SlopeMax; shallow columns whose LDD95 scans never
  stop; T42L20, T85L60 and T170L60 (on a band of rows there).
* The whole tendency (forward transforms, gradients, the column part, AlphaOptr forward and
  inverse, accumulation into the right-hand sides) at T21L8 and T42L20 in FP64, added onto
  nonzero right-hand sides: relative L2 of the tendency <= 1e-11 and max-norm <= 1e-10 of the
  largest tendency.  The FP64 transforms agree with the oracle's to ~1e-13; the slopes divide
  the horizontal gradients by vertical differences and the divergence of the fluxes cancels
  between its terms, which amplifies that to ~1e-12.  The host-array form equals the device
  form bit for bit.  BF16X6 transforms: their ~1e-7 error is relative to the field (about
  285 K), and the gradient of a field whose horizontal variation is ~100 times smaller than
  its mean carries it multiplied by that ratio and by n, so the tendency is good to ~1e-4
  (relative L2 <= 3e-4, max-norm <= 2e-3).
* At T170L60 the whole tendency conserves the tracer content of a flat-bottomed ocean.
* The call captured in a hipGraph replays to the eager result bit for bit.
* Argument errors: linear PBL tapering, km < 2, overlapping outputs, a short workspace.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
R_EARTH = 6.371e6


def _fields(p, km, seed, depth_min=3000.0, depth_max=5000.0, shallow=False, tilt=1.0):
    """Potential temperature [K], salinity, thickness factor, depth and topography of a
    stratified ocean with smooth horizontal structure and noise, on the plan's grid."""
    from lfam3 import vertical_grid
    rng = np.random.default_rng(seed)
    J, I = p.jm, p.im
    cdk, fdk, rcdk, rfdk, zck = vertical_grid(km)
    lon = p.x_Lon[None, :]
    mu = np.sin(p.y_Lat)[:, None]
    topo = depth_min + (depth_max - depth_min) * (0.5 + 0.5 * np.cos(2 * lon) * (1 - mu ** 2))
    if shallow:
        topo[:, : I // 4] = 40.0  # columns too shallow for either LDD95 scan to stop
    h = np.broadcast_to(topo, (km, J, I)).copy()
    z = h * zck[:, None, None]
    pt = 275.0 + 20.0 * np.exp(z / 800.0) + tilt * 3.0 * (mu ** 2) * np.cos(lon)[None] * np.exp(z / 1500.0)
    pt = pt + 0.02 * rng.standard_normal(pt.shape)
    sa = 34.5 + 0.5 * np.exp(z / 1000.0) + 0.3 * mu * np.sin(3 * lon)[None] * np.exp(z / 2000.0)
    sa = sa + 0.005 * rng.standard_normal(sa.shape)
    return dict(pt=pt, sa=sa, h=h, z=z, topo=topo, cdk=cdk, rfdk=rfdk)


def _dev(x, gpu):
    import torch
    return {k: torch.tensor(np.ascontiguousarray(v), device=gpu) for k, v in x.items()}


def _cfg_pair(**kw):
    import dsht
    import redigm as RG
    ocfg = RG.Config(**{k: v for k, v in kw.items()})
    c = dsht.redigm_config(**{("interior_taper" if k == "interior" else k): v for k, v in kw.items()})
    return ocfg, c


@pytest.mark.parametrize("N,I,J,km,interior,tilt,shallow", [
    (42, 128, 64, 20, "DM95", 1.0, False),
    (42, 128, 64, 20, "GKW91", 1.0, True),
    (42, 128, 64, 20, "DM95", 40.0, True),   # steep slopes: the interior taper is active
    (85, 256, 128, 60, "DM95", 5.0, False),
])
def test_col_redigm_matches_oracle(gpu, N, I, J, km, interior, tilt, shallow):
    import torch
    import dsht
    import redigm as RG
    p = dsht.Plan(N, I, J, radius=R_EARTH)
    x = _fields(p, km, seed=N + km, shallow=shallow, tilt=tilt)
    rng = np.random.default_rng(5)
    # gradient grids of plausible size (not the fields' own: the column part takes any)
    gl = rng.standard_normal((2 * km, J, I)) * 1e-6 * R_EARTH
    ga = rng.standard_normal((2 * km, J, I)) * 1e-6 * R_EARTH
    d = _dev(x, gpu)
    dg = _dev({"gl": gl, "ga": ga}, gpu)
    ocfg, cfg = _cfg_pair(interior=interior)
    fl, fa, vd = p.col_redigm_d(d["pt"], d["sa"], d["h"], d["z"], d["topo"], d["cdk"], d["rfdk"], dg["gl"], dg["ga"],
                                cfg=cfg)
    torch.cuda.synchronize()
    lat = p.y_Lat
    cs = np.array([math.cos(v) for v in lat])
    grads = (gl[:km], ga[:km], gl[km:], ga[km:])
    r = RG.column_part(x["pt"], x["sa"], x["h"], x["z"], x["topo"], grads, lat, cs, x["cdk"], x["rfdk"], ocfg,
                       R_EARTH)
    exact = 0
    total = 0
    for name, g, ref in (("flx_lon", fl, r[0]), ("flx_lat", fa, r[1]), ("vdiv", vd, r[2])):
        g = g.cpu().numpy().reshape(ref.shape)
        scale = np.max(np.abs(ref))
        assert scale > 0, name
        err = np.max(np.abs(g - ref)) / scale
        print(f"{name}: max err {err:.2e} of {scale:.3e}, bit-exact {np.mean(g == ref):.4f}", flush=True)
        assert err < 1e-13, name
        exact += int((g == ref).sum())
        total += g.size
    assert exact > 0.9 * total


def test_col_redigm_t170_band(gpu):
    """T170L60: the oracle on a band of 8 latitude rows of the device result."""
    import torch
    import dsht
    import redigm as RG
    N, I, J, km = 170, 512, 256, 60
    p = dsht.Plan(N, I, J, radius=R_EARTH)
    x = _fields(p, km, seed=11, tilt=5.0, shallow=True)
    rng = np.random.default_rng(6)
    gl = rng.standard_normal((2 * km, J, I)) * 1e-6 * R_EARTH
    ga = rng.standard_normal((2 * km, J, I)) * 1e-6 * R_EARTH
    d = _dev(x, gpu)
    dg = _dev({"gl": gl, "ga": ga}, gpu)
    ocfg, cfg = _cfg_pair()
    out = p.col_redigm_d(d["pt"], d["sa"], d["h"], d["z"], d["topo"], d["cdk"], d["rfdk"], dg["gl"], dg["ga"],
                         cfg=cfg)
    torch.cuda.synchronize()
    rows = slice(60, 68)
    sub = lambda a: np.ascontiguousarray(a[:, rows])
    lat = p.y_Lat[rows]
    cs = np.array([math.cos(v) for v in lat])
    grads = (sub(gl[:km]), sub(ga[:km]), sub(gl[km:]), sub(ga[km:]))
    r = RG.column_part(sub(x["pt"]), sub(x["sa"]), sub(x["h"]), sub(x["z"]), x["topo"][rows], grads, lat, cs,
                       x["cdk"], x["rfdk"], ocfg, R_EARTH)
    for name, g, ref in zip(("flx_lon", "flx_lat", "vdiv"), out, r):
        g = g.cpu().numpy()[:, rows].reshape(ref.shape)
        err = np.max(np.abs(g - ref)) / np.max(np.abs(ref))
        print(f"T170 {name}: max err {err:.2e}", flush=True)
        assert err < 1e-13, name


def _full_case(gpu, N, I, J, km, seed, prec=None):
    import torch
    import dsht
    import oracle
    import redigm as RG
    p = dsht.Plan(N, I, J, radius=R_EARTH)
    if prec is not None:
        p.set_precision(prec)
    x = _fields(p, km, seed=seed, tilt=2.0)
    rng = np.random.default_rng(seed + 1)
    rhs0 = rng.standard_normal((2, km, J, I)) * 1e-7
    d = _dev(x, gpu)
    rt = torch.tensor(rhs0[0], device=gpu)
    rs = torch.tensor(rhs0[1], device=gpu)
    ocfg, cfg = _cfg_pair()
    p.redigm_d(d["pt"], d["sa"], d["h"], d["z"], d["topo"], d["cdk"], d["rfdk"], rt, rs, cfg=cfg)
    torch.cuda.synchronize()
    got = np.stack([rt.cpu().numpy(), rs.cpu().numpy()])
    o = oracle.Oracle(N, I, J, R_EARTH)
    ref = RG.add_mixing_term(o, x["pt"], x["sa"], x["h"], x["z"], x["topo"], x["cdk"], x["rfdk"], ocfg, rhs=rhs0,
                             lat=p.y_Lat)
    return p, x, rhs0, got, ref, cfg


def test_redigm_tracers_in_one_array(gpu):
    """Temperature and salinity as the two halves of one array (the model's xyza_TRC): the
    library makes one forward call over both; the result equals the separate-array call."""
    import torch
    import dsht
    N, I, J, km = 42, 128, 64, 20
    p = dsht.Plan(N, I, J, radius=R_EARTH)
    x = _fields(p, km, seed=31, tilt=2.0)
    d = _dev(x, gpu)
    trc = torch.cat([d["pt"], d["sa"]]).contiguous()
    outs = []
    for pt, sa in ((d["pt"], d["sa"]), (trc[:km], trc[km:])):
        rt = torch.zeros((km, J, I), dtype=torch.float64, device=gpu)
        rs = torch.zeros_like(rt)
        p.redigm_d(pt, sa, d["h"], d["z"], d["topo"], d["cdk"], d["rfdk"], rt, rs)
        outs.append((rt, rs))
    torch.cuda.synchronize()
    for a, b in zip(outs[0], outs[1]):
        assert torch.equal(a, b)


@pytest.mark.parametrize("N,I,J,km", [(21, 64, 32, 8), (42, 128, 64, 20)])
def test_redigm_matches_oracle(gpu, N, I, J, km):
    from conftest import rel_l2
    p, x, rhs0, got, ref, cfg = _full_case(gpu, N, I, J, km, seed=N)
    for t, name in enumerate(("ptemp", "salt")):
        tend = ref[t] - rhs0[t]
        scale = np.max(np.abs(tend))
        err = np.max(np.abs(got[t] - ref[t])) / scale
        l2 = rel_l2(got[t] - rhs0[t], tend)
        print(f"T{N}L{km} {name}: tendency max {scale:.3e}, max-norm err {err:.2e}, rel L2 {l2:.2e}", flush=True)
        assert scale > 0
        assert l2 < 1e-11, name
        assert err < 1e-10, name
    # the host-array form equals the device form
    pr, sr = p.RediGM_AddMixingTerm(rhs0[0], rhs0[1], x["pt"], x["sa"], x["h"], x["z"], x["topo"], x["cdk"],
                                    x["rfdk"], cfg=cfg)
    np.testing.assert_array_equal(pr, got[0])
    np.testing.assert_array_equal(sr, got[1])


def test_redigm_bf16x6_transforms(gpu):
    import dsht
    N, I, J, km = 42, 128, 64, 20
    from conftest import rel_l2
    _, _, rhs0, got, ref, _ = _full_case(gpu, N, I, J, km, seed=3, prec=dsht.PREC_BF16X6)
    for t in range(2):
        tend = ref[t] - rhs0[t]
        err = np.max(np.abs(got[t] - ref[t])) / np.max(np.abs(tend))
        l2 = rel_l2(got[t] - rhs0[t], tend)
        print(f"bf16x6 tracer {t}: max-norm err {err:.2e}, rel L2 {l2:.2e}", flush=True)
        assert l2 < 3e-4
        assert err < 2e-3


def test_redigm_t170_conserves_content(gpu):
    import torch
    import dsht
    N, I, J, km = 170, 512, 256, 60
    p = dsht.Plan(N, I, J, radius=R_EARTH)
    x = _fields(p, km, seed=17, depth_min=4000.0, depth_max=4000.0, tilt=3.0)
    d = _dev(x, gpu)
    rt = torch.zeros((km, J, I), dtype=torch.float64, device=gpu)
    rs = torch.zeros_like(rt)
    p.redigm_d(d["pt"], d["sa"], d["h"], d["z"], d["topo"], d["cdk"], d["rfdk"], rt, rs)
    torch.cuda.synchronize()
    w = p.y_Lat_Weight / p.y_Lat_Weight.sum()
    for t, r in enumerate((rt, rs)):
        a = r.cpu().numpy()
        assert np.all(np.isfinite(a))
        lay = np.einsum("kji,j->k", a, w) / I
        lay_abs = np.einsum("kji,j->k", np.abs(a), w) / I
        content = float(np.dot(x["cdk"], lay))
        scale = float(np.dot(x["cdk"], lay_abs))
        print(f"T170L60 tracer {t}: content {content:.3e} of {scale:.3e}", flush=True)
        assert scale > 0
        assert abs(content) < 1e-11 * scale


def test_redigm_argument_errors(gpu):
    import torch
    import dsht
    N, I, J, km = 21, 64, 32, 4
    p = dsht.Plan(N, I, J, radius=R_EARTH)
    x = _fields(p, km, seed=1)
    d = _dev(x, gpu)
    rt = torch.zeros((km, J, I), dtype=torch.float64, device=gpu)
    rs = torch.zeros_like(rt)
    bad = dsht.redigm_config(pbl_taper=4)
    with pytest.raises(dsht.DshtError, match="LDD95"):
        p.redigm_d(d["pt"], d["sa"], d["h"], d["z"], d["topo"], d["cdk"], d["rfdk"], rt, rs, cfg=bad)
    with pytest.raises(dsht.DshtError, match="overlap"):
        p.redigm_d(d["pt"], d["sa"], d["h"], d["z"], d["topo"], d["cdk"], d["rfdk"], d["pt"], rs)
    short = torch.empty(10, dtype=torch.float64, device=gpu)
    with pytest.raises(dsht.DshtError, match="workspace"):
        p.redigm_d(d["pt"], d["sa"], d["h"], d["z"], d["topo"], d["cdk"], d["rfdk"], rt, rs, work=short)
    one = {k: (v[:1] if k in ("pt", "sa", "h", "z") else v) for k, v in d.items()}
    one["cdk"] = d["cdk"][:1].contiguous()
    one["rfdk"] = d["rfdk"][:2].contiguous()
    with pytest.raises(ValueError, match="km >= 2"):
        p.redigm_d(one["pt"], one["sa"], one["h"], one["z"], one["topo"], one["cdk"], one["rfdk"], rt[:1], rs[:1])
    torch.cuda.synchronize()


def test_redigm_graph_replays(gpu):
    """dsht_redigm_d captured in a hipGraph after one eager call on the capture stream (the
    plan's transform workspace of that stream is then sized; the call allocates nothing else
    and records no memset): every replay, with eager work and synchronisations between the
    replays, equals the eager result bit for bit."""
    import torch
    import dsht
    N, I, J, km = 42, 128, 64, 20
    p = dsht.Plan(N, I, J, radius=R_EARTH)
    x = _fields(p, km, seed=29, tilt=2.0)
    d = _dev(x, gpu)
    s = torch.cuda.Stream()
    rt = torch.zeros((km, J, I), dtype=torch.float64, device=gpu)
    rs = torch.zeros_like(rt)
    work = torch.empty(dsht.load().dsht_redigm_work_size(p.h, km), dtype=torch.float64, device=gpu)
    args = (d["pt"], d["sa"], d["h"], d["z"], d["topo"], d["cdk"], d["rfdk"], rt, rs)
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        p.redigm_d(*args, work=work, stream=s.cuda_stream)
    s.synchronize()
    ref_t, ref_s = rt.clone(), rs.clone()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            rt.zero_()
            rs.zero_()
            p.redigm_d(*args, work=work, stream=s.cuda_stream)
    for rep in range(3):
        rt.fill_(float("nan"))
        rs.fill_(float("nan"))
        torch.cuda.synchronize()
        (d["pt"] * 2.0).sum().item()  # eager work between the replays
        with torch.cuda.stream(s):
            g.replay()
        s.synchronize()
        assert torch.equal(rt, ref_t) and torch.equal(rs, ref_s), rep
