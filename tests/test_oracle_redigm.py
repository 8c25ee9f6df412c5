"""Pins of the oracle's Redi/GM restatement (oracle/redigm.py) on the CPU.

The reference holds no test or golden vector for LPhys_RediGM_hspm_vfvm_mod.f90, so the
restatement is pinned by what the physics and the published polynomials fix:

* the McDougall (1987) alpha/beta polynomial of EOS_JM95_alpha_beta against its published
  check value (S = 40 psu, theta = 10 degC, p = 4000 dbar: alpha/beta = 0.34763,
  beta = 0.72088e-3);
* the isoneutral slope of a stratification T(z) with a uniform horizontal gradient a is
  -a / (dT/dz) (calc_IsoNeutralSlope_new);
* the Redi flux of the tracer that defines the density vanishes, in the layers and on the
  half levels (the Redi tensor diffuses only along isopycnals), where the tapering is 1;
* the GM skew flux of that tracer is kappa_GM a on the layers and -kappa_GM |a|^2 / (dT/dz)
  on the half levels;
* the fluxes are linear in (kappa_Redi, kappa_GM);
* the tapering functions at their fixed points (DM95 at |S| = SlopeMax is 1/2, GKW91 at
  2 SlopeMax is 1/4, LDD95 is 1 in the interior and uses the slope maximum above the first
  layer deeper than the eddy displacement only);
* the vertical flux divergence telescopes (sum_k H cdk vdiv = 0) and the whole tendency
  conserves the tracer content of a flat-bottomed ocean.
A planted mistake (a sign of the slope term, the half-level weights swapped, the taper
scan not stopping) fails one of these checks.
"""
import math
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import redigm as RG  # noqa: E402
from lfam3 import vertical_grid  # noqa: E402


def test_alpha_beta_check_value():
    cfg = RG.Config()
    rg = cfg.ref_dens * cfg.grav
    z = np.array([-4000.0 / (1e-4 * rg)])  # 4000 dbar
    al, be = RG.alpha_beta(np.array([283.15]), np.array([40.0]), z, cfg)
    assert abs((al / be)[0] - 0.34763) < 1e-5
    assert abs(be[0] - 0.72088e-3) < 1e-8


def _column_setup(km=20, depth=4000.0, gamma=0.005, a=1e-6, b=-0.6e-6, I=4, lat_deg=(30.0, -45.0)):
    """A flat-bottomed stratified ocean with uniform horizontal temperature gradient (a, b)
    [K/m] and uniform salinity: T = 285 + gamma Z on every column."""
    cdk, fdk, rcdk, rfdk, zck = vertical_grid(km)
    J = len(lat_deg)
    lat = np.radians(np.array(lat_deg))
    h = np.full((km, J, I), depth)
    z = depth * zck[:, None, None] * np.ones((km, J, I))
    topo = np.full((J, I), depth)
    pt = 285.0 + gamma * z
    sa = np.full((km, J, I), 35.0)
    R = 6.371e6
    grads = (np.full((km, J, I), a * R), np.full((km, J, I), b * R), np.zeros((km, J, I)), np.zeros((km, J, I)))
    coslat = np.array([math.cos(v) for v in lat])
    return dict(pt=pt, sa=sa, h=h, z=z, topo=topo, grads=grads, lat=lat, coslat=coslat, cdk=cdk, rfdk=rfdk, R=R)


def _run(s, cfg):
    return RG.column_part(s["pt"], s["sa"], s["h"], s["z"], s["topo"], s["grads"], s["lat"], s["coslat"], s["cdk"],
                          s["rfdk"], cfg, s["R"])


def _interior(s, cfg):
    """Layers where the LDD95 taper is exactly 1 on every column."""
    km = s["pt"].shape[0]
    hg_t = (s["grads"][0] / s["R"], s["grads"][1] / s["R"])
    hg_s = (s["grads"][2] / s["R"], s["grads"][3] / s["R"])
    ds_t, ds_s = RG.dsig(s["pt"], s["rfdk"]), RG.dsig(s["sa"], s["rfdk"])
    slon, slat = RG.slopes(s["pt"], s["sa"], s["z"], s["h"], hg_t, hg_s, ds_t, ds_s, cfg)
    tap = RG.taper(slon, slat, s["z"], s["topo"], s["lat"], cfg)
    return [k for k in range(km) if np.all(tap[k] == 1.0)], slon, slat, tap


def test_slope_of_a_tilted_stratification():
    s = _column_setup()
    cfg = RG.Config()
    _, slon, slat, _ = _interior(s, cfg)
    # -a / (dT/dz) up to the EPS of the denominator (relative 1e-10 / (alpha gamma H) ~ 1e-7)
    np.testing.assert_allclose(slon, -1e-6 / 0.005, rtol=1e-6)
    np.testing.assert_allclose(slat, 0.6e-6 / 0.005, rtol=1e-6)


@pytest.mark.parametrize("interior", ["GKW91", "DM95"])
def test_redi_flux_of_the_density_tracer_vanishes(interior):
    s = _column_setup()
    cfg = RG.Config(kappa_gm=0.0, interior=interior)
    ks, _, _, tap = _interior(s, cfg)
    fa, fb, vt = _run(s, cfg)
    cs = s["coslat"][None, :, None]
    scale = cfg.kappa_redi * 1e-6
    if interior == "GKW91":
        assert len(ks) >= 10  # the taper is exactly 1 below the surface and above the bottom layers
        for k in ks:
            assert np.max(np.abs(fa[0, k] / cs)) < 1e-6 * scale
            assert np.max(np.abs(fb[0, k] / cs)) < 1e-6 * scale
        # half levels between two untapered layers carry no flux: vdiv of the layers
        # whose both half levels lie there is zero
        inner = [k for k in ks if k - 1 in ks and k + 1 in ks]
        for k in inner:
            assert np.max(np.abs(vt[0, k])) < 1e-9 * scale / 4000.0
    else:
        # DM95 leaves 1 - tap of the flux: FLon = kappa a (1 - tap), up to the EPS term of the
        # slope denominator (relative ~1e-7)
        np.testing.assert_allclose(fa[0] / cs, cfg.kappa_redi * 1e-6 * (1.0 - tap), rtol=1e-5, atol=1e-6 * scale)
    # the salinity is uniform: its fluxes are zero
    assert np.max(np.abs(fa[1])) == 0.0 and np.max(np.abs(vt[1])) == 0.0


def _gm_face_error(fluxes=None):
    """Largest relative error of the GM vertical flux divergence against its closed form for a
    horizontal gradient that varies linearly with depth, a(z) = a0 + a1 z: the half-level
    weights interpolate to the face exactly, so FSig_GM(face) = -kappa_GM |a(z_f)|^2 / gamma
    and vdiv_k = ((F(top face) - F(bottom face)) RCDK) / H on untapered layers."""
    s = _column_setup()
    km, depth, g = s["pt"].shape[0], 4000.0, 0.005
    a0, a1, b0 = 1e-6, 2e-10, -0.6e-6
    zz = s["z"]
    s["grads"] = (s["R"] * (a0 + a1 * zz), np.full_like(zz, b0 * s["R"])) + s["grads"][2:]
    cfg = RG.Config(kappa_redi=0.0, interior="GKW91")
    ks, _, _, _ = _interior(s, cfg)
    saved = RG.fluxes
    if fluxes is not None:
        RG.fluxes = fluxes
    try:
        fa, fb, vt = _run(s, cfg)
    finally:
        RG.fluxes = saved
    fk = -0.5 + 0.5 * np.cos(np.pi * np.arange(km + 1) / km)
    F = lambda zf: -cfg.kappa_gm * ((a0 + a1 * zf) ** 2 + b0 * b0) / g
    inner = [k for k in ks if k - 1 in ks and k + 1 in ks]
    assert len(inner) >= 8
    err = 0.0
    for k in inner:
        want = ((F(depth * fk[k]) - F(depth * fk[k + 1])) / s["cdk"][k]) / depth
        err = max(err, float(np.max(np.abs(vt[0, k] - want))) / abs(want))
    # the layer fluxes: kappa_GM a(z_k), kappa_GM b0
    cs = s["coslat"][None, :, None]
    for k in ks:
        np.testing.assert_allclose(fa[0, k] / cs[0], cfg.kappa_gm * (a0 + a1 * zz[k]), rtol=1e-6)
        np.testing.assert_allclose(fb[0, k] / cs[0], cfg.kappa_gm * b0, rtol=1e-6)
    return err


def test_gm_skew_flux_closed_form():
    assert _gm_face_error() < 1e-5


def _planted(fn, *edits):
    """A copy of an oracle function with source edits, for planted-mistake checks."""
    import inspect
    import textwrap
    src = textwrap.dedent(inspect.getsource(fn))
    for a, b in edits:
        assert a in src
        src = src.replace(a, b)
    ns = dict(vars(RG))
    exec(compile(src, "<planted>", "exec"), ns)
    return ns[fn.__name__]


def test_planted_mistakes_fail():
    # half-level weights swapped
    swapped = _planted(RG.fluxes, ("m1 = cdk[lo] / (cdk[up] + cdk[lo])", "m1 = cdk[up] / (cdk[up] + cdk[lo])"),
                       ("m2 = cdk[up] / (cdk[up] + cdk[lo])", "m2 = cdk[lo] / (cdk[up] + cdk[lo])"))
    assert _gm_face_error(swapped) > 1e-3
    # the sign of the slope
    s = _column_setup()
    cfg = RG.Config(kappa_gm=0.0, interior="GKW91")
    ks, _, _, _ = _interior(s, cfg)
    saved = RG.slopes
    RG.slopes = _planted(RG.slopes, ("fac = -(h / den)", "fac = h / den"))
    try:
        fa, _, _ = _run(s, cfg)
    finally:
        RG.slopes = saved
    cs = s["coslat"][None, :, None]
    assert np.max(np.abs(fa[0, ks[len(ks) // 2]] / cs)) > 1.0 * cfg.kappa_redi * 1e-6
    # the surface scan of LDD95 not stopping at the first deep layer
    km = 20
    cdk, fdk, rcdk, rfdk, zck = vertical_grid(km)
    z = 4000.0 * zck[:, None, None] * np.ones((km, 1, 1))
    topo = np.full((1, 1), 4000.0)
    lat = np.radians(np.array([30.0]))
    slon = np.full((km, 1, 1), 1e-3)
    slon[8] = 3.5e-3
    slat = np.zeros((km, 1, 1))
    good = RG.taper(slon, slat, z, topo, lat, RG.Config(interior="GKW91"))
    nostop = _planted(RG.taper, ("        done |= hit\n    smax = np.zeros_like(topo)", "    smax = np.zeros_like(topo)"))
    bad = nostop(slon, slat, z, topo, lat, RG.Config(interior="GKW91"))
    assert np.max(np.abs(bad - good)) > 1e-2


def test_fluxes_are_linear_in_the_diffusivities():
    s = _column_setup(gamma=0.002, a=3e-6, b=2e-6)
    r = _run(s, RG.Config(kappa_redi=700.0, kappa_gm=0.0))
    g = _run(s, RG.Config(kappa_redi=0.0, kappa_gm=300.0))
    both = _run(s, RG.Config(kappa_redi=700.0, kappa_gm=300.0))
    for x, y, z in zip(r, g, both):
        np.testing.assert_allclose(z, x + y, rtol=1e-12, atol=1e-20)


def test_taper_fixed_points():
    km = 20
    cdk, fdk, rcdk, rfdk, zck = vertical_grid(km)
    z = 4000.0 * zck[:, None, None] * np.ones((km, 1, 1))
    topo = np.full((1, 1), 4000.0)
    lat = np.radians(np.array([30.0]))
    mid = km // 2
    for interior, mag, want in (("DM95", 4e-3, 0.5), ("GKW91", 8e-3, 0.25)):
        cfg = RG.Config(interior=interior)
        slon = np.full((km, 1, 1), 1e-5)
        slat = np.zeros((km, 1, 1))
        slon[mid] = mag
        t = RG.taper(slon, slat, z, topo, lat, cfg)
        assert abs(t[mid, 0, 0] - want) < 1e-15
    # LDD95: the slope maximum is taken down to the first layer deeper than the eddy
    # displacement only (the scan exits there); a steep slope below it leaves the surface
    # taper unchanged
    cfg = RG.Config(interior="GKW91")
    disph = min(max(15e3, 2.0 / ((2.0 * cfg.omega) * abs(math.sin(lat[0])))), 100e3)
    dzmax = cfg.slope_max * disph
    ks = next(k for k in range(km) if -z[k, 0, 0] > dzmax)
    slon = np.full((km, 1, 1), 1e-3)
    slat = np.zeros((km, 1, 1))
    base = RG.taper(slon, slat, z, topo, lat, cfg)
    slon2 = slon.copy()
    slon2[ks + 1] = 3.5e-3
    steep = RG.taper(slon2, slat, z, topo, lat, cfg)
    np.testing.assert_array_equal(base[:ks + 1], steep[:ks + 1])
    tmp = 1e-3 * disph
    for k in range(ks + 1):
        r = min(-z[k, 0, 0] / (tmp + 1e-13), 1.0)
        assert abs(base[k, 0, 0] - 0.5 * (1.0 + math.sin(math.pi * (r - 0.5)))) < 1e-15
    # the interior is untapered
    kb = max(k for k in range(km) if z[k, 0, 0] + 4000.0 > dzmax)
    assert np.all(base[ks + 1:kb] == 1.0)
    # a steep slope above the surface hit enters the maximum
    slon3 = slon.copy()
    slon3[0] = 2e-3
    t3 = RG.taper(slon3, slat, z, topo, lat, cfg)
    assert t3[0, 0, 0] < base[0, 0, 0] < 1.0


def test_vertical_divergence_telescopes():
    s = _column_setup(gamma=0.003, a=2e-6, b=1e-6)
    rng = np.random.default_rng(3)
    s["pt"] = s["pt"] + 0.05 * rng.standard_normal(s["pt"].shape)
    s["sa"] = s["sa"] + 0.01 * rng.standard_normal(s["sa"].shape)
    s["grads"] = tuple(g * (1 + 0.3 * rng.standard_normal(g.shape)) for g in s["grads"][:2]) + \
        tuple(0.1 * s["R"] * 1e-6 * rng.standard_normal(s["pt"].shape) for _ in range(2))
    fa, fb, vt = _run(s, RG.Config())
    for t in range(2):
        tot = np.einsum("k,kji->ji", s["cdk"], s["h"] * vt[t])
        scale = np.einsum("k,kji->ji", s["cdk"], np.abs(s["h"] * vt[t]))
        assert np.all(np.abs(tot) <= 1e-13 * scale + 1e-300)
        assert np.max(scale) > 0


def test_tendency_conserves_tracer_content():
    import oracle as O
    N, I, J, km = 10, 32, 16, 8
    o = O.Oracle(N, I, J, 6.371e6)
    cdk, fdk, rcdk, rfdk, zck = vertical_grid(km)
    depth = 4000.0
    rng = np.random.default_rng(7)
    lon = np.linspace(0, 2 * np.pi, I, endpoint=False)[None, :]
    mu = np.sin(o.lat)[:, None]
    z = depth * zck[:, None, None] * np.ones((km, J, I))
    pt = 283.0 + 0.004 * z + 1.5 * (mu ** 2) * np.cos(lon)[None] * np.exp(z / 1500.0)
    sa = 35.0 + 0.2 * mu * np.sin(2 * lon)[None] * np.exp(z / 1000.0)
    pt = pt + 0.01 * rng.standard_normal(pt.shape)
    h = np.full((km, J, I), depth)
    topo = np.full((J, I), depth)
    out = RG.add_mixing_term(o, pt, sa, h, z, topo, cdk, rfdk, RG.Config())
    for t in range(2):
        content = sum(cdk[k] * o.avrlonlat(out[t, k][None])[0] for k in range(km))
        scale = sum(cdk[k] * o.avrlonlat(np.abs(out[t, k])[None])[0] for k in range(km))
        assert scale > 0
        assert abs(content) < 1e-12 * scale
