"""CPU restatement of dogcm's Redi/GM isopycnal mixing tendency on the oracle's transforms.

TEST INFRASTRUCTURE ONLY (imported by tests/ and bench.py's cpu_baseline leg as the oracle
of the device Redi/GM path); the product path never imports it.

It restates LPhys_RediGM_hspm_vfvm_AddMixingTerm (model/dogcm/src/physics/lphys/
LPhys_RediGM_hspm_vfvm_mod.f90:218-372) for the hspm_vfvm grid, in the order of operations
of each routine (left to right, parentheses as written, no fused multiply-add):

  calc_GradTRC              (:376-456)  vertical differences DSig on half levels, horizontal
                                         gradients xy_GradLon_w / xy_GradLat_w of w_xy, / RPlanet
  calc_IsoNeutralSlope_new  (:683-806)  slopes from EOSDriver_alpha_beta (EOSDriver_mod.f90:
                                         343-420) = EOS_JM95_alpha_beta (EOS_JM95_mod.f90:272-332)
  prepare_SlopeTapering     (LPhys_RediGMHelper_mod.f90:175-236): interior TaperingFunc_DM95
                                         (:258-283) or _GKW91 (:238-257) times TaperingFunc_LDD95
                                         (:285-431)
  calc_IsopycDiffFlux       (:460-556)  Redi fluxes
  calc_SkewFlux             (:560-679)  GM skew fluxes
  append_Redi_GM_RHS        (:308-370)  RHS += xy_w(w_AlphaOptr_xy(FLon cos, FLat cos)), then
                                         RHS += (FSig above - FSig below) RCDK / H

The potential density that AddMixingTerm evaluates (:263-268) feeds only the 'linear' PBL
tapering (LinearSlopeTapering), so with the DM95/GKW91 + LDD95 tapering it does not enter the
result and is not computed.  xyr_SLon/xyr_SLat (:780-797) are not read by either flux routine
and are not computed either.

Arrays: per layer (km, J, I), layer 0 on top (the reference's KS); half-level arrays
(km+1, J, I), half level h on top of layer h; 2-D fields (J, I).
"""
from dataclasses import dataclass
import math

import numpy as np


@dataclass
class Config:
    """Parameters (RediGM_nml of model/dogcm/exp/APEOGCirc/APEOGCircExp.nml:82-91; Omega:
    DOGCM_Admin_Constants_mod.f90:158; RefDens, Grav as the LF-AM3 restatement)."""
    kappa_redi: float = 800.0      # DiffCoefRedi [m2/s]
    kappa_gm: float = 800.0        # DiffCoefGM [m2/s]
    slope_max: float = 4e-3        # SlopeMax
    sd: float = 1e-3               # Sd (DM95)
    interior: str = "DM95"         # InteriorTaperingName: DM95 or GKW91
    omega: float = 2.0 * math.pi / (60.0 * 60.0 * 23.9345)   # Omega [1/s]
    ref_dens: float = 1027.0       # RefDens
    grav: float = 9.79764          # Grav


# EOS_JM95_alpha_beta coefficients (McDougall 1987 polynomial of alpha/beta and beta in
# theta [degC], S - 35 [psu], p [dbar]); EOS_JM95_mod.f90:284-313
AB = dict(ab000=0.665157e-1, ab100=0.170907e-1, ab200=-0.203814e-3, ab300=0.298357e-5,
          ab400=-0.255019e-7, ab010=0.378110e-2, ab110=-0.846960e-4, ab011=-0.164759e-6,
          ab012=-0.251520e-11, ab020=-0.678662e-5, ab001=0.380374e-4, ab101=-0.933746e-6,
          ab201=0.791325e-8, ab202=0.512857e-12, ab003=-0.302285e-13)
BB = dict(b000=0.785567e-3, b100=-0.301985e-5, b200=0.555579e-7, b300=-0.415613e-9,
          b010=-0.356603e-6, b110=0.788212e-8, b011=0.408195e-10, b012=-0.602281e-15,
          b020=0.515032e-8, b001=-0.121555e-7, b101=0.192867e-9, b201=-0.213127e-11,
          b002=0.176621e-12, b102=-0.175379e-14, b003=0.121551e-17)


def alpha_beta(theta_k, s, z, cfg):
    """EOSDriver_alpha_beta(theta, S, p = -RefDens Grav Z): theta in K converted by K2degC,
    p in Pa converted by Pa2bar and x10 to dbar (EOSDriver_mod.f90:343-420), then the
    EOS_JM95_alpha_beta polynomial (EOS_JM95_mod.f90:315-330)."""
    a = AB
    b = BB
    t = theta_k - 273.15                       # K2degC
    p = -(cfg.ref_dens * cfg.grav * z)         # -RefDens*Grav*xyz_Z
    p = (1e-05 * p) * 10.0                     # Pa2bar(p)*10d0
    sm = s - 35.0
    adb = ((a["ab000"]
            + t * (a["ab100"] + t * (a["ab200"] + t * (a["ab300"] + t * a["ab400"]))))
           + sm * (((a["ab010"] + a["ab110"] * t) + p * (a["ab011"] + a["ab012"] * p))
                   + sm * a["ab020"])
           + p * ((a["ab001"] + t * (a["ab101"] + t * (a["ab201"] + p * a["ab202"])))
                  + (p * p) * a["ab003"]))
    beta = ((b["b000"]
             + t * (b["b100"] + t * (b["b200"] + t * b["b300"])))
            + sm * (((b["b010"] + b["b110"] * t) + p * (b["b011"] + b["b012"] * p))
                    + sm * b["b020"])
            + p * ((b["b001"] + t * ((b["b101"] + t * b["b201"]) + p * b["b102"]))
                   + p * (b["b002"] + p * b["b003"])))
    return adb * beta, beta


def dsig(x, rfdk):
    """calc_GradTRC (:412-424): (km+1) half levels, DSig[h] = (x[h-1] - x[h]) * z_RFDK, the
    top and bottom half levels copied from their neighbours."""
    km = x.shape[0]
    d = np.empty((km + 1,) + x.shape[1:])
    for h in range(1, km):
        d[h] = (x[h - 1] - x[h]) * rfdk[h]
    d[0] = d[1]
    d[km] = d[km - 1]
    return d


def slopes(ptemp, salt, z, h, hg_t, hg_s, ds_t, ds_s, cfg):
    """calc_IsoNeutralSlope_new (:745-775): SLon, SLat on layers."""
    EPS = 1e-10
    al, be = alpha_beta(ptemp, salt, z, cfg)
    den = (EPS + (al * 0.5) * (ds_t[:-1] + ds_t[1:])) - (be * 0.5) * (ds_s[:-1] + ds_s[1:])
    fac = -(h / den)
    slon = fac * (al * hg_t[0] - be * hg_s[0])
    slat = fac * (al * hg_t[1] - be * hg_s[1])
    return slon, slat


def taper(slon, slat, z, topo, lat, cfg):
    """prepare_SlopeTapering (LPhys_RediGMHelper_mod.f90:175-236) for the interior schemes
    DM95 / GKW91 and the LDD95 boundary-layer scheme: xyz_T = T1 * T2."""
    mag2 = slon * slon + slat * slat
    if cfg.interior == "DM95":                 # TaperingFunc_DM95 (:274-281)
        t1 = 0.5 * (1.0 + np.tanh((cfg.slope_max - np.sqrt(mag2)) / cfg.sd))
    elif cfg.interior == "GKW91":              # TaperingFunc_GKW91 (:250-255)
        with np.errstate(divide="ignore"):
            t1 = np.minimum(1.0, (cfg.slope_max * cfg.slope_max) / mag2)
    else:
        raise ValueError(cfg.interior)
    # TaperingFunc_LDD95 (:318-431), per column
    km = slon.shape[0]
    EPS = 1e-13
    c = 2.0
    sab = np.sqrt(mag2)
    disph = np.array([c / ((2.0 * cfg.omega) * abs(math.sin(la))) for la in lat])
    disph = np.minimum(np.maximum(15e3, disph), 100e3)[:, None] * np.ones_like(topo)
    dzmax = cfg.slope_max * disph
    r = np.ones_like(slon)
    smax = np.zeros_like(topo)
    done = np.zeros(topo.shape, bool)
    for k in range(km):                        # from the surface down (:392-401)
        smax = np.where(~done & (sab[k] > smax), sab[k], smax)
        hit = ~done & ((0.0 - z[k]) > dzmax)
        if hit.any():
            tmp = np.minimum(smax, cfg.slope_max) * disph
            for kk in range(k + 1):
                r[kk] = np.where(hit, np.minimum((0.0 - z[kk]) / (tmp + EPS), 1.0), r[kk])
        done |= hit
    smax = np.zeros_like(topo)
    done = np.zeros(topo.shape, bool)
    for k in range(km - 1, -1, -1):            # from the bottom up (:420-428)
        smax = np.where(~done & (sab[k] > smax), sab[k], smax)
        hit = ~done & ((z[k] - (-topo)) > dzmax)
        if hit.any():
            tmp = np.minimum(smax, cfg.slope_max) * disph
            for kk in range(k, km):
                r[kk] = np.where(hit, np.minimum((z[kk] - (-topo)) / (tmp + EPS), 1.0), r[kk])
        done |= hit
    t2 = 0.5 * (1.0 + np.sin(np.pi * (r - 0.5)))
    return t1 * t2


def fluxes(hg, ds, h, slon, slat, tap, cdk, cfg):
    """calc_IsopycDiffFlux (:497-553) plus calc_SkewFlux (:607-676) of one tracer: the summed
    layer fluxes FLon, FLat and half-level FSig (Redi + GM, each kept separately until the
    sums of append_Redi_GM_RHS)."""
    km = h.shape[0]
    kr, kg = cfg.kappa_redi, cfg.kappa_gm
    dz = (0.5 * (ds[:-1] + ds[1:])) / h
    ts_lon = tap * slon
    ts_lat = tap * slat
    flon_r = kr * (hg[0] + ts_lon * dz)
    flat_r = kr * (hg[1] + ts_lat * dz)
    psi_lon = -((kg * tap) * slat)
    psi_lat = (kg * tap) * slon
    flon_g = -(psi_lat * dz)
    flat_g = psi_lon * dz
    fs_r = np.zeros((km + 1,) + h.shape[1:])
    fs_g = np.zeros_like(fs_r)
    for hh in range(1, km):
        up, lo = hh - 1, hh
        m1 = cdk[lo] / (cdk[up] + cdk[lo])
        m2 = cdk[up] / (cdk[up] + cdk[lo])
        sl = (m1 * tap[up]) * slon[up] + (m2 * tap[lo]) * slon[lo]
        sa = (m1 * tap[up]) * slat[up] + (m2 * tap[lo]) * slat[lo]
        glon = m1 * hg[0][up] + m2 * hg[0][lo]
        glat = m1 * hg[1][up] + m2 * hg[1][lo]
        fs_r[hh] = kr * ((sl * glon + sa * glat)
                         + ((sl * sl + sa * sa) * ds[hh]) / (m1 * h[up] + m2 * h[lo]))
        pl = m1 * psi_lon[up] + m2 * psi_lon[lo]
        pa = m1 * psi_lat[up] + m2 * psi_lat[lo]
        fs_g[hh] = pa * glon - pl * glat
    return flon_r + flon_g, flat_r + flat_g, fs_r + fs_g


def column_part(ptemp, salt, h, z, topo, grads, lat, coslat, cdk, rfdk, cfg, radius):
    """The grid-space part of AddMixingTerm given the gradient grids grads = (GradLon T,
    GradLat T, GradLon S, GradLat S) of xy_Grad{Lon,Lat}_w(w_xy(.)) (not yet divided by
    RPlanet).  Returns (fa, fb, vt): per tracer (temperature, salinity) the AlphaOptr inputs
    (FLon cos, FLat cos) and the vertical flux-divergence term (FSig above - FSig below)
    RCDK / H, each (2, km, J, I)."""
    hg_t = (grads[0] / radius, grads[1] / radius)
    hg_s = (grads[2] / radius, grads[3] / radius)
    ds_t = dsig(ptemp, rfdk)
    ds_s = dsig(salt, rfdk)
    slon, slat = slopes(ptemp, salt, z, h, hg_t, hg_s, ds_t, ds_s, cfg)
    tap = taper(slon, slat, z, topo, lat, cfg)
    rcdk = 1.0 / np.asarray(cdk)
    cs = np.asarray(coslat)[None, :, None]
    fa, fb, vt = [], [], []
    for hg, ds in ((hg_t, ds_t), (hg_s, ds_s)):
        flon, flat, fsig = fluxes(hg, ds, h, slon, slat, tap, cdk, cfg)
        fa.append(flon * cs)
        fb.append(flat * cs)
        vt.append(((fsig[:-1] - fsig[1:]) * rcdk[:, None, None]) / h)
    return np.stack(fa), np.stack(fb), np.stack(vt)


def add_mixing_term(o, ptemp, salt, h, z, topo, cdk, rfdk, cfg, rhs=None, lat=None):
    """LPhys_RediGM_hspm_vfvm_AddMixingTerm on the oracle o (an oracle.Oracle whose radius is
    RPlanet): returns the tracer right-hand sides (2, km, J, I) = rhs (zeros if None) plus
    the Redi/GM tendency.  lat: the grid latitudes to use (default: the oracle's)."""
    from oracle import GRADLON, GRADLAT
    km = ptemp.shape[0]
    lat = o.lat if lat is None else np.asarray(lat)
    coslat = np.array([math.cos(v) for v in lat])
    w = o.fwd(np.concatenate([ptemp, salt]))
    glon = o.inv(w, GRADLON)
    glat = o.inv(w, GRADLAT)
    grads = (glon[:km], glat[:km], glon[km:], glat[km:])
    fa, fb, vt = column_part(ptemp, salt, h, z, topo, grads, lat, coslat, cdk, rfdk, cfg, o.R)
    wa = o.alpha_fwd(fa.reshape(2 * km, o.J, o.I), fb.reshape(2 * km, o.J, o.I))
    hdiv = o.inv(wa).reshape(2, km, o.J, o.I)
    out = np.zeros((2, km, o.J, o.I)) if rhs is None else np.array(rhs, dtype=np.float64)
    out = out + hdiv
    return out + vt
