"""CPU restatement of dogcm's LF-AM3 dynamics step on the oracle's transforms.

TEST INFRASTRUCTURE ONLY (imported by tests/ and bench.py's cpu_baseline leg as the oracle
of the device LF-AM3 step); the product path never imports it.

It restates DOGCM_TInt_LFAM3_advance_Dyn (model/dogcm/src/tint/DOGCM_TInt_LFAM3_mod.f90:
281-880) for the configuration the device step runs: hydrostatic Boussinesq dynamics on
the hspm_vfvm grid (spectral horizontal, finite-volume sigma vertical), a rigid lid
(KinBC_Surface = Rigid: BarotSubCycNumMax = 1, :147-170), a flat bottom of depth D (xyz_H =
D, xyz_Z = D z_CK, DOGCM_Admin_Grid_mod.f90:376-388), the linear equation of state
(EOS_Linear_mod.f90 EOS_Linear_Eval) and no physics forcing (the *_RHS_phy inputs zero).
Each sub-step cites the routine it follows:

  UVBarotDiag    HBEDiagnose_UVBarot   (HBEDiagnose_hspm_vfvm_mod.f90:221-255)
  VorDivDiag     HBEDiagnose_VorDiv    (:259-286)
  OMGDiag        HBEDiagnose_OMG       (:90-143)
  HydPresDiag    HBEDiagnose_HydPres   (:290-325)
  MOMBarocRHS    HBEBaroc_MOMRHS_VorDivForm (HBEBaroc_hspm_vfvm_mod.f90:321-443)
  HTRCRHS        HBEBaroc_HTRCRHS      (:119-202) with calc_ADVFlxZ_QUICK (:233-317)
  MOMBarotRHS    HBEBarot_MOMRHS       (HBEBarot_hspm_vfvm_mod.f90:172-215)
  BarotUpdate    HBEBarot_Update_LinFreeSfc with vareps1 = 0 (:219-324)
  ImplUV         DOGCM_Phys_hspm_vfvm_ImplUV  (DOGCM_Phys_hspm_vfvm_mod.f90:666-748)
  ImplTRC        DOGCM_Phys_hspm_vfvm_ImplTRC (:544-662)

Arrays: per layer (km, J, I), layer 0 on top; 2-D fields (J, I).  The vertical grid is the
reference's: faces z_FK = g_Sig (Chebyshev-Lobatto levels of [-1, 0], at_module_omp.f90:
598-619), centres, thicknesses and centre spacings as DOGCM_GaussSpmVFvmGrid_mod.f90:392-410
builds them.
"""
from dataclasses import dataclass

import numpy as np

import oracle as O


@dataclass
class Config:
    """Parameters of the step (defaults: APEOGCircExp.nml:88-140, EOS_Linear_mod.f90)."""
    dt: float = 3600.0             # DelTime [s]
    depth: float = 5000.0          # flat-bottom depth [m]
    grav: float = 9.79764          # Grav
    ref_dens: float = 1027.0       # RefDens
    alpha: float = 0.51            # CoriolisTermACoef (gamma = 1 - 2 alpha)
    lam: float = 0.5               # VDiffTermACoef
    sub_cycles: int = 1            # SubCycleNum
    hvisc: float = 3e5             # hViscCoef [m2/s]
    hdiff: float = 0.0             # hDiffCoef
    numdiff: float = 1e12          # NumDiffCoefH (shared by momentum and tracers)
    numdiff_ord: int = 4           # NumDiffOrdH
    vvisc: float = 1e-3            # vViscCoef [m2/s]
    vdiff: float = 3e-5            # vDiffCoef
    btm_frict: float = 0.0         # xy_BtmFrictCoef (uniform)
    mom_sfc_bc: str = "N"          # DynBC_Surface (SpecStress)
    mom_btm_bc: str = "D"          # DynBC_Bottom (NoSlip)
    trc_sfc_bc: str = "N"
    trc_btm_bc: str = "N"
    ref_temp: float = 283.0        # EOS_Linear
    ref_salt: float = 35.0
    beta_t: float = 1.67e-4
    beta_s: float = 0.78e-3
    cp0: float = 3986.0
    cs0: float = 1490.0


def vertical_grid(km):
    """(cdk, fdk, rcdk, rfdk, zck) of km layers: z_FK = g_Sig faces from the top (0) down to -1,
    z_CK the face midpoints, z_CDK = z_FK(k-1) - z_FK(k), ghost centres half a layer outside,
    z_FDK(k) = z_CK(k) - z_CK(k+1) for the km+1 half levels (DOGCM_GaussSpmVFvmGrid_mod.f90:
    392-410)."""
    i = np.arange(km + 1)
    fk = -0.5 + 0.5 * np.cos(np.pi * i / km)  # at_module_omp.f90:598-601 on [-1, 0]
    cdk = fk[:-1] - fk[1:]
    ck = np.empty(km + 2)
    ck[1:-1] = 0.5 * (fk[:-1] + fk[1:])
    ck[0] = fk[0] + 0.5 * cdk[0]
    ck[-1] = fk[-1] - 0.5 * cdk[-1]
    fdk = ck[:-1] - ck[1:]
    return cdk, fdk, 1.0 / cdk, 1.0 / fdk, ck[1:-1]


class Step:
    """One LF-AM3 dynamics step for one run of km layers with rotation rate `omega`."""

    def __init__(self, o, km, cfg, omega):
        self.o, self.km, self.cfg, self.omega = o, km, cfg, float(omega)
        self.cdk, self.fdk, self.rcdk, self.rfdk, self.zck = vertical_grid(km)
        J, I = o.J, o.I
        mu = o.mu[:, None] * np.ones((1, I))
        self.cos = o.coslat[:, None] * np.ones((1, I))
        self.cori = 2.0 * self.omega * mu                      # xy_Cori = 2 Omega sin(lat)
        self.H = np.full((km, J, I), cfg.depth)                 # xyz_H = D (rigid lid, flat bottom)
        self.Z = cfg.depth * self.zck[:, None, None] * np.ones((km, J, I))
        self.geopot = cfg.grav * self.Z
        n, _ = o.nm_l()
        self.e = n * (n + 1.0) / o.R ** 2                       # -w_LaplaEigVal
        self.advfilter = o.coef_advfilter()
        self.c_visc = o.coef_hvisc(cfg.hvisc, cfg.numdiff, cfg.numdiff_ord)
        self.c_diff = o.coef_hdiff(cfg.hdiff, cfg.numdiff, cfg.numdiff_ord)
        self.kv = np.full((km, J, I), cfg.vvisc)
        self.kd = np.full((km, J, I), cfg.vdiff)
        self.bf = np.full((J, I), cfg.btm_frict)
        # EOS_Linear_Init
        self.ptemp_dep = -cfg.beta_t
        self.press_dep = 1.0 / (cfg.ref_dens * cfg.cs0 ** 2) - cfg.ref_temp * cfg.beta_t ** 2 / (cfg.ref_dens * cfg.cp0)
        self.salt_dep = cfg.beta_s

    # ------------------------------------------------------------ diagnostics
    def uvbarot(self, u, v, h):
        return O.col_uvbarot(u, v, h, self.cdk)

    def vordiv(self, u, v):
        """HBEDiagnose_VorDiv: calc_UVCosLat2VorDiv then xy_w of both."""
        wv, wd = self.o.uv2vordiv(u * self.cos, v * self.cos)
        return self.o.inv(wv), self.o.inv(wd)

    def omg(self, div, h):
        return O.col_omega(div, h, self.cdk)

    def eos(self, t, s):
        """EOS_Linear_Eval(theta, S, p = -RefDens GeoPot)."""
        c = self.cfg
        p = -c.ref_dens * self.geopot
        return c.ref_dens * (self.ptemp_dep * (t - c.ref_temp) + self.press_dep * p + self.salt_dep * (s - c.ref_salt))

    def hydpres(self, dens, h):
        return O.col_hydpres(dens, h, self.cdk, self.cfg.grav)

    # ------------------------------------------------------------ right-hand sides
    def mom_rhs(self, u, v, omg, vor, h, pres, dens, coru, corv):
        """HBEBaroc_MOMRHS_VorDivForm with zero physics tendencies."""
        o, c, R = self.o, self.cfg, self.o.R
        km = self.km
        vadv_u = np.zeros((km + 1,) + u.shape[1:])
        vadv_v = np.zeros_like(vadv_u)
        for k in range(1, km):  # half level k between layers k-1 and k
            vadv_u[k] = omg[k] * (u[k - 1] - u[k]) * self.rfdk[k]
            vadv_v[k] = omg[k] * (v[k - 1] - v[k]) * self.rfdk[k]
        wgeo = o.fwd(self.geopot)
        glat = o.inv(wgeo, O.GRADLAT)
        glon = o.inv(wgeo, O.GRADLON)
        a = vor * u + coru + 0.5 * (vadv_v[1:] + vadv_v[:-1]) / h - 0.0
        a = self.cos * (a + dens * glat / (c.ref_dens * R))
        b = vor * v + corv - 0.5 * (vadv_u[1:] + vadv_u[:-1]) / h + 0.0
        b = self.cos * (b + dens * glon / (c.ref_dens * R))
        cc = (0.5 * (u ** 2 + v ** 2) + pres / c.ref_dens) / R ** 2
        wvor = -o.alpha_fwd(a, b)
        wdiv = o.alpha_fwd(b, -a) - o.lapla(o.fwd(cc))
        return o.vordiv2uv(wvor, wdiv)

    def htrc_rhs(self, trc, u, v, omg, h):
        """HBEBaroc_HTRCRHS (QUICK vertical flux, rigid lid) with zero physics tendency."""
        o = self.o
        flx = O.col_quick(omg, trc, self.cdk, self.rfdk, linfree=False)
        tmp = h * self.cos * trc
        w = -o.alpha_fwd(u * tmp, v * tmp)
        vf = -(flx[:-1] - flx[1:]) * self.rcdk[:, None, None]
        w = w + o.fwd(vf)
        return o.inv(self.advfilter * w) + h * 0.0

    def impl_uv(self, u0, v0, urhs, vrhs, h):
        """DOGCM_Phys_hspm_vfvm_ImplUV: the vertical viscosity solve (calc_VDiffMOMEq) then
        backward-Euler lateral viscosity in vorticity/divergence (w_Filter = 1)."""
        c, o = self.cfg, self.o
        ua, _ = O.col_vdiff(u0, urhs, self.kv, h, self.cdk, self.fdk, c.dt, c.lam, c.mom_sfc_bc, c.mom_btm_bc,
                            btm_frict=self.bf)
        va, _ = O.col_vdiff(v0, vrhs, self.kv, h, self.cdk, self.fdk, c.dt, c.lam, c.mom_sfc_bc, c.mom_btm_bc,
                            btm_frict=self.bf)
        wv, wd = o.uv2vordiv(ua * self.cos, va * self.cos)
        wv = wv / (1.0 - c.dt * self.c_visc)
        wd = wd / (1.0 - c.dt * self.c_visc)
        return o.vordiv2uv(wv, wd)

    def impl_trc(self, trc0, htrc, ha, h0):
        """DOGCM_Phys_hspm_vfvm_ImplTRC: calc_VDiffTRCEq, then backward-Euler lateral
        diffusion (w_Filter = 1) and the H0/HA factor."""
        c, o = self.cfg, self.o
        ta, _ = O.col_vdiff(trc0, htrc, self.kd, h0, self.cdk, self.fdk, c.dt, c.lam, c.trc_sfc_bc, c.trc_btm_bc)
        ta = o.inv(o.fwd(ta) / (1.0 - c.dt * self.c_diff))
        return h0 * ta / ha

    # ------------------------------------------------------------ barotropic mode
    def mom_barot_rhs(self, coru_b, corv_b, sfcpres, ubf, vbf):
        """HBEBarot_MOMRHS."""
        o, c, R = self.o, self.cfg, self.o.R
        a = self.cos * (coru_b - vbf)
        b = self.cos * (corv_b + ubf)
        wvor = -o.alpha_fwd(a, b)
        wdiv = o.alpha_fwd(b, -a) - o.lapla(o.fwd(sfcpres)) / (c.ref_dens * R ** 2)
        u, v = o.vordiv2uv(wvor, wdiv)
        return u[0], v[0]

    def barot_update(self, uba, vba, spa, dts, pres_tavg):
        """HBEBarot_Update_LinFreeSfc with vareps1 = 0 (rigid lid): project the barotropic
        velocity onto its divergence-free part and add the surface-pressure correction."""
        o, c = self.o, self.cfg
        _, wdiv = o.uv2vordiv(uba * self.cos, vba * self.cos)
        wdiv = wdiv[0]
        diag = 0.0
        wtmp = -wdiv
        coef = np.empty(o.L)
        coef[0] = 0.0
        coef[1:] = 1.0 / (self.e[1:] + diag)
        wdsp = c.ref_dens * coef * wtmp / (pres_tavg * dts)
        wddiv = self.e * coef * wtmp
        dsp = o.inv(wdsp)[0]
        du, dv = o.vordiv2uv(np.zeros(o.L), wddiv)
        return uba + du[0], vba + dv[0], spa + dsp

    # ------------------------------------------------------------ the step
    def advance(self, st):
        """One step.  st: dict of the time levels B and N (UB, VB, TB, SB, PB, UN, VN, TN, SN,
        PN; P the surface pressure).  Returns the new state (B <- N, N <- A) and the A-level
        diagnostics."""
        c = self.cfg
        dt = c.dt
        f = self.cori
        H = self.H
        HA = HB = HN = H
        UB, VB, TB, SB, PB = (st[k] for k in ("UB", "VB", "TB", "SB", "PB"))
        UN, VN, TN, SN, PN = (st[k] for k in ("UN", "VN", "TN", "SN", "PN"))
        gamma = 1.0 - 2.0 * c.alpha
        # preparation
        ubn, vbn = self.uvbarot(UN, VN, HN)
        ubb, vbb = self.uvbarot(UB, VB, HB)
        # predictor
        vor, div = self.vordiv(UN, VN)
        omgn = self.omg(div, HB)
        dens = self.eos(TN, SN)
        hydn = self.hydpres(dens, HN)
        coru = f * (gamma * UN + (1.0 - gamma) * UB)
        corv = f * (gamma * VN + (1.0 - gamma) * VB)
        pres = hydn + PN
        urhs, vrhs = self.mom_rhs(UN, VN, omgn, vor, HN, pres, dens, coru, corv)
        if c.alpha > 0.0:
            cif = 0.5 * (2.0 * dt) * f
            ut = urhs
            urhs = (ut + cif * vrhs) / (1.0 + cif ** 2)
            vrhs = (vrhs - cif * ut) / (1.0 + cif ** 2)
        U = 5.0 / 6.0 * dt * urhs + (2.0 * UN + UB) / 3.0
        V = 5.0 / 6.0 * dt * vrhs + (2.0 * VN + VB) / 3.0
        ubt, vbt = self.uvbarot(U, V, HN)
        U = U - ubt + ubn
        V = V - vbt + vbn
        ht = self.htrc_rhs(TN, UN, VN, omgn, HN)
        hs = self.htrc_rhs(SN, UN, VN, omgn, HN)
        T = ((HN + 0.0) * (2.0 * TN + TB) / 3.0 + 5.0 / 6.0 * dt * ht) / (HN + 0.0)
        S = ((HN + 0.0) * (2.0 * SN + SB) / 3.0 + 5.0 / 6.0 * dt * hs) / (HN + 0.0)
        # corrector
        vor, div = self.vordiv(U, V)
        omg = self.omg(div, HB)
        dens = self.eos(T, S)
        hyd = self.hydpres(dens, H)
        alpha_, gamma_ = 0.5, 0.5
        cn, cb = gamma_ + alpha_, 1.0 - alpha_ - gamma_
        coru_bar = f * (cn * ubn + cb * ubb)
        corv_bar = f * (cn * vbn + cb * vbb)
        coru = f * (cn * UN + cb * UB)
        corv = f * (cn * VN + cb * VB)
        pres = hyd + PN
        urhs, vrhs = self.mom_rhs(U, V, omg, vor, HN, pres, dens, coru, corv)
        UA, VA = self.impl_uv(UN, VN, urhs, vrhs, HN)
        urhs = (UA - UN) / dt
        vrhs = (VA - VN) / dt
        ubf, vbf = self.uvbarot(urhs, vrhs, H)
        ubf = ubf - corv_bar
        vbf = vbf + coru_bar
        cif = alpha_ * dt * f
        ut = urhs
        urhs = (ut + cif * vrhs) / (1.0 + cif ** 2)
        vrhs = (vrhs - cif * ut) / (1.0 + cif ** 2)
        UA = UN + dt * urhs
        VA = VN + dt * vrhs
        # barotropic mode: one iteration (rigid lid)
        dts = dt / c.sub_cycles
        ubar, vbar, sp = ubn, vbn, PN
        coru_b = f * ubar
        corv_b = f * vbar
        spm = sp - PN
        ubr, vbr = self.mom_barot_rhs(coru_b, corv_b, spm, ubf, vbf)
        cif = alpha_ * dts * f
        ubar = ubar + dts * (ubr + cif * vbr) / (1.0 + cif ** 2)
        vbar = vbar + dts * (vbr - cif * ubr) / (1.0 + cif ** 2)
        uba, vba, PA = self.barot_update(ubar, vbar, sp, dts, alpha_)
        # full velocity and the mid-level fields of the tracer update
        ubt, vbt = self.uvbarot(UA, VA, HA)
        UA = UA - ubt + uba
        VA = VA - vbt + vba
        U = (5.0 * UA + 8.0 * UN - UB) / 12.0
        V = (5.0 * VA + 8.0 * VN - VB) / 12.0
        Hm = (5.0 * HA + 8.0 * HN - HB) / 12.0
        vor, div = self.vordiv(U, V)
        omg = self.omg(div, HN)
        ht = self.htrc_rhs(T, U, V, omg, Hm)
        hs = self.htrc_rhs(S, U, V, omg, Hm)
        TA = self.impl_trc(TN, ht, HA, HN)
        SA = self.impl_trc(SN, hs, HA, HN)
        new = {"UB": UN, "VB": VN, "TB": TN, "SB": SN, "PB": PN,
               "UN": UA, "VN": VA, "TN": TA, "SN": SA, "PN": PA}
        return new


def initial_state(o, km, seed, u_amp=0.05, t_amp=1.0, s_amp=0.05, depth=5000.0):
    """A seeded red-spectrum state: velocities from random vorticity/divergence spectra
    (variance ~ (n+1)^-3) scaled to an RMS of about u_amp, temperature 25 -> 2 C (in K) and
    salinity 35 -/+ 0.5 psu depth backgrounds with red-spectrum perturbations; B = N."""
    rng = np.random.default_rng(seed)
    n, _ = o.nm_l()
    sd = np.sqrt(0.5 * (n + 1.0) ** -3)
    vor = rng.standard_normal((km, o.L)) * sd
    div = rng.standard_normal((km, o.L)) * sd * 0.1
    vor[:, 0] = div[:, 0] = 0.0
    u, v = o.vordiv2uv(vor, div)
    scale = u_amp / max(np.sqrt(np.mean(u ** 2 + v ** 2)), 1e-300)
    u, v = u * scale, v * scale
    zck = vertical_grid(km)[4]
    z = depth * zck
    tp = o.inv(rng.standard_normal((km, o.L)) * sd)
    sp = o.inv(rng.standard_normal((km, o.L)) * sd)
    t = 273.15 + 2.0 + 23.0 * np.exp(z / 1000.0)[:, None, None] + t_amp * tp / tp.std()
    s = 34.5 + 1.0 * np.exp(z / 1000.0)[:, None, None] + s_amp * sp / sp.std()
    p = np.zeros((o.J, o.I))
    return {"UB": u, "VB": v, "TB": t, "SB": s, "PB": p, "UN": u.copy(), "VN": v.copy(), "TN": t.copy(),
            "SN": s.copy(), "PN": p.copy()}
