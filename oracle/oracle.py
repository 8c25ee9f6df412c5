"""ctypes wrapper of the FP64 CPU oracle (oracle/dsht_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the parity oracle.  The product path never imports this module.

Array conventions (numpy, C order):
  grid      (ncols, J, I)   -- SpmlUtil xy(0:I-1,1:J) per column, longitude fastest
  spectral  (ncols, L)      -- SPMODEL order, l = n(n+1)+m, Im parts at m < 0
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libdsht_oracle.so")

VAL, DIVLAMBDA, DIVMU, GRADLAMBDA, GRADMU, GRADLON, GRADLAT = range(7)

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def use_library(path):
    """Load the oracle from another build of the same source (e.g. one compiled with
    -march=native for the host that times the CPU baseline); before any other use."""
    global _LIB, _lib
    if _lib is not None:
        raise RuntimeError("the oracle library is already loaded")
    _LIB = path


def build_native(outdir):
    """Compile dsht_oracle.c with -O3 -march=native -fopenmp into outdir (for the host it runs
    on) and return the library path."""
    os.makedirs(outdir, exist_ok=True)
    out = os.path.join(outdir, "libdsht_oracle_native.so")
    subprocess.check_call(["gcc", "-O3", "-march=native", "-fopenmp", "-fPIC", "-std=gnu11", "-shared", "-o", out,
                           os.path.join(_HERE, "dsht_oracle.c"), "-lm"])
    return out


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        dp = ctypes.POINTER(ctypes.c_double)
        vp = ctypes.c_void_p
        i = ctypes.c_int
        d = ctypes.c_double
        L.orc_create.restype = vp
        L.orc_create.argtypes = [i, i, i, d]
        L.orc_destroy.argtypes = [vp]
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_set_threads.argtypes = [i]
        L.orc_set_threads.restype = i
        L.orc_grid.argtypes = [vp, dp, dp, dp, dp, dp, dp]
        L.orc_legendre.argtypes = [vp, i, i, dp, dp]
        L.orc_legendre.restype = i
        L.orc_fft_fwd.argtypes = [vp, dp, dp, i]
        L.orc_fft_inv.argtypes = [vp, dp, dp, i]
        for f in ("orc_fwd", "orc_inv"):
            getattr(L, f).argtypes = [vp, i, dp, dp, i]
            getattr(L, f).restype = i
        for f in ("orc_alpha_fwd", "orc_alpha_inv"):
            getattr(L, f).argtypes = [vp, dp, dp, dp, i]
            getattr(L, f).restype = i
        for f in ("orc_uv2vordiv", "orc_vordiv2uv"):
            getattr(L, f).argtypes = [vp, dp, dp, dp, dp, i]
            getattr(L, f).restype = i
        for f in ("orc_lapla", "orc_laplainv", "orc_intlonlat", "orc_avrlonlat",
                  "orc_ya_intlon", "orc_ya_avrlon"):
            getattr(L, f).argtypes = [vp, dp, dp, i]
        L.orc_interpolate.argtypes = [vp, dp, dp, dp, dp, i]
        L.orc_coef_hdiff.argtypes = [vp, d, d, i, dp]
        L.orc_coef_hvisc.argtypes = [vp, d, d, i, dp]
        L.orc_coef_advfilter.argtypes = [vp, dp]
        L.orc_coef_helmholtz.argtypes = [vp, d, d, dp]
        L.orc_diag.argtypes = [vp, dp, dp, i]
        lg = ctypes.c_long
        L.orc_col_omega.argtypes = [dp, dp, i, dp, dp, lg]
        L.orc_col_hydpres.argtypes = [dp, dp, i, dp, d, dp, lg]
        L.orc_col_uvbarot.argtypes = [dp, dp, dp, i, dp, dp, dp, lg]
        L.orc_col_quick.argtypes = [dp, dp, i, dp, dp, i, dp, lg]
        L.orc_dgtsv.argtypes = [i, dp, dp, dp, dp]
        L.orc_dgtsv.restype = i
        L.orc_col_vdiff.argtypes = [dp, dp, dp, dp, i, dp, dp, d, d, i, i, dp, lg]
        L.orc_col_vdiff.restype = lg
        L.orc_col_vdiff_mom.argtypes = [dp, dp, dp, dp, i, dp, dp, d, d, i, i, dp, dp, lg]
        L.orc_col_vdiff_mom.restype = lg
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def set_threads(n):
    return lib().orc_set_threads(int(n))


class Oracle:
    """One (nmax, im, jm, radius) transform setup on the CPU."""

    def __init__(self, nmax, im, jm, radius=1.0):
        L = lib()
        self.h = L.orc_create(nmax, im, jm, float(radius))
        if not self.h:
            raise ValueError(L.orc_last_error().decode())
        self.N, self.I, self.J, self.R = nmax, im, jm, float(radius)
        self.L = (nmax + 1) ** 2
        g = [np.zeros(jm) for _ in range(4)] + [np.zeros(im), np.zeros(im)]
        L.orc_grid(self.h, _p(g[0]), _p(g[1]), _p(g[2]), _p(g[4]), _p(g[3]), _p(g[5]))
        self.mu, self.wq, self.lat, self.y_lat_weight, self.lon, self.x_lon_weight = g
        self.coslat = np.sqrt((1 - self.mu) * (1 + self.mu))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_destroy(self.h)
            self.h = None

    # index maps (SPMODEL l_nm / nm_l, 0-based)
    def l_nm(self, n, m):
        return n * (n + 1) + m

    def nm_l(self):
        n = np.floor(np.sqrt(np.arange(self.L))).astype(int)
        m = np.arange(self.L) - n * (n + 1)
        return n, m

    def legendre(self, m, n):
        p = np.zeros(self.J)
        h = np.zeros(self.J)
        if lib().orc_legendre(self.h, m, n, _p(p), _p(h)) != 0:
            raise ValueError(lib().orc_last_error().decode())
        return p, h

    def fft_fwd(self, x):
        x = _c(x)
        nrow = x.size // self.I
        G = np.zeros((nrow, self.N + 1, 2))
        lib().orc_fft_fwd(self.h, _p(x), _p(G), nrow)
        return G[..., 0] + 1j * G[..., 1]

    def fft_inv(self, G):
        G = np.ascontiguousarray(np.stack([G.real, G.imag], -1), dtype=np.float64)
        nrow = G.size // (2 * (self.N + 1))
        x = np.zeros((nrow, self.I))
        lib().orc_fft_inv(self.h, _p(G), _p(x), nrow)
        return x

    def _cols_grid(self, a):
        a = _c(a)
        return a, a.size // (self.I * self.J)

    def _cols_spec(self, a):
        a = _c(a)
        return a, a.size // self.L

    def fwd(self, xy, kind=VAL):
        xy, nc = self._cols_grid(xy)
        w = np.zeros((nc, self.L))
        if lib().orc_fwd(self.h, kind, _p(xy), _p(w), nc) != 0:
            raise ValueError(lib().orc_last_error().decode())
        return w

    def inv(self, w, kind=VAL):
        w, nc = self._cols_spec(w)
        xy = np.zeros((nc, self.J, self.I))
        if lib().orc_inv(self.h, kind, _p(w), _p(xy), nc) != 0:
            raise ValueError(lib().orc_last_error().decode())
        return xy

    def alpha_fwd(self, a, b):
        a, nc = self._cols_grid(a)
        b = _c(b)
        w = np.zeros((nc, self.L))
        lib().orc_alpha_fwd(self.h, _p(a), _p(b), _p(w), nc)
        return w

    def alpha_inv(self, wa, wb):
        wa, nc = self._cols_spec(wa)
        wb = _c(wb)
        xy = np.zeros((nc, self.J, self.I))
        lib().orc_alpha_inv(self.h, _p(wa), _p(wb), _p(xy), nc)
        return xy

    def uv2vordiv(self, ucos, vcos):
        ucos, nc = self._cols_grid(ucos)
        vcos = _c(vcos)
        vor = np.zeros((nc, self.L))
        div = np.zeros((nc, self.L))
        lib().orc_uv2vordiv(self.h, _p(ucos), _p(vcos), _p(vor), _p(div), nc)
        return vor, div

    def vordiv2uv(self, vor, div):
        vor, nc = self._cols_spec(vor)
        div = _c(div)
        u = np.zeros((nc, self.J, self.I))
        v = np.zeros((nc, self.J, self.I))
        lib().orc_vordiv2uv(self.h, _p(vor), _p(div), _p(u), _p(v), nc)
        return u, v

    def lapla(self, w):
        w, nc = self._cols_spec(w)
        o = np.zeros((nc, self.L))
        lib().orc_lapla(self.h, _p(w), _p(o), nc)
        return o

    def laplainv(self, w):
        w, nc = self._cols_spec(w)
        o = np.zeros((nc, self.L))
        lib().orc_laplainv(self.h, _p(w), _p(o), nc)
        return o

    def intlonlat(self, xy):
        xy, nc = self._cols_grid(xy)
        o = np.zeros(nc)
        lib().orc_intlonlat(self.h, _p(xy), _p(o), nc)
        return o

    def avrlonlat(self, xy):
        xy, nc = self._cols_grid(xy)
        o = np.zeros(nc)
        lib().orc_avrlonlat(self.h, _p(xy), _p(o), nc)
        return o

    def ya_intlon(self, xy):
        xy, nc = self._cols_grid(xy)
        o = np.zeros((nc, self.J))
        lib().orc_ya_intlon(self.h, _p(xy), _p(o), nc)
        return o

    def ya_avrlon(self, xy):
        xy, nc = self._cols_grid(xy)
        o = np.zeros((nc, self.J))
        lib().orc_ya_avrlon(self.h, _p(xy), _p(o), nc)
        return o

    def interpolate(self, w, lon, lat):
        w = _c(w)
        lon = _c(np.atleast_1d(lon))
        lat = _c(np.atleast_1d(lat))
        o = np.zeros(lon.size)
        lib().orc_interpolate(self.h, _p(w), _p(lon), _p(lat), _p(o), lon.size)
        return o

    def coef_hdiff(self, kappa, numdiff, ord_):
        c = np.zeros(self.L)
        lib().orc_coef_hdiff(self.h, kappa, numdiff, int(ord_), _p(c))
        return c

    def coef_hvisc(self, nu, numdiff, ord_):
        c = np.zeros(self.L)
        lib().orc_coef_hvisc(self.h, nu, numdiff, int(ord_), _p(c))
        return c

    def coef_advfilter(self):
        c = np.zeros(self.L)
        lib().orc_coef_advfilter(self.h, _p(c))
        return c

    def coef_helmholtz(self, diag, c0):
        c = np.zeros(self.L)
        lib().orc_coef_helmholtz(self.h, float(diag), float(c0), _p(c))
        return c

    def diag(self, d, w):
        w, nc = self._cols_spec(np.array(w, dtype=np.float64, copy=True))
        d = _c(d)
        lib().orc_diag(self.h, _p(d), _p(w), nc)
        return w


# ---------------------------------------------------------------- vertical column diagnostics
# Fields are (km, ...) arrays, layer 0 on top; cdk the km layer thicknesses in sigma.

def _col(*arrs):
    a = [_c(x) for x in arrs]
    km = a[0].shape[0]
    if any(x.shape != a[0].shape for x in a):
        raise ValueError("column fields differ in shape")
    return a, km, a[0].size // km


def col_omega(div, h, cdk):
    """HBEDiagnose_OMG: the (km+1, ...) half-level omega; half level k on top of layer k."""
    (div, h), km, npts = _col(div, h)
    cdk = _c(cdk)
    omg = np.empty((km + 1,) + div.shape[1:])
    lib().orc_col_omega(_p(div), _p(h), km, _p(cdk), _p(omg), npts)
    return omg


def col_hydpres(dens, h, cdk, grav):
    """HBEDiagnose_HydPres: (km, ...) hydrostatic pressure at the layer centres."""
    (dens, h), km, npts = _col(dens, h)
    cdk = _c(cdk)
    p = np.empty_like(dens)
    lib().orc_col_hydpres(_p(dens), _p(h), km, _p(cdk), float(grav), _p(p), npts)
    return p


def col_uvbarot(u, v, h, cdk):
    """HBEDiagnose_UVBarot: thickness-weighted layer means of u and v."""
    (u, v, h), km, npts = _col(u, v, h)
    cdk = _c(cdk)
    ub, vb = np.empty(u.shape[1:]), np.empty(u.shape[1:])
    lib().orc_col_uvbarot(_p(u), _p(v), _p(h), km, _p(cdk), _p(ub), _p(vb), npts)
    return ub, vb


def col_quick(omg, trc, cdk, rfdk, linfree=False):
    """calc_ADVFlxZ_QUICK: the (km+1, ...) vertical advective flux at the half levels."""
    (trc,), km, npts = _col(trc)
    omg = _c(omg)
    if omg.shape != (km + 1,) + trc.shape[1:]:
        raise ValueError("omega must have km+1 half levels")
    cdk, rfdk = _c(cdk), _c(rfdk)
    flx = np.empty_like(omg)
    lib().orc_col_quick(_p(omg), _p(trc), km, _p(cdk), _p(rfdk), int(bool(linfree)), _p(flx), npts)
    return flx


def dgtsv(dl, d, du, b):
    """LAPACK DGTSV restated (NRHS = 1): (x, number of row interchanges)."""
    dl, d, du, b = (np.array(a, dtype=np.float64, copy=True) for a in (dl, d, du, b))
    sw = lib().orc_dgtsv(d.size, _p(dl), _p(d), _p(du), _p(b))
    if sw < 0:
        raise ZeroDivisionError(f"zero pivot at row {-sw - 1}")
    return b, sw


def col_vdiff(q0, rhs, kv, e3, cdk, fdk, dt, alpha, sfc_bc="D", btm_bc="N", btm_frict=None):
    """calc_VDiffTRCEq, or with btm_frict calc_VDiffMOMEq: (q0 + the implicit vertical
    diffusion increment, row interchanges)."""
    (q0, rhs, kv, e3), km, npts = _col(q0, rhs, kv, e3)
    cdk, fdk = _c(cdk), _c(fdk)
    qa = np.empty_like(q0)
    if btm_frict is None:
        sw = lib().orc_col_vdiff(_p(q0), _p(rhs), _p(kv), _p(e3), km, _p(cdk), _p(fdk), float(dt), float(alpha),
                                 ord(sfc_bc), ord(btm_bc), _p(qa), npts)
    else:
        bf = _c(btm_frict).reshape(-1)
        if bf.size != npts:
            raise ValueError("btm_frict must have one value per grid point")
        sw = lib().orc_col_vdiff_mom(_p(q0), _p(rhs), _p(kv), _p(e3), km, _p(cdk), _p(fdk), float(dt),
                                     float(alpha), ord(sfc_bc), ord(btm_bc), _p(bf), _p(qa), npts)
    return qa, int(sw)


def spec_energy(w, nmax, enstrophy=False):
    """nma_/na_EnergyFromStreamfunc_wa and the enstrophy forms (SPMODEL, cascaded at
    SpmlUtil_mod.f90:131-132) on the unit sphere for this library's normalisation: E(n, +-m) =
    ((w^2) c_m) (n(n+1)/2), c_0 = 1, c_m = 2; the enstrophy with (n(n+1))^2 / 2.  w: (ncols, L).
    Returns nma (ncols, 2N+1, N+1) [c, m+N, n] (zero where |m| > n) and na (ncols, N+1), the
    sums over m = -n..n in ascending order."""
    w = np.asarray(w, dtype=np.float64).reshape(-1, (nmax + 1) ** 2)
    nc = w.shape[0]
    nma = np.zeros((nc, 2 * nmax + 1, nmax + 1))
    na = np.zeros((nc, nmax + 1))
    for n in range(nmax + 1):
        nn1 = float(n) * float(n + 1)
        f = 0.5 * nn1 * nn1 if enstrophy else 0.5 * nn1
        acc = np.zeros(nc)
        for m in range(-n, n + 1):
            x = w[:, n * (n + 1) + m]
            v = ((x * x) * (1.0 if m == 0 else 2.0)) * f
            nma[:, m + nmax, n] = v
            acc = acc + v
        na[:, n] = acc
    return nma, na
