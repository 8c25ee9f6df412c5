"""Python host wrapper of libdsht.so (the C ABI in include/dsht.h).

Mirrors the horizontal SpmlUtil_mod API of dogcm (model/dogcm/src/util/math/
SpmlUtil_mod.f90:71-135): the same operator names with the same argument meaning, so a
test written against SpmlUtil reads the same here.  Arrays follow the Fortran layouts
in C order:

    grid      numpy (ncols, J, I) or (J, I)  == Fortran xya(0:I-1, J, ncols)
    spectral  numpy (ncols, L)   or (L,)     == Fortran wa(L, ncols), SPMODEL l order

The device-resident entry points (``*_d``) take torch CUDA tensors and run on the
current torch stream.  There is no CPU fallback: if libdsht.so is missing or no GPU is
present the calls raise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libdsht.so")

# kinds
VAL, DIVLAMBDA, DIVMU, GRADLAMBDA, GRADMU, GRADLON, GRADLAT = range(7)
# precision forms
PREC_FP64, PREC_BF16X3, PREC_FP16, PREC_BF16X6, PREC_FP16X3 = range(5)
PREC_NAMES = {PREC_FP64: "fp64", PREC_BF16X3: "bf16x3", PREC_FP16: "fp16",
              PREC_BF16X6: "bf16x6", PREC_FP16X3: "fp16x3"}
# coefficient tables
(COEF_LAPLA, COEF_LAPLAINV, COEF_HDIFF, COEF_HVISC, COEF_ADVFILTER, COEF_HELMHOLTZ,
 COEF_BEDIFF) = range(7)

# every entry point declared in include/dsht.h
EXPORTS = [
    "dsht_create", "dsht_destroy", "dsht_last_error", "dsht_get_dims", "dsht_set_precision",
    "dsht_get_precision", "dsht_grid_info", "dsht_sig_grid", "dsht_nm_l", "dsht_l_nm", "dsht_rn", "dsht_coef",
    "dsht_w_xy", "dsht_xy_w", "dsht_xy_vgrad_w", "dsht_w_alpha_xy", "dsht_xy_alpha_w", "dsht_uvcoslat2vordiv",
    "dsht_vordiv2uv", "dsht_w_diag_w", "dsht_xy_diag_xy", "dsht_intlonlat", "dsht_avrlonlat",
    "dsht_ya_intlon", "dsht_ya_avrlon", "dsht_interpolate", "dsht_fwd_d", "dsht_inv_d", "dsht_inv_vgrad_d", "dsht_inv_noalloc_d",
    "dsht_fwd_f32_d", "dsht_inv_f32_d", "dsht_inv_vgrad_f32_d",
    "dsht_alpha_fwd_d", "dsht_alpha_inv_d", "dsht_uv2vordiv_d", "dsht_vordiv2uv_d",
    "dsht_diag_d", "dsht_intlonlat_d", "dsht_avrlonlat_d", "dsht_ya_intlon_d",
    "dsht_interpolate_d", "dsht_coef_d", "dsht_col_omega_d", "dsht_col_hydpres_d", "dsht_col_uvbarot_d",
    "dsht_col_quick_d", "dsht_col_vdiff_d", "dsht_col_vdiff_mom_d", "dsht_col_omega", "dsht_col_hydpres",
    "dsht_col_uvbarot", "dsht_col_quick", "dsht_col_vdiff", "dsht_prof_enable", "dsht_prof_reset", "dsht_prof_query",
    "dsht_set_tuning", "dsht_get_tuning", "dsht_batch_create", "dsht_batch_add", "dsht_batch_finalize",
    "dsht_batch_info", "dsht_batch_ngroups", "dsht_batch_buffers", "dsht_batch_step",
    "dsht_batch_step_sequential", "dsht_batch_capture", "dsht_batch_launch", "dsht_batch_destroy",
    "dsht_lfam3_default_config", "dsht_lfam3_create", "dsht_lfam3_destroy", "dsht_lfam3_set_rotation",
    "dsht_lfam3_state", "dsht_lfam3_step", "dsht_lfam3_copy_run", "dsht_lfam3_units",
    "dsht_redigm_default_config", "dsht_col_redigm_d", "dsht_redigm_work_size", "dsht_redigm_d", "dsht_redigm",
    "dsht_spec_energy", "dsht_spec_energy_d",
]
SPEC_ENERGY, SPEC_ENSTROPHY = 0, 1
TUNE_CHUNK_COLS, TUNE_LEG_CB, TUNE_FFT, TUNE_LEG, TUNE_FFT_ORDER = 0, 1, 2, 3, 5


class LfAm3Config(ctypes.Structure):
    """dsht_lfam3_config (include/dsht.h)."""
    _fields_ = [("km", ctypes.c_int), ("max_runs", ctypes.c_int), ("sub_cycles", ctypes.c_int),
                ("numdiff_ord", ctypes.c_int),
                ("dt", ctypes.c_double), ("depth", ctypes.c_double), ("grav", ctypes.c_double),
                ("ref_dens", ctypes.c_double), ("alpha", ctypes.c_double), ("lam", ctypes.c_double),
                ("hvisc", ctypes.c_double), ("hdiff", ctypes.c_double), ("numdiff", ctypes.c_double),
                ("vvisc", ctypes.c_double), ("vdiff", ctypes.c_double), ("btm_frict", ctypes.c_double),
                ("mom_sfc_bc", ctypes.c_int), ("mom_btm_bc", ctypes.c_int), ("trc_sfc_bc", ctypes.c_int),
                ("trc_btm_bc", ctypes.c_int),
                ("ref_temp", ctypes.c_double), ("ref_salt", ctypes.c_double), ("beta_t", ctypes.c_double),
                ("beta_s", ctypes.c_double), ("cp0", ctypes.c_double), ("cs0", ctypes.c_double)]


class RediGMConfig(ctypes.Structure):
    """dsht_redigm_config (include/dsht.h)."""
    _fields_ = [("kappa_redi", ctypes.c_double), ("kappa_gm", ctypes.c_double), ("slope_max", ctypes.c_double),
                ("sd", ctypes.c_double), ("interior_taper", ctypes.c_int), ("pbl_taper", ctypes.c_int),
                ("omega", ctypes.c_double), ("ref_dens", ctypes.c_double), ("grav", ctypes.c_double)]


TAPER_GKW91, TAPER_DM95, TAPER_LDD95 = 1, 2, 3

LF_U, LF_V, LF_TRC, LF_SFCPRES = range(4)
LF_B, LF_N, LF_A = range(3)


class DshtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"dsht error {code}: {msg}")
        self.code = code


_lib = None


def load():
    """Loads libdsht.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run python dennou-ogcm_amd/build.py")
    L = ctypes.CDLL(LIB_PATH)
    vp, i, d, lg = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_long
    pp = ctypes.POINTER(ctypes.c_void_p)
    ip = ctypes.POINTER(ctypes.c_int)
    sig = {
        "dsht_create": ([i, i, i, i, d, pp], i),
        "dsht_destroy": ([vp], i),
        "dsht_last_error": ([], ctypes.c_char_p),
        "dsht_get_dims": ([vp, ip, ip, ip, ip], i),
        "dsht_set_precision": ([vp, i], i),
        "dsht_get_precision": ([vp], i),
        "dsht_grid_info": ([vp, vp, vp, vp, vp, vp, vp], i),
        "dsht_sig_grid": ([i, i, d, d, vp, vp], i),
        "dsht_nm_l": ([vp, vp, vp], i),
        "dsht_l_nm": ([vp, i, i], i),
        "dsht_rn": ([vp, vp], i),
        "dsht_coef": ([vp, i, vp, i, vp], i),
        "dsht_w_xy": ([vp, i, vp, vp, i], i),
        "dsht_xy_w": ([vp, i, vp, vp, i], i),
        "dsht_xy_vgrad_w": ([vp, vp, vp, vp, vp, i], i),
        "dsht_w_alpha_xy": ([vp, vp, vp, vp, i], i),
        "dsht_xy_alpha_w": ([vp, vp, vp, vp, i], i),
        "dsht_uvcoslat2vordiv": ([vp, vp, vp, vp, vp, i], i),
        "dsht_vordiv2uv": ([vp, vp, vp, vp, vp, i], i),
        "dsht_w_diag_w": ([vp, vp, vp, vp, i], i),
        "dsht_xy_diag_xy": ([vp, vp, vp, vp, i], i),
        "dsht_intlonlat": ([vp, vp, vp, i], i),
        "dsht_avrlonlat": ([vp, vp, vp, i], i),
        "dsht_ya_intlon": ([vp, vp, vp, i], i),
        "dsht_ya_avrlon": ([vp, vp, vp, i], i),
        "dsht_interpolate": ([vp, vp, vp, vp, vp, i], i),
        "dsht_fwd_d": ([vp, i, vp, lg, vp, lg, lg, i, vp, vp], i),
        "dsht_inv_d": ([vp, i, vp, lg, lg, vp, lg, i, vp, vp], i),
        "dsht_inv_noalloc_d": ([vp, i, vp, lg, lg, vp, lg, i, vp, vp], i),
        "dsht_fwd_f32_d": ([vp, i, vp, lg, vp, lg, lg, i, vp, vp], i),
        "dsht_inv_f32_d": ([vp, i, vp, lg, lg, vp, lg, i, vp, vp], i),
        "dsht_inv_vgrad_f32_d": ([vp, vp, lg, lg, vp, vp, vp, lg, i, vp, vp], i),
        "dsht_inv_vgrad_d": ([vp, vp, lg, lg, vp, vp, vp, lg, i, vp, vp], i),
        "dsht_alpha_fwd_d": ([vp, vp, vp, lg, vp, lg, lg, i, vp, vp], i),
        "dsht_alpha_inv_d": ([vp, vp, vp, lg, lg, vp, lg, i, vp], i),
        "dsht_uv2vordiv_d": ([vp, vp, vp, lg, vp, vp, lg, lg, i, vp, vp], i),
        "dsht_vordiv2uv_d": ([vp, vp, vp, lg, lg, vp, vp, lg, i, vp], i),
        "dsht_diag_d": ([vp, vp, vp, vp, lg, lg, i, vp], i),
        "dsht_intlonlat_d": ([vp, vp, lg, vp, i, vp], i),
        "dsht_avrlonlat_d": ([vp, vp, lg, vp, i, vp], i),
        "dsht_ya_intlon_d": ([vp, vp, lg, vp, i, vp], i),
        "dsht_interpolate_d": ([vp, vp, lg, vp, vp, vp, i, vp], i),
        "dsht_coef_d": ([vp, i, vp, i, vp, vp], i),
        "dsht_col_omega_d": ([vp, vp, vp, lg, i, vp, vp, lg, vp], i),
        "dsht_col_hydpres_d": ([vp, vp, vp, lg, i, vp, ctypes.c_double, vp, lg, vp], i),
        "dsht_col_uvbarot_d": ([vp, vp, vp, vp, lg, i, vp, vp, vp, vp], i),
        "dsht_col_quick_d": ([vp, vp, lg, vp, lg, i, vp, vp, i, vp, lg, vp], i),
        "dsht_col_vdiff_d": ([vp, vp, vp, vp, vp, lg, i, vp, vp, ctypes.c_double, ctypes.c_double, i, i, vp, lg,
                              vp, lg, vp], i),
        "dsht_col_vdiff_mom_d": ([vp, vp, vp, vp, vp, lg, i, vp, vp, ctypes.c_double, ctypes.c_double, i, i, vp, vp,
                                  lg, vp, lg, vp], i),
        "dsht_col_omega": ([vp, vp, vp, i, vp, vp], i),
        "dsht_col_hydpres": ([vp, vp, vp, i, vp, ctypes.c_double, vp], i),
        "dsht_col_uvbarot": ([vp, vp, vp, vp, i, vp, vp, vp], i),
        "dsht_col_quick": ([vp, vp, vp, i, vp, vp, i, vp], i),
        "dsht_col_vdiff": ([vp, vp, vp, vp, vp, i, vp, vp, ctypes.c_double, ctypes.c_double, i, i, vp, vp], i),
        "dsht_prof_enable": ([i], i),
        "dsht_prof_reset": ([], i),
        "dsht_prof_query": ([ctypes.c_char_p, vp, vp, vp, vp], i),
        "dsht_set_tuning": ([vp, i, i], i),
        "dsht_get_tuning": ([vp, i], i),
        "dsht_batch_create": ([pp], i),
        "dsht_batch_add": ([vp, vp, i, ip], i),
        "dsht_batch_finalize": ([vp], i),
        "dsht_batch_info": ([vp, i, ip, ip, ip, ctypes.POINTER(ctypes.c_long)], i),
        "dsht_batch_ngroups": ([vp], i),
        "dsht_batch_buffers": ([vp, i, pp, pp, pp, pp], i),
        "dsht_batch_step": ([vp, vp], i),
        "dsht_batch_step_sequential": ([vp, vp], i),
        "dsht_batch_capture": ([vp, vp], i),
        "dsht_batch_launch": ([vp, vp], i),
        "dsht_batch_destroy": ([vp], i),
        "dsht_lfam3_default_config": ([ctypes.POINTER(LfAm3Config)], i),
        "dsht_lfam3_create": ([vp, ctypes.POINTER(LfAm3Config), pp], i),
        "dsht_lfam3_destroy": ([vp], i),
        "dsht_lfam3_set_rotation": ([vp, i, d], i),
        "dsht_lfam3_state": ([vp, i, i, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_long)], i),
        "dsht_lfam3_step": ([vp, i, vp], i),
        "dsht_lfam3_copy_run": ([vp, i, i, vp], i),
        "dsht_lfam3_units": ([vp, i, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long)], i),
        "dsht_spec_energy": ([vp, i, vp, i, vp, vp], i),
        "dsht_spec_energy_d": ([vp, i, vp, lg, lg, i, vp, vp, vp], i),
        "dsht_redigm_default_config": ([ctypes.POINTER(RediGMConfig)], i),
        "dsht_col_redigm_d": ([vp, ctypes.POINTER(RediGMConfig), vp, vp, vp, vp, lg, i, vp, vp, vp, vp, vp, vp, vp,
                               vp, lg, vp], i),
        "dsht_redigm_work_size": ([vp, i], lg),
        "dsht_redigm_d": ([vp, ctypes.POINTER(RediGMConfig), vp, vp, vp, vp, lg, i, vp, vp, vp, vp, vp, lg, vp, lg,
                           vp], i),
        "dsht_redigm": ([vp, ctypes.POINTER(RediGMConfig), vp, vp, vp, vp, i, vp, vp, vp, vp, vp], i),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise DshtError(rc, load().dsht_last_error().decode())


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _tptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream=None):
    if stream is not None:
        return ctypes.c_void_p(stream)
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def prof_enable(on=True):
    _check(load().dsht_prof_enable(1 if on else 0))


def prof_reset():
    _check(load().dsht_prof_reset())


def prof_query(name):
    """(total_ms, count, algorithmic_bytes, algorithmic_flops) of one kernel class."""
    ms, by, fl = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    n = ctypes.c_long()
    _check(load().dsht_prof_query(name.encode(), ctypes.byref(ms), ctypes.byref(n),
                                  ctypes.byref(by), ctypes.byref(fl)))
    return ms.value, n.value, by.value, fl.value


def sig_grid(kmax, tmax, sig_min=-1.0, sig_max=0.0):
    """(g_Sig, g_Sig_WEIGHT): the kmax+1 Chebyshev-Lobatto levels of [sig_min, sig_max] and
    their Clenshaw-Curtis weights, the vertical half of get_SpmlGridInfo
    (SpmlUtil_mod.f90:288,291; at_module_omp.f90:598-619).  Host-only."""
    sig, wgt = np.zeros(kmax + 1), np.zeros(kmax + 1)
    _check(load().dsht_sig_grid(int(kmax), int(tmax), float(sig_min), float(sig_max), _ptr(sig), _ptr(wgt)))
    return sig, wgt


class Plan:
    """One transform set-up: replaces SpmlUtil_Init(iMax, jMax, kMax, nMax, tMax, RPlanet)."""

    def __init__(self, nmax, im, jm, kmax=1, radius=1.0):
        L = load()
        h = ctypes.c_void_p()
        _check(L.dsht_create(int(nmax), int(im), int(jm), int(kmax), float(radius), ctypes.byref(h)))
        self.h = h
        self.nmax, self.im, self.jm, self.kmax, self.radius = nmax, im, jm, kmax, float(radius)
        self.lm = (nmax + 1) ** 2
        x = [np.zeros(im), np.zeros(jm), np.zeros(im), np.zeros(jm),
             np.zeros((jm, im)), np.zeros((jm, im))]
        _check(L.dsht_grid_info(h, *[_ptr(a) for a in x]))
        (self.x_Lon, self.y_Lat, self.x_Lon_Weight, self.y_Lat_Weight,
         self.xy_Lon, self.xy_Lat) = x
        self.xy_CosLat = np.cos(self.xy_Lat)

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            load().dsht_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -------------------------------------------------- configuration / maps
    def set_precision(self, prec):
        _check(load().dsht_set_precision(self.h, int(prec)))

    @property
    def precision(self):
        return load().dsht_get_precision(self.h)

    def nm_l(self):
        n = np.zeros(self.lm, dtype=np.int32)
        m = np.zeros(self.lm, dtype=np.int32)
        _check(load().dsht_nm_l(self.h, _ptr(n), _ptr(m)))
        return n, m

    def l_nm(self, n, m):
        """1-based SPMODEL index (0 for an invalid pair)."""
        return load().dsht_l_nm(self.h, int(n), int(m))

    @property
    def rn(self):
        r = np.zeros(self.lm)
        _check(load().dsht_rn(self.h, _ptr(r)))
        return r

    def coef(self, which, params=()):
        p = _f64(np.asarray(params, dtype=np.float64).reshape(-1)) if len(params) else np.zeros(1)
        out = np.zeros(self.lm)
        _check(load().dsht_coef(self.h, int(which), _ptr(p), len(params), _ptr(out)))
        return out

    # -------------------------------------------------- host API helpers
    def _grid_in(self, a):
        a = _f64(a)
        if a.size % (self.im * self.jm):
            raise ValueError(f"grid array size {a.size} is not a multiple of I*J")
        return a, a.size // (self.im * self.jm)

    def _spec_in(self, a):
        a = _f64(a)
        if a.size % self.lm:
            raise ValueError(f"spectral array size {a.size} is not a multiple of L")
        return a, a.size // self.lm

    def _fwd(self, xy, kind):
        a, nc = self._grid_in(xy)
        w = np.empty((nc, self.lm)) if a.ndim == 3 else np.empty(self.lm)
        _check(load().dsht_w_xy(self.h, kind, _ptr(a), _ptr(w), nc))
        return w

    def _inv(self, w, kind):
        a, nc = self._spec_in(w)
        xy = np.empty((nc, self.jm, self.im)) if a.ndim == 2 else np.empty((self.jm, self.im))
        _check(load().dsht_xy_w(self.h, kind, _ptr(a), _ptr(xy), nc))
        return xy

    # -------------------------------------------------- SpmlUtil / SPMODEL names
    def w_xy(self, xy):
        return self._fwd(xy, VAL)

    wa_xya = w_xy

    def xy_w(self, w):
        return self._inv(w, VAL)

    xya_wa = xy_w

    def w_DivLambda_xy(self, xy):
        return self._fwd(xy, DIVLAMBDA)

    def w_DivMu_xy(self, xy):
        return self._fwd(xy, DIVMU)

    def xy_GradLambda_w(self, w):
        return self._inv(w, GRADLAMBDA)

    def xy_GradMu_w(self, w):
        return self._inv(w, GRADMU)

    def xy_GradLon_w(self, w):
        return self._inv(w, GRADLON)

    def xy_GradLat_w(self, w):
        return self._inv(w, GRADLAT)

    def xy_GradLonLat_w(self, w, value=True):
        """(xy_w(w), xy_GradLon_w(w), xy_GradLat_w(w)) from one fused pass; the first item is
        None when value is False."""
        a, nc = self._spec_in(w)
        shape = (nc, self.jm, self.im) if a.ndim == 2 else (self.jm, self.im)
        xy = np.empty(shape) if value else None
        lon, lat = np.empty(shape), np.empty(shape)
        _check(load().dsht_xy_vgrad_w(self.h, _ptr(a), _ptr(xy) if value else None, _ptr(lon), _ptr(lat), nc))
        return xy, lon, lat

    xya_GradLonLat_wa = xy_GradLonLat_w

    xya_GradLambda_wa = xy_GradLambda_w
    xya_GradMu_wa = xy_GradMu_w
    xya_GradLon_wa = xy_GradLon_w
    xya_GradLat_wa = xy_GradLat_w

    def w_AlphaOptr_xy(self, xy_A, xy_B):
        a, nc = self._grid_in(xy_A)
        b, nc2 = self._grid_in(xy_B)
        if nc != nc2:
            raise ValueError("A and B differ in size")
        w = np.empty((nc, self.lm)) if a.ndim == 3 else np.empty(self.lm)
        _check(load().dsht_w_alpha_xy(self.h, _ptr(a), _ptr(b), _ptr(w), nc))
        return w

    wa_AlphaOptr_xya = w_AlphaOptr_xy

    def xy_AlphaOptr_w(self, w_A, w_B):
        a, nc = self._spec_in(w_A)
        b, nc2 = self._spec_in(w_B)
        if nc != nc2:
            raise ValueError("A and B differ in size")
        xy = np.empty((nc, self.jm, self.im)) if a.ndim == 2 else np.empty((self.jm, self.im))
        _check(load().dsht_xy_alpha_w(self.h, _ptr(a), _ptr(b), _ptr(xy), nc))
        return xy

    xya_AlphaOptr_wa = xy_AlphaOptr_w

    def calc_UVCosLat2VorDiv(self, xy_UCosLat, xy_VCosLat):
        u, nc = self._grid_in(xy_UCosLat)
        v, _ = self._grid_in(xy_VCosLat)
        shape = (nc, self.lm) if u.ndim == 3 else (self.lm,)
        vor, div = np.empty(shape), np.empty(shape)
        _check(load().dsht_uvcoslat2vordiv(self.h, _ptr(u), _ptr(v), _ptr(vor), _ptr(div), nc))
        return vor, div

    def calc_VorDiv2UV(self, w_Vor, w_Div):
        z, nc = self._spec_in(w_Vor)
        d, _ = self._spec_in(w_Div)
        shape = (nc, self.jm, self.im) if z.ndim == 2 else (self.jm, self.im)
        u, v = np.empty(shape), np.empty(shape)
        _check(load().dsht_vordiv2uv(self.h, _ptr(z), _ptr(d), _ptr(u), _ptr(v), nc))
        return u, v

    def w_diag_w(self, d, w):
        a, nc = self._spec_in(w)
        dd = _f64(d)
        out = np.empty_like(a)
        _check(load().dsht_w_diag_w(self.h, _ptr(dd), _ptr(a), _ptr(out), nc))
        return out

    def w_Lapla_w(self, w):
        return self.w_diag_w(self.coef(COEF_LAPLA), w)

    def w_LaplaInv_w(self, w):
        return self.w_diag_w(self.coef(COEF_LAPLAINV), w)

    def xy_diag_xy(self, d, xy):
        """xy_w(d * w_xy(xy)) in one call (the implicit-diffusion epilogue)."""
        a, nc = self._grid_in(xy)
        dd = _f64(d)
        out = np.empty_like(a)
        _check(load().dsht_xy_diag_xy(self.h, _ptr(dd), _ptr(a), _ptr(out), nc))
        return out

    def _sum(self, fn, xy, per_lat):
        a, nc = self._grid_in(xy)
        out = np.empty((nc, self.jm) if per_lat else nc)
        _check(getattr(load(), fn)(self.h, _ptr(a), _ptr(out), nc))
        if a.ndim == 2:
            return out[0] if per_lat else float(out[0])
        return out

    def IntLonLat_xy(self, xy):
        return self._sum("dsht_intlonlat", xy, False)

    def AvrLonLat_xy(self, xy):
        return self._sum("dsht_avrlonlat", xy, False)

    def ya_IntLon_xya(self, xya):
        return self._sum("dsht_ya_intlon", xya, True)

    def ya_AvrLon_xya(self, xya):
        return self._sum("dsht_ya_avrlon", xya, True)

    def Interpolate_w(self, w, alon, alat):
        ww = _f64(w)
        lon = _f64(np.atleast_1d(alon))
        lat = _f64(np.atleast_1d(alat))
        out = np.empty(lon.size)
        _check(load().dsht_interpolate(self.h, _ptr(ww), _ptr(lon), _ptr(lat), _ptr(out), lon.size))
        return out if np.ndim(alon) else float(out[0])

    # -------------------------------------------------- device-resident API (torch)
    # grids: tensor (ncols, J, I) float64 CUDA; spectra: tensor (L, ncols) ("LC", fast) or
    # (ncols, L) ("CL", Fortran order) float64 CUDA.
    def _spec_strides(self, t, ncols, layout):
        if layout == "LC":
            return t.stride(0), 1
        return 1, t.stride(0)

    def fwd_d(self, kind, xy, w, layout="LC", diag=None, stream=None):
        nc = xy.shape[0]
        sl, sc = self._spec_strides(w, nc, layout)
        _check(load().dsht_fwd_d(self.h, kind, _tptr(xy), xy.stride(0), _tptr(w), sl, sc, nc,
                                 _tptr(diag) if diag is not None else None, _stream(stream)))
        return w

    def inv_d(self, kind, w, xy, layout="LC", diag=None, stream=None):
        nc = xy.shape[0]
        sl, sc = self._spec_strides(w, nc, layout)
        _check(load().dsht_inv_d(self.h, kind, _tptr(w), sl, sc, _tptr(xy), xy.stride(0), nc,
                                 _tptr(diag) if diag is not None else None, _stream(stream)))
        return xy

    def inv_noalloc_d(self, kind, w, xy, layout="LC", diag=None, stream=None):
        """inv_d through dsht_inv_noalloc_d: no device allocation (static scratch shared by every
        call on the GPU; FP64 form only)"""
        nc = xy.shape[0]
        sl, sc = self._spec_strides(w, nc, layout)
        _check(load().dsht_inv_noalloc_d(self.h, kind, _tptr(w), sl, sc, _tptr(xy), xy.stride(0), nc,
                                         _tptr(diag) if diag is not None else None, _stream(stream)))
        return xy

    def inv_vgrad_d(self, w, xy, xy_lon, xy_lat, layout="LC", diag=None, stream=None):
        """value, GradLon and GradLat grids of the same spectra in one fused pass; any one or
        two of the outputs may be None"""
        outs = [t for t in (xy, xy_lon, xy_lat) if t is not None]
        if not outs:
            raise ValueError("inv_vgrad_d: every output is None")
        nc = outs[0].shape[0]
        if any(t.shape[0] != nc or t.stride(0) != outs[0].stride(0) for t in outs):
            raise ValueError("inv_vgrad_d: outputs differ in columns or pitch")
        sl, sc = self._spec_strides(w, nc, layout)
        tp = lambda t: _tptr(t) if t is not None else None
        _check(load().dsht_inv_vgrad_d(self.h, _tptr(w), sl, sc, tp(xy), tp(xy_lon), tp(xy_lat),
                                       outs[0].stride(0), nc, tp(diag), _stream(stream)))
        return xy, xy_lon, xy_lat

    # FP32 grid I/O of the split forms: grids are float32 device tensors (ncols, J, I)
    def fwd_f32_d(self, kind, xy, w, layout="LC", diag=None, stream=None):
        nc = xy.shape[0]
        sl, sc = self._spec_strides(w, nc, layout)
        _check(load().dsht_fwd_f32_d(self.h, kind, _tptr(xy), xy.stride(0), _tptr(w), sl, sc, nc,
                                     _tptr(diag) if diag is not None else None, _stream(stream)))
        return w

    def inv_f32_d(self, kind, w, xy, layout="LC", diag=None, stream=None):
        nc = xy.shape[0]
        sl, sc = self._spec_strides(w, nc, layout)
        _check(load().dsht_inv_f32_d(self.h, kind, _tptr(w), sl, sc, _tptr(xy), xy.stride(0), nc,
                                     _tptr(diag) if diag is not None else None, _stream(stream)))
        return xy

    def inv_vgrad_f32_d(self, w, xy, xy_lon, xy_lat, layout="LC", diag=None, stream=None):
        outs = [t for t in (xy, xy_lon, xy_lat) if t is not None]
        if not outs:
            raise ValueError("inv_vgrad_f32_d: every output is None")
        nc = outs[0].shape[0]
        if any(t.shape[0] != nc or t.stride(0) != outs[0].stride(0) for t in outs):
            raise ValueError("inv_vgrad_f32_d: outputs differ in columns or pitch")
        sl, sc = self._spec_strides(w, nc, layout)
        tp = lambda t: _tptr(t) if t is not None else None
        _check(load().dsht_inv_vgrad_f32_d(self.h, _tptr(w), sl, sc, tp(xy), tp(xy_lon), tp(xy_lat),
                                           outs[0].stride(0), nc, tp(diag), _stream(stream)))
        return xy, xy_lon, xy_lat

    def alpha_fwd_d(self, a, b, w, layout="LC", diag=None, stream=None):
        nc = a.shape[0]
        sl, sc = self._spec_strides(w, nc, layout)
        _check(load().dsht_alpha_fwd_d(self.h, _tptr(a), _tptr(b), a.stride(0), _tptr(w), sl, sc, nc,
                                       _tptr(diag) if diag is not None else None, _stream(stream)))
        return w

    def alpha_inv_d(self, wa, wb, xy, layout="LC", stream=None):
        nc = xy.shape[0]
        sl, sc = self._spec_strides(wa, nc, layout)
        _check(load().dsht_alpha_inv_d(self.h, _tptr(wa), _tptr(wb), sl, sc, _tptr(xy), xy.stride(0), nc,
                                       _stream(stream)))
        return xy

    def uv2vordiv_d(self, ucos, vcos, vor, div, layout="LC", diag=None, stream=None):
        nc = ucos.shape[0]
        sl, sc = self._spec_strides(vor, nc, layout)
        _check(load().dsht_uv2vordiv_d(self.h, _tptr(ucos), _tptr(vcos), ucos.stride(0), _tptr(vor),
                                       _tptr(div), sl, sc, nc,
                                       _tptr(diag) if diag is not None else None, _stream(stream)))
        return vor, div

    def vordiv2uv_d(self, vor, div, u, v, layout="LC", stream=None):
        nc = u.shape[0]
        sl, sc = self._spec_strides(vor, nc, layout)
        _check(load().dsht_vordiv2uv_d(self.h, _tptr(vor), _tptr(div), sl, sc, _tptr(u), _tptr(v),
                                       u.stride(0), nc, _stream(stream)))
        return u, v

    def diag_d(self, d, w, out, layout="LC", stream=None):
        nc = w.shape[1] if layout == "LC" else w.shape[0]
        sl, sc = self._spec_strides(w, nc, layout)
        _check(load().dsht_diag_d(self.h, _tptr(d), _tptr(w), _tptr(out), sl, sc, nc, _stream(stream)))
        return out

    def avrlonlat_d(self, xy, out, stream=None):
        _check(load().dsht_avrlonlat_d(self.h, _tptr(xy), xy.stride(0), _tptr(out), xy.shape[0],
                                       _stream(stream)))
        return out

    def intlonlat_d(self, xy, out, stream=None):
        _check(load().dsht_intlonlat_d(self.h, _tptr(xy), xy.stride(0), _tptr(out), xy.shape[0],
                                       _stream(stream)))
        return out

    # ------------------------------------------------ vertical column diagnostics (device)
    def _col_in(self, cdk, *ts):
        """(km, ldg) of same-shape (km, J, I) float64 device grids (layer 0 on top) and a
        float64 device vector cdk of km layer thicknesses."""
        import torch
        km = ts[0].shape[0]
        if cdk.dtype != torch.float64 or not cdk.is_cuda or cdk.dim() != 1 or cdk.numel() != km \
                or cdk.stride(0) != 1:
            raise ValueError(f"cdk must be a contiguous float64 device vector of {km} layer thicknesses")
        for t in ts:
            if t.dtype != torch.float64 or not t.is_cuda:
                raise TypeError("column diagnostics take float64 device tensors")
            if tuple(t.shape) != (km, self.jm, self.im) or t.stride(2) != 1 or t.stride(1) != self.im:
                raise ValueError(f"grids must be (km, {self.jm}, {self.im}) with contiguous layers")
            if t.stride(0) != ts[0].stride(0):
                raise ValueError("grids differ in layer pitch")
        return km, ts[0].stride(0)

    def _col_out(self, out, nlev, like):
        import torch
        if out is None:
            return torch.empty((nlev, self.jm, self.im), dtype=torch.float64, device=like.device)
        if tuple(out.shape) != (nlev, self.jm, self.im) or out.stride(2) != 1 or out.stride(1) != self.im:
            raise ValueError(f"output must be ({nlev}, {self.jm}, {self.im}) with contiguous levels")
        return out

    def col_omega_d(self, div, h, cdk, out=None, stream=None):
        """HBEDiagnose_OMG (HBEDiagnose_hspm_vfvm_mod.f90:90-143): the (km+1, J, I) half-level
        omega from the layer divergence and thickness factor h; half level k is on top of
        layer k and out[km] = 0."""
        km, ldg = self._col_in(cdk, div, h)
        out = self._col_out(out, km + 1, div)
        _check(load().dsht_col_omega_d(self.h, _tptr(div), _tptr(h), ldg, km, _tptr(cdk), _tptr(out),
                                       out.stride(0), _stream(stream)))
        return out

    def col_hydpres_d(self, dens, h, cdk, grav, out=None, stream=None):
        """HBEDiagnose_HydPres (HBEDiagnose_hspm_vfvm_mod.f90:290-325): (km, J, I) hydrostatic
        pressure at the layer centres from the density anomaly."""
        km, ldg = self._col_in(cdk, dens, h)
        out = self._col_out(out, km, dens)
        _check(load().dsht_col_hydpres_d(self.h, _tptr(dens), _tptr(h), ldg, km, _tptr(cdk), float(grav),
                                         _tptr(out), out.stride(0), _stream(stream)))
        return out

    def col_uvbarot_d(self, u, v, h, cdk, ub=None, vb=None, stream=None):
        """HBEDiagnose_UVBarot (HBEDiagnose_hspm_vfvm_mod.f90:221-255): the thickness-weighted
        layer means of u and v, (J, I) each."""
        import torch
        km, ldg = self._col_in(cdk, u, v, h)
        res = []
        for o in (ub, vb):
            if o is None:
                o = torch.empty((self.jm, self.im), dtype=torch.float64, device=u.device)
            elif tuple(o.shape) != (self.jm, self.im) or not o.is_contiguous():
                raise ValueError(f"outputs must be contiguous ({self.jm}, {self.im}) grids")
            res.append(o)
        _check(load().dsht_col_uvbarot_d(self.h, _tptr(u), _tptr(v), _tptr(h), ldg, km, _tptr(cdk),
                                         _tptr(res[0]), _tptr(res[1]), _stream(stream)))
        return res[0], res[1]

    def col_quick_d(self, omg, trc, cdk, rfdk, linfree=False, out=None, stream=None):
        """calc_ADVFlxZ_QUICK (HBEBaroc_hspm_vfvm_mod.f90:233-317): the (km+1, J, I) vertical
        advective flux at the half levels from the (km+1, J, I) half-level omega and the
        (km, J, I) tracer; rfdk the km+1 reciprocal half-level spacings."""
        import torch
        km, ldg = self._col_in(cdk, trc)
        if tuple(omg.shape) != (km + 1, self.jm, self.im) or omg.stride(2) != 1 or omg.stride(1) != self.im \
                or omg.dtype != torch.float64 or not omg.is_cuda:
            raise ValueError(f"omega must be a float64 device grid ({km + 1}, {self.jm}, {self.im})")
        if rfdk.dtype != torch.float64 or not rfdk.is_cuda or rfdk.dim() != 1 or rfdk.numel() != km + 1 \
                or rfdk.stride(0) != 1:
            raise ValueError(f"rfdk must be a contiguous float64 device vector of {km + 1} values")
        out = self._col_out(out, km + 1, trc)
        _check(load().dsht_col_quick_d(self.h, _tptr(omg), omg.stride(0), _tptr(trc), ldg, km, _tptr(cdk),
                                       _tptr(rfdk), int(bool(linfree)), _tptr(out), out.stride(0),
                                       _stream(stream)))
        return out

    def col_vdiff_d(self, q0, rhs, kv, e3, cdk, fdk, dt, alpha, sfc_bc="D", btm_bc="N", out=None, work=None,
                    stream=None, btm_frict=None):
        """calc_VDiffTRCEq (DOGCM_Phys_hspm_vfvm_mod.f90:752-861): q0 + the implicit vertical
        diffusion increment, (km, J, I); fdk the km+1 half-level spacings, sfc_bc/btm_bc 'D' or
        'N'.  work: optional float64 device scratch of >= 3*km*J*I values.  With btm_frict (a
        (J, I) device grid of xy_BtmFrictCoef) the momentum form calc_VDiffMOMEq (:865-975)."""
        import torch
        km, ldg = self._col_in(cdk, q0, rhs, kv, e3)
        if fdk.dtype != torch.float64 or not fdk.is_cuda or fdk.dim() != 1 or fdk.numel() != km + 1 \
                or fdk.stride(0) != 1:
            raise ValueError(f"fdk must be a contiguous float64 device vector of {km + 1} values")
        out = self._col_out(out, km, q0)
        need = 3 * km * self.jm * self.im
        if work is None:
            work = torch.empty(need, dtype=torch.float64, device=q0.device)
            if stream is not None:
                # the scratch is used on the caller's stream: keep the caching allocator from
                # handing its block out again before that stream's work is done
                work.record_stream(torch.cuda.ExternalStream(stream))
        elif work.dtype != torch.float64 or not work.is_cuda or not work.is_contiguous():
            raise ValueError("work must be a contiguous float64 device tensor")
        if btm_frict is None:
            _check(load().dsht_col_vdiff_d(self.h, _tptr(q0), _tptr(rhs), _tptr(kv), _tptr(e3), ldg, km, _tptr(cdk),
                                           _tptr(fdk), float(dt), float(alpha), ord(sfc_bc), ord(btm_bc), _tptr(out),
                                           out.stride(0), _tptr(work), work.numel(), _stream(stream)))
            return out
        if tuple(btm_frict.shape) != (self.jm, self.im) or not btm_frict.is_contiguous() \
                or btm_frict.dtype != torch.float64 or not btm_frict.is_cuda:
            raise ValueError(f"btm_frict must be a contiguous float64 device grid ({self.jm}, {self.im})")
        _check(load().dsht_col_vdiff_mom_d(self.h, _tptr(q0), _tptr(rhs), _tptr(kv), _tptr(e3), ldg, km, _tptr(cdk),
                                           _tptr(fdk), float(dt), float(alpha), ord(sfc_bc), ord(btm_bc),
                                           _tptr(btm_frict), _tptr(out), out.stride(0), _tptr(work), work.numel(),
                                           _stream(stream)))
        return out

    # column kernels on host arrays (numpy (km, J, I) fields, layer 0 on top)
    def _col_host(self, *arrs):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in arrs]
        km = a[0].shape[0]
        for x in a:
            if x.shape != (km, self.jm, self.im):
                raise ValueError(f"fields must be (km, {self.jm}, {self.im})")
        return a, km

    @staticmethod
    def _col_vec(name, v, n):
        """A host layer vector of exactly n values (the C entry points copy n values; a
        halo-sized KA vector would be read shifted)."""
        v = _f64(np.asarray(v, dtype=np.float64).reshape(-1))
        if v.size != n:
            raise ValueError(f"{name} has {v.size} values, expected {n} (pass the KS:KE section, not a halo array)")
        return v

    def xyr_OMG(self, xyz_Div, xyz_H, z_CDK):
        """HBEDiagnose_OMG on host arrays: (km+1, J, I) half-level omega."""
        (div, h), km = self._col_host(xyz_Div, xyz_H)
        cdk = self._col_vec("z_CDK", z_CDK, km)
        out = np.empty((km + 1, self.jm, self.im))
        _check(load().dsht_col_omega(self.h, _ptr(div), _ptr(h), km, _ptr(cdk), _ptr(out)))
        return out

    def xyz_HydPres(self, xyz_DensEdd, xyz_H, z_CDK, grav):
        (dens, h), km = self._col_host(xyz_DensEdd, xyz_H)
        cdk = self._col_vec("z_CDK", z_CDK, km)
        out = np.empty((km, self.jm, self.im))
        _check(load().dsht_col_hydpres(self.h, _ptr(dens), _ptr(h), km, _ptr(cdk), float(grav), _ptr(out)))
        return out

    def xy_UVBarot(self, xyz_U, xyz_V, xyz_H, z_CDK):
        (u, v, h), km = self._col_host(xyz_U, xyz_V, xyz_H)
        cdk = self._col_vec("z_CDK", z_CDK, km)
        ub, vb = np.empty((self.jm, self.im)), np.empty((self.jm, self.im))
        _check(load().dsht_col_uvbarot(self.h, _ptr(u), _ptr(v), _ptr(h), km, _ptr(cdk), _ptr(ub), _ptr(vb)))
        return ub, vb

    def xyr_ADVFlxZ_QUICK(self, xyr_OMG, xyz_TRC, z_CDK, z_RFDK, linfree=False):
        (trc,), km = self._col_host(xyz_TRC)
        omg = np.ascontiguousarray(xyr_OMG, dtype=np.float64)
        if omg.shape != (km + 1, self.jm, self.im):
            raise ValueError("omega must have km+1 half levels")
        cdk = self._col_vec("z_CDK", z_CDK, km)
        rfdk = self._col_vec("z_RFDK", z_RFDK, km + 1)
        out = np.empty((km + 1, self.jm, self.im))
        _check(load().dsht_col_quick(self.h, _ptr(omg), _ptr(trc), km, _ptr(cdk), _ptr(rfdk), int(bool(linfree)),
                                     _ptr(out)))
        return out

    def xyz_VDiff(self, xyz_q0, xyz_RHS, xyz_Kv, xyz_H, z_CDK, z_FDK, dt, alpha, sfc_bc="D", btm_bc="N",
                  xy_BtmFrictCoef=None):
        """calc_VDiffTRCEq on host arrays, or calc_VDiffMOMEq with xy_BtmFrictCoef."""
        (q0, rhs, kv, e3), km = self._col_host(xyz_q0, xyz_RHS, xyz_Kv, xyz_H)
        cdk = self._col_vec("z_CDK", z_CDK, km)
        fdk = self._col_vec("z_FDK", z_FDK, km + 1)
        bf = None if xy_BtmFrictCoef is None else np.ascontiguousarray(xy_BtmFrictCoef, dtype=np.float64)
        if bf is not None and bf.shape != (self.jm, self.im):
            raise ValueError(f"xy_BtmFrictCoef must be ({self.jm}, {self.im})")
        out = np.empty((km, self.jm, self.im))
        _check(load().dsht_col_vdiff(self.h, _ptr(q0), _ptr(rhs), _ptr(kv), _ptr(e3), km, _ptr(cdk), _ptr(fdk),
                                     float(dt), float(alpha), ord(sfc_bc), ord(btm_bc),
                                     _ptr(bf) if bf is not None else None, _ptr(out)))
        return out

    # energy and enstrophy spectra of streamfunctions (nma_/na_EnergyFromStreamfunc_wa,
    # nma_/na_EnstrophyFromStreamfunc_wa, SpmlUtil_mod.f90:131-132)
    def spec_energy(self, wa, kind=SPEC_ENERGY):
        """(nma, na) of host spectra wa (ncols, L): nma (ncols, 2N+1, N+1) with nma[c, m+N, n]
        (zero where |m| > n), na (ncols, N+1) their sums over m; unit sphere."""
        w = _f64(np.asarray(wa, dtype=np.float64).reshape(-1, self.lm))
        nc = w.shape[0]
        nma = np.zeros((nc, 2 * self.nmax + 1, self.nmax + 1))
        na = np.zeros((nc, self.nmax + 1))
        _check(load().dsht_spec_energy(self.h, int(kind), _ptr(w), nc, _ptr(nma), _ptr(na)))
        return nma, na

    def nma_EnergyFromStreamfunc_wa(self, wa):
        return self.spec_energy(wa, SPEC_ENERGY)[0]

    def na_EnergyFromStreamfunc_wa(self, wa):
        return self.spec_energy(wa, SPEC_ENERGY)[1]

    def nma_EnstrophyFromStreamfunc_wa(self, wa):
        return self.spec_energy(wa, SPEC_ENSTROPHY)[0]

    def na_EnstrophyFromStreamfunc_wa(self, wa):
        return self.spec_energy(wa, SPEC_ENSTROPHY)[1]

    # Redi/GM isopycnal mixing (LPhys_RediGM_hspm_vfvm_AddMixingTerm,
    # LPhys_RediGM_hspm_vfvm_mod.f90:218-372)
    def _redigm_in(self, ptemp, salt, h, z, topo, cdk, rfdk):
        import torch
        km, ldg = self._col_in(cdk, ptemp, salt, h, z)
        if km < 2:
            raise ValueError("the Redi/GM mixing needs km >= 2")
        if tuple(topo.shape) != (self.jm, self.im) or not topo.is_contiguous() or topo.dtype != torch.float64 \
                or not topo.is_cuda:
            raise ValueError(f"topo must be a contiguous float64 device grid ({self.jm}, {self.im})")
        if rfdk.dtype != torch.float64 or not rfdk.is_cuda or rfdk.dim() != 1 or rfdk.numel() != km + 1 \
                or rfdk.stride(0) != 1:
            raise ValueError(f"rfdk must be a contiguous float64 device vector of {km + 1} values")
        return km, ldg

    def col_redigm_d(self, ptemp, salt, h, z, topo, cdk, rfdk, grad_lon, grad_lat, cfg=None, stream=None):
        """The grid-space part of the Redi/GM mixing: from (km, J, I) ptemp [K], salt, h, z and
        the (2km, J, I) GradLon / GradLat grids of w_xy(ptemp) then w_xy(salt) (not divided by
        the radius), returns (flx_lon, flx_lat, vdiv), each (2km, J, I): the AlphaOptr inputs
        (FLonRedi + FLonGM) cos(lat), (FLatRedi + FLatGM) cos(lat) and the vertical flux
        divergence of each tracer layer."""
        import torch
        km, ldg = self._redigm_in(ptemp, salt, h, z, topo, cdk, rfdk)
        for g in (grad_lon, grad_lat):
            if tuple(g.shape) != (2 * km, self.jm, self.im) or g.stride(0) != ldg or g.stride(2) != 1 \
                    or g.stride(1) != self.im or g.dtype != torch.float64 or not g.is_cuda:
                raise ValueError(f"gradients must be float64 device grids (2km, {self.jm}, {self.im}) with the "
                                 "layer pitch of the fields")
        outs = [torch.empty((2 * km, self.jm, self.im), dtype=torch.float64, device=ptemp.device) for _ in range(3)]
        cfg = redigm_config() if cfg is None else cfg
        _check(load().dsht_col_redigm_d(self.h, ctypes.byref(cfg), _tptr(ptemp), _tptr(salt), _tptr(h), _tptr(z),
                                        ldg, km, _tptr(topo), _tptr(cdk), _tptr(rfdk), _tptr(grad_lon),
                                        _tptr(grad_lat), _tptr(outs[0]), _tptr(outs[1]), _tptr(outs[2]),
                                        outs[0].stride(0), _stream(stream)))
        return tuple(outs)

    def redigm_d(self, ptemp, salt, h, z, topo, cdk, rfdk, ptemp_rhs, salt_rhs, cfg=None, work=None, stream=None):
        """LPhys_RediGM_hspm_vfvm_AddMixingTerm on device grids: adds the Redi/GM tendency to
        the (km, J, I) ptemp_rhs and salt_rhs in place and returns them.  work: optional float64
        device scratch of >= dsht_redigm_work_size doubles."""
        import torch
        km, ldg = self._redigm_in(ptemp, salt, h, z, topo, cdk, rfdk)
        ldo = ptemp_rhs.stride(0)
        for r in (ptemp_rhs, salt_rhs):
            if tuple(r.shape) != (km, self.jm, self.im) or r.stride(0) != ldo or r.stride(2) != 1 \
                    or r.stride(1) != self.im or r.dtype != torch.float64 or not r.is_cuda:
                raise ValueError(f"tendencies must be float64 device grids (km, {self.jm}, {self.im}) of one pitch")
        need = load().dsht_redigm_work_size(self.h, km)
        if need < 0:
            _check(-1)
        if work is None:
            work = torch.empty(need, dtype=torch.float64, device=ptemp.device)
            if stream is not None:
                work.record_stream(torch.cuda.ExternalStream(stream))
        elif work.dtype != torch.float64 or not work.is_cuda or not work.is_contiguous():
            raise ValueError("work must be a contiguous float64 device tensor")
        cfg = redigm_config() if cfg is None else cfg
        _check(load().dsht_redigm_d(self.h, ctypes.byref(cfg), _tptr(ptemp), _tptr(salt), _tptr(h), _tptr(z), ldg,
                                    km, _tptr(topo), _tptr(cdk), _tptr(rfdk), _tptr(ptemp_rhs), _tptr(salt_rhs),
                                    ldo, _tptr(work), work.numel(), _stream(stream)))
        return ptemp_rhs, salt_rhs

    def RediGM_AddMixingTerm(self, xyz_PTemp_RHS, xyz_Salt_RHS, xyz_PTemp, xyz_Salt, xyz_H, xyz_Z, xy_Topo,
                             z_CDK, z_RFDK, cfg=None):
        """LPhys_RediGM_hspm_vfvm_AddMixingTerm on host arrays ((km, J, I) fields, layer 0 on
        top; z_RFDK the km+1 half levels): returns the updated (PTemp_RHS, Salt_RHS)."""
        (pr, sr, pt, sa, h, z), km = self._col_host(xyz_PTemp_RHS, xyz_Salt_RHS, xyz_PTemp, xyz_Salt, xyz_H, xyz_Z)
        topo = np.ascontiguousarray(xy_Topo, dtype=np.float64)
        if topo.shape != (self.jm, self.im):
            raise ValueError(f"xy_Topo must be ({self.jm}, {self.im})")
        cdk = self._col_vec("z_CDK", z_CDK, km)
        rfdk = self._col_vec("z_RFDK", z_RFDK, km + 1)
        pr, sr = pr.copy(), sr.copy()
        cfg = redigm_config() if cfg is None else cfg
        _check(load().dsht_redigm(self.h, ctypes.byref(cfg), _ptr(pt), _ptr(sa), _ptr(h), _ptr(z), km, _ptr(topo),
                                  _ptr(cdk), _ptr(rfdk), _ptr(pr), _ptr(sr)))
        return pr, sr

    def set_tuning(self, key, value):
        _check(load().dsht_set_tuning(self.h, int(key), int(value)))

    def get_tuning(self, key):
        return load().dsht_get_tuning(self.h, int(key))

    def coef_d(self, which, params, out, stream=None):
        p = np.asarray(params, dtype=np.float64).reshape(-1)
        pp = _f64(p) if p.size else np.zeros(1)
        _check(load().dsht_coef_d(self.h, int(which), _ptr(pp), int(p.size), _tptr(out), _stream(stream)))
        return out


class Batch:
    """Independent instances of uneven truncation/layer count advanced together
    (dsht_batch_*).  ``add(plan, km)`` returns an instance id; after ``finalize()`` each
    truncation group exposes torch views of its device buffers via ``group_views``."""

    NPROG, NTEND, NINV = 4, 6, 3

    def __init__(self):
        self.h = ctypes.c_void_p()
        _check(load().dsht_batch_create(ctypes.byref(self.h)))
        self.plans = []

    def add(self, plan, km):
        i = ctypes.c_int()
        _check(load().dsht_batch_add(self.h, plan.h, int(km), ctypes.byref(i)))
        self.plans.append((plan, km))
        return i.value

    def finalize(self):
        _check(load().dsht_batch_finalize(self.h))

    def info(self, inst):
        g, c0, gc, by = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_long()
        _check(load().dsht_batch_info(self.h, int(inst), ctypes.byref(g), ctypes.byref(c0), ctypes.byref(gc),
                                      ctypes.byref(by)))
        return g.value, c0.value, gc.value, by.value

    def group_views(self, group, plan, cols):
        """torch tensors over the group's device buffers (no copies)."""
        import torch
        ptrs = [ctypes.c_void_p() for _ in range(4)]
        _check(load().dsht_batch_buffers(self.h, int(group), *[ctypes.byref(p) for p in ptrs]))
        L, J, I = plan.lm, plan.jm, plan.im
        shapes = [(L, self.NPROG * cols), (self.NINV, self.NPROG * cols, J, I), (self.NTEND * cols, J, I),
                  (L, self.NTEND * cols)]
        return [_tensor_from_ptr(p.value, s) for p, s in zip(ptrs, shapes)]

    def ngroups(self):
        return load().dsht_batch_ngroups(self.h)

    def step(self, stream=None):
        _check(load().dsht_batch_step(self.h, _stream(stream)))

    def step_sequential(self, stream=None):
        _check(load().dsht_batch_step_sequential(self.h, _stream(stream)))

    def capture(self, stream):
        _check(load().dsht_batch_capture(self.h, ctypes.c_void_p(stream)))

    def launch(self, stream=None):
        _check(load().dsht_batch_launch(self.h, _stream(stream)))

    def close(self):
        if self.h and self.h.value:
            load().dsht_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _tensor_from_ptr(ptr, shape):
    """A float64 CUDA tensor viewing device memory owned by the library."""
    import torch

    n = 1
    for s in shape:
        n *= s

    class _Holder:
        pass

    h = _Holder()
    h.__cuda_array_interface__ = {"shape": tuple(shape), "typestr": "<f8", "data": (ptr, False), "version": 2,
                                  "strides": None}
    return torch.as_tensor(h, device="cuda")


def redigm_config(**kw):
    """dsht_redigm_default_config with the given fields replaced (tapering names 'DM95',
    'GKW91', 'LDD95' or the TAPER_* numbers)."""
    c = RediGMConfig()
    _check(load().dsht_redigm_default_config(ctypes.byref(c)))
    names = {"GKW91": TAPER_GKW91, "DM95": TAPER_DM95, "LDD95": TAPER_LDD95}
    for k, v in kw.items():
        if not hasattr(c, k):
            raise AttributeError(f"dsht_redigm_config has no field {k}")
        setattr(c, k, names[v] if isinstance(v, str) else v)
    return c


def lfam3_config(**kw):
    """dsht_lfam3_default_config with the given fields replaced (boundary types as 'D'/'N')."""
    c = LfAm3Config()
    _check(load().dsht_lfam3_default_config(ctypes.byref(c)))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise AttributeError(f"dsht_lfam3_config has no field {k}")
        setattr(c, k, ord(v) if k.endswith("_bc") and isinstance(v, str) else v)
    return c


class LfAm3:
    """A pool of up to max_runs independent LF-AM3 runs of one plan and km layers
    (dsht_lfam3_*, the device form of DOGCM_TInt_LFAM3_advance_Dyn).  State views
    (torch, float64 device): U, V (max_runs*km, J, I); TRC (2*max_runs*km, J, I) with run r's
    temperature at rows (2r)*km.. and salinity at (2r+1)*km..; SFCPRES (max_runs, J, I)."""

    def __init__(self, plan, cfg):
        self.plan, self.cfg = plan, cfg
        self.h = ctypes.c_void_p()
        _check(load().dsht_lfam3_create(plan.h, ctypes.byref(cfg), ctypes.byref(self.h)))
        self.km, self.max_runs = cfg.km, cfg.max_runs

    def state(self, field, level):
        ptr, n = ctypes.c_void_p(), ctypes.c_long()
        _check(load().dsht_lfam3_state(self.h, int(field), int(level), ctypes.byref(ptr), ctypes.byref(n)))
        return _tensor_from_ptr(ptr.value, (n.value, self.plan.jm, self.plan.im))

    def run_view(self, field, level, run):
        """Views of one run's slice: U/V (km, J, I); TRC (2, km, J, I); SFCPRES (J, I)."""
        t = self.state(field, level)
        k = self.km
        if field == LF_TRC:
            return t[2 * run * k:(2 * run + 2) * k].view(2, k, self.plan.jm, self.plan.im)
        if field == LF_SFCPRES:
            return t[run]
        return t[run * k:(run + 1) * k]

    def set_rotation(self, run, omega):
        _check(load().dsht_lfam3_set_rotation(self.h, int(run), float(omega)))

    def step(self, nruns, stream=None):
        _check(load().dsht_lfam3_step(self.h, int(nruns), _stream(stream)))

    def copy_run(self, dst, src, stream=None):
        _check(load().dsht_lfam3_copy_run(self.h, int(dst), int(src), _stream(stream)))

    def units(self, nruns):
        f, i = ctypes.c_long(), ctypes.c_long()
        _check(load().dsht_lfam3_units(self.h, int(nruns), ctypes.byref(f), ctypes.byref(i)))
        return f.value, i.value

    def close(self):
        if self.h and self.h.value:
            load().dsht_lfam3_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
