"""CPU-side checks of the C ABI: the library loads, exports every entry point that
include/dsht.h declares, and rejects bad arguments before touching a device."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "dsht.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dsht_\w+)\s*\(", text)))


def test_header_matches_wrapper_list():
    import dsht
    assert sorted(dsht.EXPORTS) == _declared()


def test_library_exports_every_declared_symbol():
    import dsht
    lib = dsht.load()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_create_rejects_bad_sizes_without_a_device():
    import dsht
    lib = dsht.load()
    h = ctypes.c_void_p()
    for args in [(42, 64, 64), (42, 128, 63), (42, 127, 64), (-1, 16, 8), (42, 128, 40)]:
        rc = lib.dsht_create(args[0], args[1], args[2], 1, 1.0, ctypes.byref(h))
        assert rc == 1, args  # DSHT_EARG
        assert lib.dsht_last_error()
    assert lib.dsht_create(21, 64, 32, 1, -1.0, ctypes.byref(h)) == 1


def test_null_plan_is_a_state_error():
    import dsht
    lib = dsht.load()
    assert lib.dsht_w_xy(None, 0, None, None, 1) == 4  # DSHT_ESTATE
    assert lib.dsht_fwd_d(None, 0, None, 1, None, 1, 1, 1, None, None) == 4
    assert lib.dsht_l_nm(None, 1, 0) == 0


def test_fortran_shim_compiles():
    """The ISO_C_BINDING SpmlUtil_mod module and its test program build with amdflang."""
    fc = "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        pytest.skip("amdflang not installed")
    import subprocess
    import tempfile
    src = os.path.join(ROOT, "dennou-ogcm_amd", "fortran", "SpmlUtil_mod.f90")
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([fc, "-c", src, "-module-dir", d, "-o", os.path.join(d, "m.o")],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("kmax,tmax", [(2, 2), (8, 8), (20, 20), (60, 60), (60, 40), (12, 7)])
def test_sig_grid_clenshaw_curtis(kmax, tmax):
    """g_Sig / g_Sig_WEIGHT (host-only, no device): Chebyshev-Lobatto levels of [-1, 0] from
    the top, sig_i = -1/2 + cos(pi i/kmax)/2, and weights that integrate every polynomial of
    degree <= tmax (odd part: every degree up to kmax by symmetry) exactly over [-1, 0]."""
    import numpy as np
    import dsht
    sig, w = dsht.sig_grid(kmax, tmax)
    i = np.arange(kmax + 1)
    assert np.abs(sig - (-0.5 + 0.5 * np.cos(np.pi * i / kmax))).max() < 1e-15
    assert sig[0] == 0.0 and sig[-1] == -1.0
    assert np.abs(w - w[::-1]).max() < 1e-15  # symmetric about the interval's midpoint
    for deg in range(tmax + 1):
        exact = (0.0 - (-1.0) ** (deg + 1)) / (deg + 1)  # int_{-1}^{0} x^deg dx
        assert abs(np.dot(w, sig ** deg) - exact) < 1e-14, (deg, np.dot(w, sig ** deg), exact)
    if tmax == kmax:
        # a planted mistake (ends not halved) breaks the constant's integral
        bad = w.copy()
        bad[0] *= 2
        bad[-1] *= 2
        assert abs(bad.sum() - 1.0) > 1e-8


def test_sig_grid_rejects_what_at_initial_rejects():
    import dsht
    for args in [(0, 1), (5, 4), (8, 10), (8, 0)]:
        with pytest.raises(dsht.DshtError):
            dsht.sig_grid(*args)


def test_host_column_wrappers_check_vector_lengths():
    """The host-array column wrappers refuse layer vectors of the wrong length (a short one
    would be over-read by the C entry point, a halo-sized KA one read shifted) before any
    library call, so this runs without a device."""
    import numpy as np
    import dsht
    p = object.__new__(dsht.Plan)  # argument checks only: no plan handle is needed
    p.im, p.jm = 8, 4
    km = 3
    f = np.zeros((km, 4, 8))
    half = np.zeros((km + 1, 4, 8))
    good, ka = np.ones(km), np.ones(km + 2)
    with pytest.raises(ValueError, match="z_CDK"):
        p.xyr_OMG(f, f, ka)
    with pytest.raises(ValueError, match="z_CDK"):
        p.xyz_HydPres(f, f, good[:2], 9.8)
    with pytest.raises(ValueError, match="z_CDK"):
        p.xy_UVBarot(f, f, f, ka)
    with pytest.raises(ValueError, match="z_RFDK"):
        p.xyr_ADVFlxZ_QUICK(half, f, good, np.ones(km))
    with pytest.raises(ValueError, match="z_FDK"):
        p.xyz_VDiff(f, f, f, f, good, np.ones(km + 2), 1.0, 1.0)
    with pytest.raises(ValueError, match="z_CDK"):
        p.xyz_VDiff(f, f, f, f, ka, np.ones(km + 1), 1.0, 1.0)


def test_lfam3_config_layout_and_defaults():
    """dsht_lfam3_default_config fills the C struct the ctypes mirror describes (host-only):
    the defaults of APEOGCircExp.nml and EOS_Linear_Init land in the right fields."""
    import dsht
    c = dsht.lfam3_config()
    assert (c.km, c.max_runs, c.sub_cycles, c.numdiff_ord) == (20, 1, 1, 4)
    assert (c.alpha, c.lam, c.grav, c.ref_dens) == (0.51, 0.5, 9.79764, 1027.0)
    assert (c.hvisc, c.vvisc, c.vdiff, c.numdiff) == (3e5, 1e-3, 3e-5, 1e12)
    assert (chr(c.mom_sfc_bc), chr(c.mom_btm_bc), chr(c.trc_sfc_bc), chr(c.trc_btm_bc)) == ("N", "D", "N", "N")
    assert (c.ref_temp, c.ref_salt, c.beta_t, c.beta_s, c.cp0, c.cs0) == (283.0, 35.0, 1.67e-4, 0.78e-3, 3986.0, 1490.0)
    c2 = dsht.lfam3_config(km=60, dt=900.0, mom_btm_bc="N")
    assert c2.km == 60 and c2.dt == 900.0 and c2.mom_btm_bc == ord("N")
