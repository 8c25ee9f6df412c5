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
