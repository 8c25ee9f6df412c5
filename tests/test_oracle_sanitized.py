"""The CPU oracle (oracle/dsht_oracle.c) under AddressSanitizer and UndefinedBehaviorSanitizer:
built into a temporary directory with -fsanitize=address,undefined (undefined behaviour is
fatal), loaded in a fresh Python process with the sanitizer runtimes preloaded, and driven
through every transform kind, the operators, the means and the column loops at T5, T21 and
T42.  Any out-of-bounds access, use after free or undefined operation ends that process
with an error report."""
import os
import subprocess
import sys
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

DRIVER = r'''
import sys
sys.path.insert(0, {oracle_dir!r}); sys.path.insert(0, {tests_dir!r})
import numpy as np
import oracle as O
O.use_library({lib!r})
from conftest import red_spectrum
for (N, I, J) in [(5, 16, 8), (21, 64, 32), (42, 128, 64)]:
    o = O.Oracle(N, I, J, 6.371e6)
    w = red_spectrum(N, 3, seed=N)
    for k in (O.VAL, O.GRADLAMBDA, O.GRADMU, O.GRADLON, O.GRADLAT):
        xy = o.inv(w, k)
    for k in (O.VAL, O.DIVLAMBDA, O.DIVMU):
        o.fwd(xy, k)
    o.alpha_fwd(xy, xy); o.alpha_inv(w, w); o.uv2vordiv(xy, xy); o.vordiv2uv(w, w)
    o.intlonlat(xy); o.avrlonlat(xy); o.ya_intlon(xy)
    o.interpolate(w[0], np.array([0.1, 2.0]), np.array([0.3, -1.0]))
    km = 7
    rng = np.random.default_rng(0)
    f = lambda: rng.standard_normal((km, J, I))
    cdk = np.full(km, 1.0 / km)
    omg = O.col_omega(f() * 1e-6, 1 + 0 * f(), cdk)
    O.col_hydpres(f(), 1 + 0 * f(), cdk, 9.8)
    O.col_uvbarot(f(), f(), 1 + 0 * f(), cdk)
    O.col_quick(omg, f(), cdk, np.full(km + 1, float(km)))
    O.col_vdiff(f(), f() * 1e-5, np.abs(f()) * 1e-3, 100 + np.abs(f()), cdk, np.full(km + 1, 1.0 / km),
                3600.0, 1.0, "D", "N")
print("sanitized oracle: ok")
'''


def _runtime(name):
    p = subprocess.run(["gcc", "-print-file-name=" + name], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


def test_oracle_under_asan_and_ubsan():
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("gcc sanitizer runtimes not installed")
    with tempfile.TemporaryDirectory(prefix="dsht_oracle_san_") as tmp:
        lib = os.path.join(tmp, "liborc_san.so")
        subprocess.check_call(["gcc", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                               "-fno-omit-frame-pointer", "-fPIC", "-std=gnu11", "-shared", "-o", lib,
                               os.path.join(ROOT, "oracle", "dsht_oracle.c"), "-lm"])
        env = dict(os.environ, LD_PRELOAD=asan + ":" + ubsan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
                   UBSAN_OPTIONS="print_stacktrace=1")
        code = DRIVER.format(oracle_dir=os.path.join(ROOT, "oracle"), tests_dir=HERE, lib=lib)
        r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=600)
        assert r.returncode == 0 and "sanitized oracle: ok" in r.stdout, (r.stdout[-2000:], r.stderr[-4000:])
