"""Builds libdsht.so (HIP, gfx950) in-tree and the Fortran SpmlUtil_mod shim.

Usage: python dennou-ogcm_amd/build.py   (also called by __graft_entry__.build())
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdsht.so")
FDIR = os.path.join(HERE, "fortran")

SOURCES = ["api.hip", "tables.hip", "fft.hip", "fft_reg.hip", "legendre_f64.hip", "legendre_f64d.hip", "legendre_mixed.hip", "legendre_mx2.hip", "ops.hip", "column.hip", "batch.hip", "lfam3.hip", "redigm.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIPFLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _compile(src):
    out = os.path.join(OBJ, src.replace(".hip", ".o"))
    cmd = [HIPCC] + HIPFLAGS + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(l for l in (r.stderr or "").splitlines()
                     if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{errs}")
    return out


def build_lib(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.run(cmd, check=True, capture_output=True)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_fortran(verbose=True):
    """Compiles the ISO_C_BINDING SpmlUtil_mod shim and its test program with amdflang."""
    fc = os.environ.get("FC", "/opt/rocm/bin/amdflang")
    if not os.path.exists(fc):
        if verbose:
            print("amdflang not found; Fortran shim not built")
        return None
    os.makedirs(OBJ, exist_ok=True)
    mod = os.path.join(FDIR, "SpmlUtil_mod.f90")
    test = os.path.join(FDIR, "spml_dropin_test.f90")
    exe = os.path.join(HERE, "spml_dropin_test")
    # -fopenmp: the test program drives the shim from concurrent OpenMP threads, as the
    # reference's !$omp parallel do over layers does
    subprocess.run([fc, "-O2", "-fopenmp", "-c", mod, "-module-dir", OBJ, "-o", os.path.join(OBJ, "SpmlUtil_mod.o")],
                   check=True)
    subprocess.run([fc, "-O2", "-fopenmp", "-c", test, "-module-dir", OBJ, "-o",
                    os.path.join(OBJ, "spml_dropin_test.o")], check=True)
    subprocess.run([fc, "-O2", "-fopenmp", "-o", exe, os.path.join(OBJ, "spml_dropin_test.o"),
                    os.path.join(OBJ, "SpmlUtil_mod.o"), "-L" + HERE, "-ldsht",
                    "-Wl,-rpath," + HERE], check=True)
    if verbose:
        print("built", exe)
    return exe


if __name__ == "__main__":
    build_lib()
    if os.path.exists(os.path.join(FDIR, "SpmlUtil_mod.f90")):
        build_fortran()
    sys.exit(0)
