"""Build recipes: the HIP solver library (gfx950), the CPU oracle used as the
checker, and the Fortran ``linsol`` shim.

Everything is built in-tree so the shared objects travel with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "nga2_amd")
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
INCLUDE = os.path.join(ROOT, "include")

HIP_SOURCES = ["ngap_level.hip", "ngap_smooth.hip", "ngap_api.hip", "ngap_proj.hip"]
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
    # packed FP32 instructions stay off: some of their forms have returned wrong results without an error
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")
    return r.stdout


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build_solver_lib():
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libngapsolve.so")
    tmp = out + ".tmp"
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    _run([hipcc()] + HIPCC_FLAGS + ["-I" + INCLUDE, "-I" + CSRC] + srcs + ["-o", tmp])
    os.replace(tmp, out)
    return out


NET_SOURCES = ["ngap_net.hip", "ngap_netmodel.hip"]


def build_net_lib():
    """libngapnet.so: the U-Net's BF16 MFMA convolution and training kernels."""
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libngapnet.so")
    tmp = out + ".tmp"
    srcs = [os.path.join(CSRC, s) for s in NET_SOURCES]
    _run([hipcc()] + HIPCC_FLAGS + ["-I" + INCLUDE, "-I" + CSRC] + srcs + ["-o", tmp])
    os.replace(tmp, out)
    return out


FFT_SOURCES = ["ngap_fft.hip"]


def build_fft_lib():
    """libngapfft.so: fourier3d on the GPU (rocFFT transforms through hipFFT)."""
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libngapfft.so")
    tmp = out + ".tmp"
    srcs = [os.path.join(CSRC, s) for s in FFT_SOURCES]
    _run([hipcc()] + HIPCC_FLAGS + ["-I" + INCLUDE, "-I" + CSRC] + srcs + ["-lhipfft", "-o", tmp])
    os.replace(tmp, out)
    return out


def build_oracle():
    _run(["make", "-B", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liborc.so")


def flang():
    for c in (os.environ.get("FC_AMD"), "/opt/rocm/bin/amdflang", shutil.which("amdflang"), shutil.which("flang")):
        if c and os.path.exists(c):
            return c
    return None


FORTRAN_DIR = os.path.join(PKG, "fortran")
FORTRAN_SOURCES = ["ngap_c.f90", "ngap_str_class.f90", "ngap_fourier3d_class.f90", "ngap_uns_class.f90"]
FORTRAN_TEST_SOURCES = ["standalone/precision.f90", "standalone/config_class.f90", "standalone/linsol_class.f90"]


def build_fortran():
    """Compile the Fortran shim (modules) against the standalone linsol/config
    restatement and link the demo driver. Returns the driver path, or None when
    no Fortran compiler is present."""
    fc = flang()
    if fc is None or not os.path.exists(os.path.join(FORTRAN_DIR, "ngap_c.f90")):
        return None
    bdir = os.path.join(LIBDIR, "fortran")
    os.makedirs(bdir, exist_ok=True)
    srcs = [os.path.join(FORTRAN_DIR, s) for s in FORTRAN_TEST_SOURCES + FORTRAN_SOURCES]
    libs = ["-L" + LIBDIR, "-lngapsolve", "-lngapfft", "-lngapuns", "-Wl,-rpath," + LIBDIR]
    exes = []
    for drv, name in (("demo_driver.f90", "ngap_demo"), ("fft_demo.f90", "ngap_fft_demo"),
                      ("uns_demo.f90", "ngap_uns_demo")):
        exe = os.path.join(bdir, name)
        _run([fc, "-O2", "-J", bdir] + srcs + [os.path.join(FORTRAN_DIR, drv)] + libs + ["-o", exe + ".tmp"], cwd=bdir)
        os.replace(exe + ".tmp", exe)
        exes.append(exe)
    return exes[0]

UNS_SOURCES = ["ngap_amg.cpp", "ngap_uns.hip"]


def build_uns_lib():
    """libngapuns.so: hypre_uns (CSR + algebraic multigrid) on the GPU."""
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libngapuns.so")
    tmp = out + ".tmp"
    srcs = [os.path.join(CSRC, s) for s in UNS_SOURCES]
    _run([hipcc()] + HIPCC_FLAGS + ["-I" + INCLUDE, "-I" + CSRC] + srcs + ["-o", tmp])
    os.replace(tmp, out)
    return out


def build_all(fortran=True):
    out = {"solver": build_solver_lib(), "net": build_net_lib(), "fft": build_fft_lib(), "uns": build_uns_lib(),
           "oracle": build_oracle()}
    if fortran:
        out["fortran"] = build_fortran()
    return out
