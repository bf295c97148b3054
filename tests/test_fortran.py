"""The Fortran drop-in (nga2_amd/fortran/ngap_str_class.f90): it compiles with
amdflang against the linsol contract, links libngapsolve.so, and its
self-checking driver solves manufactured problems on the GPU in FP64 and
mixed precision (periodic box and channel with walls)."""
import os
import subprocess

import pytest

from conftest import ROOT

DEMO = os.path.join(ROOT, "nga2_amd", "lib", "fortran", "ngap_demo")


def _need_demo():
    if not os.path.exists(DEMO):
        from nga2_amd.build import build_fortran, flang

        if flang() is None:
            pytest.skip("no Fortran compiler in this image")
        build_fortran()
    assert os.path.exists(DEMO)


def test_fortran_shim_builds_and_fails_cleanly_without_gpu():
    """Without a usable device the shim reports the library's error through
    die (error stop 2), as the reference does for misuse."""
    _need_demo()
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present; the GPU test covers the driver")
    r = subprocess.run([DEMO, "8"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 2
    assert "[ngap_str init] HIP runtime error" in r.stdout


@pytest.mark.gpu
def test_fortran_driver_on_gpu():
    _need_demo()
    print("\n[fortran demo 32^3]", flush=True)
    r = subprocess.run([DEMO, "32"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout
    assert "NGAP_DEMO OK" in r.stdout
    assert r.stdout.count(" it=") == 4
