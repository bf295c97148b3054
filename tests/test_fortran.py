"""The Fortran drop-ins (nga2_amd/fortran/ngap_str_class.f90 for hypre_str,
ngap_fourier3d_class.f90 for fourier3d, ngap_uns_class.f90 for hypre_uns): they compile with amdflang against the
linsol contract and link libngapsolve.so / libngapfft.so. The self-checking
drivers solve manufactured problems on the GPU: PCG in FP64 and mixed
precision (periodic box and channel with walls), and the FFT direct solve
(anisotropic 3-D box, one-cell-thick 2-D box), and six hypre_uns methods on a
periodic box around a solid block."""
import os
import subprocess

import pytest

from conftest import ROOT

DEMO = os.path.join(ROOT, "nga2_amd", "lib", "fortran", "ngap_demo")
FFT_DEMO = os.path.join(ROOT, "nga2_amd", "lib", "fortran", "ngap_fft_demo")
UNS_DEMO = os.path.join(ROOT, "nga2_amd", "lib", "fortran", "ngap_uns_demo")


def _need_demo():
    if not (os.path.exists(DEMO) and os.path.exists(FFT_DEMO) and os.path.exists(UNS_DEMO)):
        from nga2_amd.build import build_fortran, flang

        if flang() is None:
            pytest.skip("no Fortran compiler in this image")
        build_fortran()
    assert os.path.exists(DEMO) and os.path.exists(FFT_DEMO) and os.path.exists(UNS_DEMO)


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
    r = subprocess.run([FFT_DEMO, "8"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 2
    assert "[ngap_fourier3d init] HIP runtime error" in r.stdout
    r = subprocess.run([UNS_DEMO, "8"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 2
    assert "[ngap_uns init] HIP runtime error" in r.stdout


@pytest.mark.gpu
def test_fortran_driver_on_gpu():
    _need_demo()
    print("\n[fortran demo 32^3]", flush=True)
    r = subprocess.run([DEMO, "32"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout
    assert "NGAP_DEMO OK" in r.stdout
    assert r.stdout.count(" it=") == 4


@pytest.mark.gpu
def test_fortran_fft_driver_on_gpu():
    _need_demo()
    print("\n[fortran fft demo 32^3]", flush=True)
    r = subprocess.run([FFT_DEMO, "32"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout
    assert "NGAP_FFT_DEMO OK" in r.stdout
    assert r.stdout.count(" it=1 ") == 2


@pytest.mark.gpu
def test_fortran_uns_driver_on_gpu():
    _need_demo()
    print("\n[fortran uns demo 24^3]", flush=True)
    r = subprocess.run([UNS_DEMO, "24"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout
    assert "NGAP_UNS_DEMO OK" in r.stdout
    assert r.stdout.count(" it=") == 6
