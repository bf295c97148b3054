"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the host C/C++ code
(the reference's FSANITIZER build option, tools/GNUMake/comps/gnu.mak:85-91):
the CPU oracle (oracle/pcg_pfmg.c) and the algebraic multigrid setup of
libngapuns.so (nga2_amd/csrc/ngap_amg.cpp), each compiled with gcc/g++ with a
driver under tests/sanitize/ and run; any report ends the run with a nonzero
status (-fno-sanitize-recover=all), and leaks are reported too."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]


def _run(cmd, env=None, timeout=600):
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, env=env)
    return p.returncode, p.stdout + p.stderr


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_oracle_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "orc_san")
    rc, out = _run(["gcc", "-std=c11"] + SAN + ["-fopenmp", os.path.join(ROOT, "tests", "sanitize", "orc_main.c"),
                                                os.path.join(ROOT, "oracle", "pcg_pfmg.c"), "-lm", "-o", exe])
    assert rc == 0, out
    env = dict(os.environ, OMP_NUM_THREADS="2", ASAN_OPTIONS="detect_leaks=1")
    rc, out = _run([exe], env=env)
    print(out)
    assert rc == 0, out
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_amg_setup_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "amg_san")
    rc, out = _run(["g++", "-std=c++17"] + SAN + ["-pthread", "-I" + os.path.join(ROOT, "include"),
                                                  "-I" + os.path.join(ROOT, "nga2_amd", "csrc"),
                                                  os.path.join(ROOT, "tests", "sanitize", "amg_main.cpp"),
                                                  os.path.join(ROOT, "nga2_amd", "csrc", "ngap_amg.cpp"), "-o", exe])
    assert rc == 0, out
    rc, out = _run([exe], env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1"))
    print(out)
    assert rc == 0, out
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out
