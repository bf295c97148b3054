"""Failure handling of host code, forced on the CPU.

The algebraic multigrid hierarchy of libngapuns.so is built on host threads
(ngap_amg_build). An allocation that fails inside a worker thread, or a
thread that cannot be started, must come back as NGAP_ERR_ALLOC instead of
ending the process (an exception escaping a std::thread calls
std::terminate). The test runs the build in a child process whose address
space is capped just above what it already uses.
"""
import os
import subprocess
import sys
import textwrap

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = textwrap.dedent(r"""
    import ctypes as C, resource, sys
    import numpy as np
    sys.path.insert(0, %(root)r)
    from nga2_amd import uns
    L = uns.lib()
    n = 48
    N = n ** 3
    idx = np.arange(N).reshape(n, n, n)
    rows, cols, vals = [idx.ravel()], [idx.ravel()], [np.full(N, 6.0)]
    for ax in range(3):
        for s in (1, -1):
            rows.append(idx.ravel()); cols.append(np.roll(idx, s, axis=ax).ravel()); vals.append(np.full(N, -1.0))
    r = np.concatenate(rows); c = np.concatenate(cols); v = np.concatenate(vals)
    o = np.lexsort((c, r)); r, c, v = r[o], c[o], v[o]
    rp = np.zeros(N + 1, dtype=np.int32); np.add.at(rp, r + 1, 1); rp = np.cumsum(rp).astype(np.int32)
    c = np.ascontiguousarray(c, dtype=np.int32); v = np.ascontiguousarray(v, dtype=np.float64)
    op = uns.UnsOpts(); L.ngap_uns_default_opts(C.byref(op))
    def vm_bytes():
        for line in open("/proc/self/status"):
            if line.startswith("VmSize:"):
                return int(line.split()[1]) * 1024
    extra = int(sys.argv[1]) * 2 ** 20
    if extra > 0:
        lim = vm_bytes() + extra
        resource.setrlimit(resource.RLIMIT_AS, (lim, lim))
    H = C.c_void_p()
    rc = L.ngap_amg_build(C.byref(H), N, rp.ctypes.data_as(C.POINTER(C.c_int)),
                          c.ctypes.data_as(C.POINTER(C.c_int)), v.ctypes.data_as(C.POINTER(C.c_double)), C.byref(op))
    if rc == 0:
        L.ngap_amg_free(H)
    print("RC", rc, flush=True)
""") % {"root": ROOT}


def _run(extra_mib):
    env = dict(os.environ, OMP_NUM_THREADS="1")
    return subprocess.run([sys.executable, "-c", CHILD, str(extra_mib)], capture_output=True, text=True, timeout=300,
                          env=env)


def test_amg_build_unlimited_succeeds():
    p = _run(0)
    assert p.returncode == 0, p.stderr
    assert "RC 0" in p.stdout


@pytest.mark.parametrize("extra_mib", [4, 24, 64])
def test_amg_build_out_of_memory_returns_code(extra_mib):
    """Under an address-space cap the build runs out of memory somewhere in
    its threaded phases: it must return NGAP_ERR_ALLOC (4) and the process
    must exit normally."""
    from nga2_amd._lib import ERR_ALLOC

    p = _run(extra_mib)
    assert p.returncode == 0, f"child ended with {p.returncode}: {p.stderr[-2000:]}"
    rc = int(p.stdout.split("RC")[-1])
    assert rc in (0, ERR_ALLOC), rc
    if extra_mib <= 4:
        assert rc == ERR_ALLOC
