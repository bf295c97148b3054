"""ctypes binding of the FP64 CPU oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker. The product package
(nga2_amd) never imports it.

The oracle restates PCG + a PFMG-style V(1,1) preconditioner
(NGA2 src/solver/hypre_str_class.f90:264-284); the algorithm is written out in
the header of oracle/pcg_pfmg.c.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liborc.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: {path} (run __graft_entry__.build())")
        L = C.CDLL(path)
        L.orc_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp,
                                 C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_create.restype = C.c_int
        L.orc_set_skip_relax.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_skip_relax.restype = None
        L.orc_relax.argtypes = [C.c_void_p, C.c_int]
        L.orc_relax.restype = C.c_int
        L.orc_set_relax_level.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_set_relax_level.restype = C.c_int
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_nlevels.argtypes = [C.c_void_p]
        L.orc_coarse_singular.argtypes = [C.c_void_p]
        L.orc_coarse_dense.argtypes = [C.c_void_p]
        L.orc_level_info.argtypes = [C.c_void_p, C.c_int, _ip, _ip]
        L.orc_level_coefs.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_apply.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.orc_apply_exact.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_apply_exact.restype = None
        L.orc_rbgs.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_int]
        L.orc_resid_restrict.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp]
        L.orc_prolong_add.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.orc_coarse_solve.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_vcycle.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_vcycle.restype = None
        L.orc_solve.argtypes = [C.c_void_p, _dp, _dp, C.c_int, C.c_double, C.c_int, _ip, _dp, _dp]
        L.orc_pcg_dense.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_double, C.c_int, _ip, _dp, _dp]
        L.orc_num_threads.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_threads.restype = None
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(_dp)


def interior_opr(opr, cfg):
    """opr(7,imino:imaxo,...) (Fortran order) -> contiguous [N][7] over interior cells, x fastest."""
    it = cfg.interior
    sub = opr[(slice(None),) + it]                     # (7, nx, ny, nz), F-order strides
    return np.ascontiguousarray(np.transpose(sub, (3, 2, 1, 0)))  # (nz, ny, nx, 7): C-order == [N][7], x fastest


def interior(field, cfg):
    """interior values as a flat x-fastest float64 array"""
    return np.ascontiguousarray(np.transpose(field[cfg.interior], (2, 1, 0))).ravel()


def to_field(flat, cfg, out=None):
    if out is None:
        out = cfg.field()
    out[cfg.interior] = np.transpose(flat.reshape(cfg.nz, cfg.ny, cfg.nx), (2, 1, 0))
    return out


class Oracle:
    """FP64 CPU PCG-PFMG on the interior of one grid."""

    def __init__(self, n3, per3, opr7, maxlevel=0, coarsest_max=128, nu_pre=1, nu_post=1, skip_relax=1):
        self.n3 = tuple(int(v) for v in n3)
        self.N = self.n3[0] * self.n3[1] * self.n3[2]
        op = np.ascontiguousarray(opr7, dtype=np.float64).reshape(self.N, 7)
        self._opr = op
        h = C.c_void_p()
        rc = lib().orc_create(C.byref(h), self.n3[0], self.n3[1], self.n3[2], int(per3[0]), int(per3[1]),
                              int(per3[2]), _p(op), int(maxlevel), int(coarsest_max), int(nu_pre), int(nu_post))
        if rc != 0:
            raise RuntimeError(f"orc_create failed: {rc}")
        self.h = h
        lib().orc_set_skip_relax(h, int(skip_relax))

    def set_relax_levels(self, flags):
        """Set every level's smoothing flag (a test hook for other skip patterns)."""
        for l, f in enumerate(flags):
            lib().orc_set_relax_level(self.h, l, int(f))

    def relax_levels(self):
        """Per level, 1 when the V-cycle smooths it (orc_set_skip_relax)."""
        return [lib().orc_relax(self.h, l) for l in range(self.nlevels)]

    @classmethod
    def from_cfg(cls, cfg, opr, **kw):
        return cls((cfg.nx, cfg.ny, cfg.nz), cfg.periodic, interior_opr(opr, cfg), **kw)

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ hierarchy
    @property
    def nlevels(self):
        return lib().orc_nlevels(self.h)

    @property
    def coarse_singular(self):
        return bool(lib().orc_coarse_singular(self.h))

    @property
    def coarse_dense(self):
        return bool(lib().orc_coarse_dense(self.h))

    def level_info(self, l):
        d = (C.c_int * 3)()
        cd = C.c_int()
        rc = lib().orc_level_info(self.h, l, d, C.byref(cd))
        if rc:
            raise IndexError(l)
        return (d[0], d[1], d[2]), cd.value

    def level_size(self, l):
        d, _ = self.level_info(l)
        return d[0] * d[1] * d[2]

    def level_coefs(self, l):
        n = self.level_size(l)
        arrs = [np.zeros(n) for _ in range(6)]
        lib().orc_level_coefs(self.h, l, *[_p(a) for a in arrs])
        return dict(zip(("c", "ax", "ay", "az", "pm", "pp"), arrs))

    # ------------------------------------------------------------ pieces
    def apply(self, l, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        lib().orc_apply(self.h, l, _p(x), _p(y))
        return y

    def apply_exact(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        lib().orc_apply_exact(self.h, _p(x), _p(y))
        return y

    def rbgs(self, l, f, v, order=0):
        f = np.ascontiguousarray(f, dtype=np.float64)
        v = np.array(v, dtype=np.float64, copy=True)
        lib().orc_rbgs(self.h, l, _p(f), _p(v), order)
        return v

    def resid_restrict(self, l, f, v):
        f = np.ascontiguousarray(f, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        fc = np.zeros(self.level_size(l + 1))
        lib().orc_resid_restrict(self.h, l, _p(f), _p(v), _p(fc))
        return fc

    def prolong_add(self, l, vc, v):
        vc = np.ascontiguousarray(vc, dtype=np.float64)
        v = np.array(v, dtype=np.float64, copy=True)
        lib().orc_prolong_add(self.h, l, _p(vc), _p(v))
        return v

    def coarse_solve(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        v = np.zeros_like(f)
        lib().orc_coarse_solve(self.h, _p(f), _p(v))
        return v

    def vcycle(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros_like(r)
        lib().orc_vcycle(self.h, _p(r), _p(z))
        return z

    def solve(self, rhs, guess=None, maxit=200, rcvg=1e-10, flexible=False):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        x = np.zeros(self.N) if guess is None else np.array(guess, dtype=np.float64, copy=True)
        it = C.c_int()
        rerr = C.c_double()
        trerr = C.c_double()
        rc = lib().orc_solve(self.h, _p(rhs), _p(x), int(maxit), float(rcvg), int(flexible), C.byref(it),
                             C.byref(rerr), C.byref(trerr))
        if rc:
            raise RuntimeError(f"orc_solve failed: {rc}")
        return x, dict(it=it.value, rerr=rerr.value, true_rerr=trerr.value)


def pcg_dense(A, b, Minv=None, maxit=1000, rcvg=1e-12, flexible=False, x0=None):
    """The oracle's PCG loop on a dense SPD matrix (row-major)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    n = len(b)
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    M = None if Minv is None else np.ascontiguousarray(Minv, dtype=np.float64)
    it = C.c_int()
    rerr = C.c_double()
    trerr = C.c_double()
    lib().orc_pcg_dense(n, _p(A), None if M is None else _p(M), _p(b), _p(x), maxit, rcvg, int(flexible),
                        C.byref(it), C.byref(rerr), C.byref(trerr))
    return x, dict(it=it.value, rerr=rerr.value, true_rerr=trerr.value)


def num_threads():
    return lib().orc_num_threads()


def set_threads(n):
    """Threads of the oracle's following OpenMP regions."""
    lib().orc_set_threads(int(n))
