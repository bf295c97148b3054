"""Python mirror of NGA2's ``linsol`` / ``hypre_str`` interface over libngapsolve.

``NgapStr(cfg, name, method, nst)`` keeps the reference constructor signature
(src/solver/hypre_str_class.f90:68-103), the public fields of ``linsol``
(src/solver/linsol_class.f90:16-58: ``opr``, ``rhs``, ``sol``, ``stc``,
``stmap``, ``maxit``, ``rcvg``, ``acvg``, ``it``, ``rerr``, ``aerr``,
``setup_done``), ``maxlevel`` (:34) and the lifecycle
``init`` / ``setup`` / ``solve`` / ``destroy`` / ``log`` / ``print`` /
``print_short`` (:109-463). Misuse raises ``NgapError`` where the reference
calls ``die`` (:311, :330).

Arrays are numpy arrays in Fortran order with the reference's bounds:
``opr[st, i, j, k]`` for Fortran ``opr(st+1, imino+i, jmino+j, kmino+k)``.
"""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np

from . import _lib
from ._lib import FP64, MIXED, PCG_PFMG, NgapError, check, lib

__all__ = ["NgapStr", "DeviceSolve", "FP64", "MIXED", "PCG_PFMG", "NgapError"]


def grid_of(cfg):
    g = _lib.Grid()
    g.nx, g.ny, g.nz = cfg.nx, cfg.ny, cfg.nz
    g.imin_, g.imax_ = cfg.imin, cfg.imax
    g.jmin_, g.jmax_ = cfg.jmin, cfg.jmax
    g.kmin_, g.kmax_ = cfg.kmin, cfg.kmax
    g.imino_, g.imaxo_ = cfg.imino, cfg.imaxo
    g.jmino_, g.jmaxo_ = cfg.jmino, cfg.jmaxo
    g.kmino_, g.kmaxo_ = cfg.kmino, cfg.kmaxo
    g.xper, g.yper, g.zper = int(cfg.xper), int(cfg.yper), int(cfg.zper)
    return g


class NgapStr:
    """Drop-in for ``hypre_str`` with methods ``pcg_pfmg`` and ``gmres_pfmg``
    on one MI355X (``opts['kdim']``: GMRES restart length, default 5)."""

    def __init__(self, cfg, name, method, nst):
        self.cfg = cfg
        self.name = str(name).strip()
        self.method = int(method)
        self.nst = int(nst)
        self.stc = np.zeros((self.nst, 3), dtype=np.int32, order="F")
        self.stmap = None
        self.opr = np.zeros((self.nst,) + cfg.shape, dtype=np.float64, order="F")
        self.rhs = cfg.field()
        self.sol = cfg.field()
        self.it = 0
        self.aerr = 0.0
        self.rerr = 0.0
        self.maxit = 100
        self.rcvg = 1e-6
        self.acvg = 0.0
        self.maxlevel = 0
        self.setup_done = False
        # build-specific knobs (no counterpart in hypre_str)
        self.precision = FP64
        self.opts = {}
        self.verbose = 0
        self._h = C.c_void_p()

    # ----------------------------------------------------------------- lifecycle
    def init(self):
        """hypre_str_init (:109-145): stencil map and grid. Call after stc is set."""
        s = self.stc
        lo = s.min(axis=0)
        hi = s.max(axis=0)
        self.stmap = np.zeros(tuple(hi - lo + 1), dtype=np.int32)
        for st in range(self.nst):
            self.stmap[tuple(s[st] - lo)] = st + 1
        if not self._h:
            stc = np.asfortranarray(self.stc, dtype=np.int32)
            check("ngap_create", lib().ngap_create(C.byref(self._h), C.byref(grid_of(self.cfg)), self.nst,
                                                   stc.ctypes.data_as(C.POINTER(C.c_int)), self.method))
            self._stc_keep = stc
        check("ngap_init", lib().ngap_init(self._h))

    def _opts(self):
        o = _lib.default_opts()
        o.precision = int(self.precision)
        o.verbose = int(self.verbose)
        for k, v in self.opts.items():
            setattr(o, k, int(v))
        return o

    def setup(self):
        """hypre_str_setup (:149-318): operator deep copy and hierarchy build."""
        if not self._h:
            raise NgapError("ngap_setup", _lib.ERR_NOT_INIT)
        opr = self.opr if self.opr.flags.f_contiguous else np.asfortranarray(self.opr)
        check("ngap_setup", lib().ngap_setup(self._h, opr.ctypes.data, int(self.maxit), float(self.rcvg),
                                             int(self.maxlevel), C.byref(self._opts())))
        self.setup_done = True

    def solve(self):
        """hypre_str_solve (:322-386): sol holds the guess, returns the solution
        with ghosts synced (cfg%sync, :380)."""
        if not self.setup_done:
            raise NgapError("ngap_solve", _lib.ERR_NOT_SETUP)
        self.it = -1
        self.aerr = self.rerr = sys.float_info.max
        rhs = self.rhs if self.rhs.flags.f_contiguous else np.asfortranarray(self.rhs)
        if not self.sol.flags.f_contiguous:
            self.sol = np.asfortranarray(self.sol)
        it = C.c_int()
        rerr = C.c_double()
        aerr = C.c_double()
        check("ngap_solve", lib().ngap_solve(self._h, rhs.ctypes.data, self.sol.ctypes.data, C.byref(it),
                                             C.byref(rerr), C.byref(aerr)))
        self.it, self.rerr, self.aerr = it.value, rerr.value, aerr.value
        self.cfg.sync(self.sol)
        if self.verbose > 0:
            self.log()
        if self.verbose > 1:
            self.print_short()

    def set_tolerance(self, maxit=None, rcvg=None):
        if maxit is not None:
            self.maxit = int(maxit)
        if rcvg is not None:
            self.rcvg = float(rcvg)
        if self._h:
            check("ngap_set_tolerance", lib().ngap_set_tolerance(self._h, int(self.maxit), float(self.rcvg)))

    def destroy(self):
        """hypre_str_destroy (:390-422)."""
        if self._h:
            lib().ngap_destroy(self._h)
            self._h = C.c_void_p()
        self.setup_done = False

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def stats(self):
        s = _lib.Stats()
        check("ngap_last_stats", lib().ngap_last_stats(self._h, C.byref(s)))
        return s.as_dict()

    # ----------------------------------------------------------------- printing
    def _lines(self):
        return [f"Structured MI355X solver [{self.name}] for config [{self.cfg.name}]",
                f" >     method = {self.method}",
                f" >   it/maxit = {self.it}/{self.maxit}",
                f" >  aerr/acvg = {self.aerr:12.5e}/{self.acvg:12.5e}",
                f" >  rerr/rcvg = {self.rerr:12.5e}/{self.rcvg:12.5e}"]

    def log(self):
        for ln in self._lines():
            print(ln, file=sys.stderr)

    def print(self):
        for ln in self._lines():
            print(ln)

    def print_short(self):
        print(f"Structured MI355X solver [{self.name:>16}] for config [{self.cfg.name:>16}] -> it/maxit = "
              f"{self.it:3d}/{self.maxit:3d} and rerr/rcvg = {self.rerr:12.5e}/{self.rcvg:12.5e}")

    # ----------------------------------------------------------------- test hooks
    def hooks(self):
        return KernelHooks(self._h)


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class KernelHooks:
    """Per-kernel entry points (include/ngapsolve_test.h) on host arrays."""

    def __init__(self, h):
        self.h = h

    @property
    def nlevels(self):
        return lib().ngapt_nlevels(self.h)

    def level_info(self, l):
        d = (C.c_int * 3)()
        cd = C.c_int()
        check("ngapt_level_info", lib().ngapt_level_info(self.h, l, d, C.byref(cd)))
        return (d[0], d[1], d[2]), cd.value

    def level_size(self, l):
        d, _ = self.level_info(l)
        return d[0] * d[1] * d[2]

    def level_coefs(self, l):
        n = self.level_size(l)
        arrs = [np.zeros(n) for _ in range(6)]
        check("ngapt_level_coefs", lib().ngapt_level_coefs(self.h, l, *[_dp(a) for a in arrs]))
        return dict(zip(("c", "ax", "ay", "az", "pm", "pp"), arrs))

    def apply(self, l, x, prec=0):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        check("ngapt_apply", lib().ngapt_apply(self.h, l, prec, _dp(x), _dp(y)))
        return y

    def rbgs(self, l, f, v, order=0, prec=0):
        f = np.ascontiguousarray(f, dtype=np.float64)
        v = np.array(v, dtype=np.float64, copy=True)
        check("ngapt_rbgs", lib().ngapt_rbgs(self.h, l, prec, _dp(f), _dp(v), order))
        return v

    def resid_restrict(self, l, f, v, prec=0):
        f = np.ascontiguousarray(f, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        fc = np.zeros(self.level_size(l + 1))
        check("ngapt_resid_restrict", lib().ngapt_resid_restrict(self.h, l, prec, _dp(f), _dp(v), _dp(fc)))
        return fc

    def prolong_add(self, l, vc, v, prec=0):
        vc = np.ascontiguousarray(vc, dtype=np.float64)
        v = np.array(v, dtype=np.float64, copy=True)
        check("ngapt_prolong_add", lib().ngapt_prolong_add(self.h, l, prec, _dp(vc), _dp(v)))
        return v

    def coarse_solve(self, f, prec=0):
        f = np.ascontiguousarray(f, dtype=np.float64)
        v = np.zeros_like(f)
        check("ngapt_coarse_solve", lib().ngapt_coarse_solve(self.h, prec, _dp(f), _dp(v)))
        return v

    def vcycle(self, r, prec=0):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros_like(r)
        check("ngapt_vcycle", lib().ngapt_vcycle(self.h, prec, _dp(r), _dp(z)))
        return z

    def apply_exact(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        check("ngapt_apply_exact", lib().ngapt_apply_exact(self.h, _dp(x), _dp(y)))
        return y

    def dot(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        out = C.c_double()
        check("ngapt_dot", lib().ngapt_dot(self.h, len(x), _dp(x), _dp(y), C.byref(out)))
        return out.value

    def time_kernel(self, kid, prec=0, reps=20):
        ms = C.c_double()
        b = C.c_double()
        check("ngapt_time_kernel", lib().ngapt_time_kernel(self.h, kid, prec, reps, C.byref(ms), C.byref(b)))
        return ms.value, b.value


class DeviceSolve:
    """Solver whose opr/rhs/sol live in device memory (torch CUDA tensors with
    the same ghosted Fortran layout as NgapStr's arrays). Uses the _device
    entry points; no host copies inside setup/solve."""

    def __init__(self, cfg, stc=None, method=PCG_PFMG):
        self.cfg = cfg
        self._h = C.c_void_p()
        stc = np.asfortranarray(np.array(stc if stc is not None else
                                         [[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1],
                                          [0, 0, -1]], dtype=np.int32))
        self._stc = stc
        check("ngap_create", lib().ngap_create(C.byref(self._h), C.byref(grid_of(cfg)), stc.shape[0],
                                               stc.ctypes.data_as(C.POINTER(C.c_int)), method))
        check("ngap_init", lib().ngap_init(self._h))

    @staticmethod
    def _torch_sync(t):
        # torch runs on its own HIP runtime: finish its queued work on t's
        # device before the library reads the buffer
        import torch

        torch.cuda.synchronize(t.device)

    def setup(self, d_opr, maxit=200, rcvg=1e-10, maxlevel=0, **opts):
        self._torch_sync(d_opr)
        o = _lib.default_opts(**opts)
        check("ngap_setup_device", lib().ngap_setup_device(self._h, C.c_void_p(d_opr.data_ptr()), int(maxit),
                                                           float(rcvg), int(maxlevel), C.byref(o)))

    def solve(self, d_rhs, d_sol):
        self._torch_sync(d_rhs)
        it = C.c_int()
        rerr = C.c_double()
        aerr = C.c_double()
        check("ngap_solve_device", lib().ngap_solve_device(self._h, C.c_void_p(d_rhs.data_ptr()),
                                                           C.c_void_p(d_sol.data_ptr()), C.byref(it),
                                                           C.byref(rerr), C.byref(aerr)))
        return it.value, rerr.value

    def set_tolerance(self, maxit, rcvg):
        check("ngap_set_tolerance", lib().ngap_set_tolerance(self._h, int(maxit), float(rcvg)))

    def residual(self, d_rhs, d_sol, d_res):
        """d_res (interior, x fastest) = rhs - A sol; returns ||res||_2."""
        self._torch_sync(d_rhs)
        nrm = C.c_double()
        check("ngap_residual_device", lib().ngap_residual_device(self._h, C.c_void_p(d_rhs.data_ptr()),
                                                                 C.c_void_p(d_sol.data_ptr()),
                                                                 C.c_void_p(d_res.data_ptr()), C.byref(nrm)))
        return nrm.value

    def apply(self, d_x, d_y):
        """d_y = A d_x over the interior cells (x fastest), FP64 device tensors."""
        self._torch_sync(d_x)
        check("ngap_apply_device", lib().ngap_apply_device(self._h, C.c_void_p(d_x.data_ptr()),
                                                           C.c_void_p(d_y.data_ptr())))

    def set_stage(self, stage, flags=_lib.STAGE_GUESS, precond_iters=0):
        """Attach a correction stage, given as the (function pointer, context)
        pair of unet.UNet.stage(); None removes it."""
        fn, ctx = stage if stage is not None else (None, None)
        self._stage_ref = stage  # the context must outlive the handle's use of it
        check("ngap_set_stage", lib().ngap_set_stage(self._h, fn, ctx, int(flags) if fn else 0,
                                                     int(precond_iters) if fn else 0))

    def stats(self):
        s = _lib.Stats()
        check("ngap_last_stats", lib().ngap_last_stats(self._h, C.byref(s)))
        return s.as_dict()

    def hooks(self):
        return KernelHooks(self._h)

    def destroy(self):
        if self._h:
            lib().ngap_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
