"""fourier3d on the GPU (libngapfft.so, include/ngapfft.h).

Fourier3d mirrors NGA2's fourier3d linsol type (src/solver/fourier3d_class.f90):
the constructor fourier3d(cfg, name, nst) with its checks (:109-165), the
public arrays opr/rhs/sol/stc, and init/setup/solve/destroy (:171-1221), with
it = 1 and rerr = aerr = 0 as the reference leaves them (:131-134). Where the
reference dies, these raise: ValueError from the constructor's checks,
NgapError with ERR_NOT_CIRCULANT / ERR_SINGULAR from setup, ERR_NOT_SETUP
from a solve before setup.

DeviceFourier is the same solve on torch device tensors (the _device entry
points): the ghosted opr/rhs/sol in the layout of NgapStr / DeviceSolve.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import NgapError, check
from .linsol import grid_of

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libngapfft.so")
EXPORTS = ["ngap_fft_create", "ngap_fft_setup", "ngap_fft_setup_device", "ngap_fft_solve", "ngap_fft_solve_device",
           "ngap_fft_destroy", "ngap_fft_last_ms"]

_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
_L = None


def lib():
    """Load libngapfft.so; raises if it has not been built."""
    global _L
    if _L is not None:
        return _L
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libngapfft.so not found at {LIB_PATH}; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    L.ngap_fft_create.argtypes = [C.POINTER(_vp), C.POINTER(_lib.Grid), C.c_int, C.POINTER(C.c_int)]
    L.ngap_fft_setup.argtypes = [_vp, _vp, C.c_double]
    L.ngap_fft_setup_device.argtypes = [_vp, _vp, C.c_double]
    L.ngap_fft_solve.argtypes = [_vp, _vp, _vp]
    L.ngap_fft_solve_device.argtypes = [_vp, _vp, _vp]
    L.ngap_fft_destroy.argtypes = [_vp]
    L.ngap_fft_last_ms.argtypes = [_vp, _dp]
    _L = L
    return L


EPS = float(np.finfo(np.float64).eps)


def min_meshsize(cfg):
    """config_class.f90:112-125: min over interior cells of the geometric mean
    of the spacings of the directions with more than one cell."""
    it = cfg.interior
    dd = [(np.asarray(d)[s], n) for d, s, n in ((cfg.dx, it[0], cfg.nx), (cfg.dy, it[1], cfg.ny), (cfg.dz, it[2], cfg.nz))]
    pw = [1 if n > 1 else 0 for _, n in dd]
    if sum(pw) == 0:
        return 1.0
    m = (dd[0][0][:, None, None] ** pw[0] * dd[1][0][None, :, None] ** pw[1] * dd[2][0][None, None, :] ** pw[2])
    return float(np.min(m ** (1.0 / sum(pw))))


def circ_tol(cfg):
    """The circulance tolerance of fourier3d_class.f90:1124: 6 eps / min_meshsize^4."""
    return 6.0 * EPS / min_meshsize(cfg) ** 4


def _uniform(d):
    """sgrid_class.f90:266-268: |max dx - min dx| < 10 eps over the ghosted array."""
    d = np.asarray(d, dtype=np.float64)
    return bool(abs(d.max() - d.min()) < 10.0 * EPS)


def check_grid(cfg):
    """The constructor's checks (fourier3d_class.f90:144-162): every direction
    with more than one cell periodic and uniform, and more directions with
    more than one cell than decomposed ones (none on one GPU), so a 1x1x1
    grid is refused as the reference refuses it."""
    for n, per, d, ax in ((cfg.nx, cfg.xper, cfg.dx, "x"), (cfg.ny, cfg.yper, cfg.dy, "y"),
                          (cfg.nz, cfg.zper, cfg.dz, "z")):
        if n > 1 and not (per and _uniform(d)):
            raise ValueError(f"[fourier3d constructor] {ax}-direction needs to be periodic and uniform")
    if sum(n > 1 for n in (cfg.nx, cfg.ny, cfg.nz)) == 0:
        raise ValueError("[fourier3d constructor] Need at least one NON-decomposed direction")


class Fourier3d:
    """fourier3d(cfg, name, nst) with NGA2's public fields and lifecycle."""

    def __init__(self, cfg, name, nst):
        check_grid(cfg)
        self.cfg = cfg
        self.name = str(name).strip()
        self.method = 0
        self.nst = int(nst)
        self.stc = np.zeros((self.nst, 3), dtype=np.int32)
        self.opr = np.zeros((self.nst,) + cfg.shape, dtype=np.float64, order="F")
        self.rhs = cfg.field()
        self.sol = cfg.field()
        self.it, self.aerr, self.rerr = 1, 0.0, 0.0
        self.setup_done = False
        self._h = None

    def init(self):
        L = lib()
        stc = np.asfortranarray(self.stc, dtype=np.int32)
        h = _vp()
        check("ngap_fft_create", L.ngap_fft_create(C.byref(h), C.byref(grid_of(self.cfg)), self.nst,
                                                   stc.ctypes.data_as(C.POINTER(C.c_int))))
        self._h = h
        self.stmap = {tuple(int(v) for v in self.stc[s]): s + 1 for s in range(self.nst)}

    def setup(self):
        if self._h is None:
            raise NgapError("ngap_fft_setup", _lib.ERR_NOT_INIT)
        opr = np.asfortranarray(self.opr, dtype=np.float64)
        self.setup_done = False
        check("ngap_fft_setup", lib().ngap_fft_setup(self._h, opr.ctypes.data, circ_tol(self.cfg)))
        self.setup_done = True

    def solve(self):
        if self._h is None:
            raise NgapError("ngap_fft_solve", _lib.ERR_NOT_SETUP)
        rhs = np.asfortranarray(self.rhs, dtype=np.float64)
        sol = np.asfortranarray(self.sol, dtype=np.float64)
        check("ngap_fft_solve", lib().ngap_fft_solve(self._h, rhs.ctypes.data, sol.ctypes.data))
        self.sol[...] = sol

    def last_ms(self):
        v = C.c_double()
        check("ngap_fft_last_ms", lib().ngap_fft_last_ms(self._h, C.byref(v)))
        return v.value

    def destroy(self):
        if self._h is not None:
            lib().ngap_fft_destroy(self._h)
            self._h = None
        self.setup_done = False

    def print_short(self):
        print(f"fourier3d solver [{self.name:16s}] for config [{self.cfg.name:16s}]")

    def print(self):
        print(f"fourier3d solver [{self.name}] for config [{self.cfg.name}]")

    log = print


class DeviceFourier:
    """The same solve with operands in device memory (torch tensors)."""

    def __init__(self, cfg, stc=None):
        import torch

        check_grid(cfg)
        self.cfg = cfg
        if stc is None:
            stc = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]])
        stc = np.asfortranarray(np.asarray(stc, dtype=np.int32))
        self.nst = stc.shape[0]
        torch.cuda.synchronize()
        h = _vp()
        check("ngap_fft_create", lib().ngap_fft_create(C.byref(h), C.byref(grid_of(cfg)), self.nst,
                                                       stc.ctypes.data_as(C.POINTER(C.c_int))))
        self._h = h

    def setup(self, d_opr, tol=None):
        import torch

        torch.cuda.synchronize()
        check("ngap_fft_setup_device", lib().ngap_fft_setup_device(
            self._h, C.c_void_p(d_opr.data_ptr()), circ_tol(self.cfg) if tol is None else float(tol)))

    def solve(self, d_rhs, d_sol):
        import torch

        torch.cuda.synchronize()
        check("ngap_fft_solve_device", lib().ngap_fft_solve_device(self._h, C.c_void_p(d_rhs.data_ptr()),
                                                                   C.c_void_p(d_sol.data_ptr())))

    def last_ms(self):
        v = C.c_double()
        check("ngap_fft_last_ms", lib().ngap_fft_last_ms(self._h, C.byref(v)))
        return v.value

    def destroy(self):
        if self._h is not None:
            lib().ngap_fft_destroy(self._h)
            self._h = None
