"""Device-side projection (include/ngapproj.h): NGA2's operator assembly,
divergence, pressure gradient, mean removal and ghost fill on device arrays.

The arrays are torch device tensors holding the Fortran-ordered ghosted
fields (nxo*nyo*nzo doubles, i fastest), the layout of NgapStr's opr/rhs/sol
and of DeviceSolve. The metrics come from nga2_amd.incomp.Metrics (NGA2's
masked divp/divu) and are uploaded once.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib
from .linsol import grid_of

_vp = C.c_void_p
_dp = C.POINTER(C.c_double)


def _bind(L):
    if getattr(L, "_proj_bound", False):
        return L
    L.ngap_proj_create.argtypes = [C.POINTER(_vp), C.POINTER(_lib.Grid)] + [_dp] * 8
    L.ngap_proj_destroy.argtypes = [_vp]
    L.ngap_proj_laplacian.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.ngap_proj_div.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.ngap_proj_pgrad.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.ngap_proj_shift_p.argtypes = [_vp, _vp, _dp]
    L.ngap_proj_sync.argtypes = [_vp, _vp]
    L.ngap_proj_rhs.argtypes = [_vp, _vp, C.c_double, C.c_double, _vp]
    L.ngap_proj_correct.argtypes = [_vp, C.c_double, C.c_double] + [_vp] * 11
    L.ngap_proj_pgrad_jump.argtypes = [_vp] + [_vp] * 7
    L.ngap_proj_csf_jump.argtypes = [_vp, C.c_double, C.c_double] + [_vp] * 13
    L.ngap_proj_set_bconds.argtypes = [_vp, C.c_int, _vp, _dp, _dp, _dp]
    L.ngap_proj_correct_mfr.argtypes = [_vp, C.c_double, _vp, _vp, _vp, _dp]
    L._proj_bound = True
    return L


EXPORTS = ["ngap_proj_create", "ngap_proj_destroy", "ngap_proj_laplacian", "ngap_proj_div", "ngap_proj_pgrad",
           "ngap_proj_shift_p", "ngap_proj_sync", "ngap_proj_rhs", "ngap_proj_correct", "ngap_proj_pgrad_jump",
           "ngap_proj_csf_jump", "ngap_proj_set_bconds", "ngap_proj_correct_mfr", "ngap_proj_shift_p_noalloc"]


class Bcond(C.Structure):
    """ngap_bcond: one boundary condition of correct_mfr."""
    _fields_ = [("face", C.c_int), ("rdir", C.c_int), ("can_correct", C.c_int), ("n_in", C.c_int),
                ("n_all", C.c_int), ("ijk", C.POINTER(C.c_int))]


def _pairs(a, b):
    """NGA2's (0:1, i, j, k) metric layout: the pair index fastest."""
    return np.ascontiguousarray(np.stack((np.asarray(a), np.asarray(b)), axis=0).ravel(order="F"))


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


class DeviceProjection:
    """Handle to ngap_proj for one config and its masked metrics."""

    def __init__(self, met):
        cfg = met.cfg
        self.cfg = cfg
        L = _bind(lib())
        dp = [_pairs(*met.divp[ax]) for ax in range(3)]
        du = [_pairs(*met.divu[ax]) for ax in range(3)]
        vol = np.ascontiguousarray(np.asarray(cfg.vol, dtype=np.float64).ravel(order="F"))
        vf = np.ascontiguousarray(np.asarray(cfg.VF, dtype=np.float64).ravel(order="F"))
        self._keep = dp + du + [vol, vf]
        self._h = _vp()
        args = [a.ctypes.data_as(_dp) for a in dp + du] + [vol.ctypes.data_as(_dp), vf.ctypes.data_as(_dp)]
        check("ngap_proj_create", L.ngap_proj_create(C.byref(self._h), C.byref(grid_of(cfg)), *args))

    @staticmethod
    def _torch_sync():
        import torch

        torch.cuda.synchronize()

    def laplacian(self, d_opr, rho_faces=None):
        """d_opr (7*No) = the pressure operator; rho_faces = (rho_U, rho_V, rho_W) device tensors or None."""
        self._torch_sync()
        ru, rv, rw = rho_faces if rho_faces is not None else (None, None, None)
        check("ngap_proj_laplacian", lib().ngap_proj_laplacian(self._h, _ptr(ru), _ptr(rv), _ptr(rw), _ptr(d_opr)))

    def div(self, dU, dV, dW, d_div):
        self._torch_sync()
        check("ngap_proj_div", lib().ngap_proj_div(self._h, _ptr(dU), _ptr(dV), _ptr(dW), _ptr(d_div)))

    def pgrad(self, dP, dPx, dPy, dPz):
        self._torch_sync()
        check("ngap_proj_pgrad", lib().ngap_proj_pgrad(self._h, _ptr(dP), _ptr(dPx), _ptr(dPy), _ptr(dPz)))

    def shift_p(self, dP):
        """Remove the VF-vol weighted mean in place; returns the removed value."""
        self._torch_sync()
        m = C.c_double()
        check("ngap_proj_shift_p", lib().ngap_proj_shift_p(self._h, _ptr(dP), C.byref(m)))
        return m.value

    def sync(self, dA):
        self._torch_sync()
        check("ngap_proj_sync", lib().ngap_proj_sync(self._h, _ptr(dA)))

    def rhs(self, d_div, rho, dt, d_rhs):
        """d_rhs = -vol*div*rho/dt (rho = 1 for the two-phase form)."""
        self._torch_sync()
        check("ngap_proj_rhs", lib().ngap_proj_rhs(self._h, _ptr(d_div), float(rho), float(dt), _ptr(d_rhs)))

    def correct(self, dt, rho, grads, vel, rho_faces=None, P=None, sol=None):
        """U -= dt*Px/rho (or /rho_U), likewise V, W; P += sol when given."""
        self._torch_sync()
        ru, rv, rw = rho_faces if rho_faces is not None else (None, None, None)
        check("ngap_proj_correct", lib().ngap_proj_correct(
            self._h, float(dt), float(rho), _ptr(ru), _ptr(rv), _ptr(rw), *[_ptr(g) for g in grads],
            *[_ptr(v) for v in vel], _ptr(P), _ptr(sol)))

    def pgrad_jump(self, dP, dPj, grads):
        self._torch_sync()
        check("ngap_proj_pgrad_jump", lib().ngap_proj_pgrad_jump(self._h, _ptr(dP), *[_ptr(a) for a in dPj],
                                                                 *[_ptr(g) for g in grads]))

    def csf_jump(self, sigma, dt, SD, curv, VF, rho_faces, Pj, DPj, div):
        """Pj (in: previous, out: new), DPj (out), div (in/out), all device tensors."""
        self._torch_sync()
        check("ngap_proj_csf_jump", lib().ngap_proj_csf_jump(
            self._h, float(sigma), float(dt), _ptr(SD), _ptr(curv), _ptr(VF), *[_ptr(r) for r in rho_faces],
            *[_ptr(a) for a in Pj], *[_ptr(a) for a in DPj], _ptr(div)))

    def set_bconds(self, bconds):
        """bconds: dicts with face, rdir, can_correct, n_in, ijk ((n_all, 3) Fortran indices)."""
        cfg = self.cfg
        self._bc_keep = []
        arr = (Bcond * max(1, len(bconds)))()
        for t, bc in enumerate(bconds):
            ijk = np.ascontiguousarray(np.asarray(bc["ijk"], dtype=np.int32).reshape(-1, 3))
            self._bc_keep.append(ijk)
            arr[t] = Bcond(int(bc["face"]), int(bc["rdir"]), int(bool(bc["can_correct"])), int(bc["n_in"]),
                           int(ijk.shape[0]), ijk.ctypes.data_as(C.POINTER(C.c_int)))
        d = [np.ascontiguousarray(np.asarray(a, dtype=np.float64)) for a in (cfg.dx, cfg.dy, cfg.dz)]
        self._bc_keep += d
        check("ngap_proj_set_bconds", lib().ngap_proj_set_bconds(self._h, len(bconds), C.cast(arr, _vp),
                                                                 *[a.ctypes.data_as(_dp) for a in d]))

    def correct_mfr(self, rho, vel):
        """correct_mfr on device velocities; returns the flow-rate mismatch."""
        self._torch_sync()
        e = C.c_double()
        check("ngap_proj_correct_mfr", lib().ngap_proj_correct_mfr(self._h, float(rho), *[_ptr(v) for v in vel],
                                                                   C.byref(e)))
        return e.value

    def destroy(self):
        if self._h:
            lib().ngap_proj_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
