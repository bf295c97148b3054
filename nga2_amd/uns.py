"""hypre_uns on the GPU (libngapuns.so, include/ngapuns.h).

NgapUns mirrors NGA2's hypre_uns linsol type (src/solver/hypre_uns_class.f90):
the constructor hypre_uns(cfg, name, method, nst) (:84-119), the method
constants (:17-24), the public arrays opr/rhs/sol/stc, maxit/rcvg, it/rerr/aerr,
maxlevel (kept and unused, as the reference leaves it, :262), the cell
numbering ind and ncell, and init/setup/solve/destroy (:124-544). Where the
reference dies, these raise NgapError: ERR_NOT_SETUP for a solve before setup
(:437), ERR_METHOD for an unknown method (:419) and for pcg_parasail and
gmres_pilut, which this build does not provide.

init() numbers the cells from opr(1,:,:,:) as it is at that moment (prep_umap,
:562-607): set opr(1) (or the whole operator) before calling it. solve()
writes the participating cells of sol and syncs the ghosts (:485-497); other
cells keep their values.

DeviceUns is the same solve on torch device tensors (the _device entry point).
Options (NgapUns.opts, DeviceUns keywords) are the fields of ngap_uns_opts;
precision=1 runs the multigrid cycle in FP32 under the FP64 iteration.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import NgapError, check
from .linsol import grid_of

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libngapuns.so")
EXPORTS = [
    "ngap_uns_default_opts", "ngap_uns_create", "ngap_uns_init", "ngap_uns_setup", "ngap_uns_solve",
    "ngap_uns_solve_device", "ngap_uns_vcycle_device", "ngap_uns_destroy", "ngap_uns_rows", "ngap_uns_index",
    "ngap_uns_last_stats", "ngap_uns_hierarchy", "ngap_uns_umap_host", "ngap_uns_assemble_host", "ngap_amg_build",
    "ngap_amg_nlevels", "ngap_amg_level_sizes", "ngap_amg_level", "ngap_amg_free",
]

# hypre_uns_class.f90:17-24
AMG, PCG_AMG, GMRES_AMG, BICGSTAB_AMG, PCG, PCG_PARASAIL, GMRES, GMRES_PILUT = 1, 2, 3, 4, 5, 6, 7, 8
METHOD_NAMES = {AMG: "amg", PCG_AMG: "pcg_amg", GMRES_AMG: "gmres_amg", BICGSTAB_AMG: "bicgstab_amg", PCG: "pcg",
                PCG_PARASAIL: "pcg_parasail", GMRES: "gmres", GMRES_PILUT: "gmres_pilut"}

_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_L = None


class UnsOpts(C.Structure):
    _fields_ = [("strong_threshold", C.c_double), ("max_row_sum", C.c_double), ("pmax", C.c_int),
                ("max_coarse", C.c_int), ("max_levels", C.c_int), ("nu", C.c_int), ("kdim", C.c_int),
                ("graphs", C.c_int), ("verbose", C.c_int), ("precision", C.c_int)]


class UnsStats(C.Structure):
    _fields_ = [("status", C.c_int), ("it", C.c_int), ("nlevels", C.c_int), ("coarse_singular", C.c_int),
                ("rows", C.c_longlong), ("nnz", C.c_longlong), ("op_complexity", C.c_double), ("rerr", C.c_double),
                ("aerr", C.c_double), ("setup_ms", C.c_double), ("hier_ms", C.c_double), ("solve_ms", C.c_double),
                ("vcycle_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def lib():
    """Load libngapuns.so; raises if it has not been built."""
    global _L
    if _L is not None:
        return _L
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libngapuns.so not found at {LIB_PATH}; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    L.ngap_uns_default_opts.argtypes = [C.POINTER(UnsOpts)]
    L.ngap_uns_create.argtypes = [C.POINTER(_vp), C.POINTER(_lib.Grid), C.c_int, _ip, C.c_int]
    L.ngap_uns_init.argtypes = [_vp, _vp]
    L.ngap_uns_setup.argtypes = [_vp, _vp, C.c_int, C.c_double, C.POINTER(UnsOpts)]
    L.ngap_uns_solve.argtypes = [_vp, _vp, _vp, _ip, _dp, _dp]
    L.ngap_uns_solve_device.argtypes = [_vp, _vp, _vp, _ip, _dp, _dp]
    L.ngap_uns_vcycle_device.argtypes = [_vp, _vp, _vp]
    L.ngap_uns_destroy.argtypes = [_vp]
    L.ngap_uns_rows.argtypes = [_vp]
    L.ngap_uns_rows.restype = C.c_longlong
    L.ngap_uns_index.argtypes = [_vp, _ip]
    L.ngap_uns_last_stats.argtypes = [_vp, C.POINTER(UnsStats)]
    L.ngap_uns_hierarchy.argtypes = [_vp]
    L.ngap_uns_hierarchy.restype = _vp
    L.ngap_uns_umap_host.argtypes = [C.POINTER(_lib.Grid), _vp, C.c_int, _ip]
    L.ngap_uns_umap_host.restype = C.c_longlong
    L.ngap_uns_assemble_host.argtypes = [C.POINTER(_lib.Grid), C.c_int, _ip, _ip, _vp, _ip, _ip, _dp]
    L.ngap_uns_assemble_host.restype = C.c_longlong
    L.ngap_amg_build.argtypes = [C.POINTER(_vp), C.c_int, _ip, _ip, _dp, C.POINTER(UnsOpts)]
    L.ngap_amg_nlevels.argtypes = [_vp]
    L.ngap_amg_level_sizes.argtypes = [_vp, C.c_int, _ip, _ip, _ip]
    L.ngap_amg_level.argtypes = [_vp, C.c_int, _ip, _ip, _dp, _ip, _ip, _ip, _dp]
    L.ngap_amg_free.argtypes = [_vp]
    _L = L
    return L


def default_opts(**kw):
    o = UnsOpts()
    lib().ngap_uns_default_opts(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, type(getattr(o, k))(v))
    return o


def _ptr(a, t=_ip):
    return a.ctypes.data_as(t)


# ------------------------------------------------------------------ host pieces (no GPU)
def umap_host(cfg, opr):
    """prep_umap: (ind over the ghosted box, int32, Fortran order; ncell)."""
    opr = np.asfortranarray(opr, dtype=np.float64)
    ind = np.zeros(cfg.shape, dtype=np.int32, order="F")
    n = lib().ngap_uns_umap_host(C.byref(grid_of(cfg)), opr.ctypes.data, opr.shape[0], _ptr(ind))
    if n < 0:
        raise NgapError("ngap_uns_umap_host", int(n))
    return ind, int(n)


def assemble_host(cfg, stc, ind, opr):
    """The CSR rows of hypre_uns_setup (rowptr, col, val)."""
    stc = np.asfortranarray(stc, dtype=np.int32)
    opr = np.asfortranarray(opr, dtype=np.float64)
    ind = np.asfortranarray(ind, dtype=np.int32)
    g = grid_of(cfg)
    L = lib()
    nnz = L.ngap_uns_assemble_host(C.byref(g), stc.shape[0], _ptr(stc), _ptr(ind), opr.ctypes.data, None, None, None)
    if nnz < 0:
        raise NgapError("ngap_uns_assemble_host", int(nnz))
    n = int((ind[cfg.interior] >= 0).sum())
    rp = np.zeros(n + 1, dtype=np.int32)
    col = np.zeros(max(nnz, 1), dtype=np.int32)
    val = np.zeros(max(nnz, 1), dtype=np.float64)
    L.ngap_uns_assemble_host(C.byref(g), stc.shape[0], _ptr(stc), _ptr(ind), opr.ctypes.data, _ptr(rp), _ptr(col),
                             _ptr(val, _dp))
    return rp, col[:nnz], val[:nnz]


class Hierarchy:
    """A multigrid hierarchy built by the library (ngap_amg_build), read back level by level."""

    def __init__(self, handle, owned=True):
        self._H = handle
        self._owned = owned

    @classmethod
    def build(cls, rowptr, col, val, **opts):
        rp = np.ascontiguousarray(rowptr, dtype=np.int32)
        cl = np.ascontiguousarray(col, dtype=np.int32)
        vl = np.ascontiguousarray(val, dtype=np.float64)
        H = _vp()
        o = default_opts(**opts)
        check("ngap_amg_build", lib().ngap_amg_build(C.byref(H), len(rp) - 1, _ptr(rp), _ptr(cl), _ptr(vl, _dp),
                                                     C.byref(o)))
        return cls(H)

    def nlevels(self):
        return int(lib().ngap_amg_nlevels(self._H))

    def level(self, l):
        """dict with A (rowptr, col, val), and for l < nlevels-1 cf and P (rowptr, col, val)."""
        L = lib()
        n, nnz, pnnz = C.c_int(), C.c_int(), C.c_int()
        check("ngap_amg_level_sizes", L.ngap_amg_level_sizes(self._H, l, C.byref(n), C.byref(nnz), C.byref(pnnz)))
        n, nnz, pnnz = n.value, nnz.value, pnnz.value
        rp = np.zeros(n + 1, dtype=np.int32)
        col = np.zeros(max(nnz, 1), dtype=np.int32)
        val = np.zeros(max(nnz, 1))
        cf = np.zeros(n, dtype=np.int32)
        prp = np.zeros(n + 1, dtype=np.int32)
        pcol = np.zeros(max(pnnz, 1), dtype=np.int32)
        pval = np.zeros(max(pnnz, 1))
        check("ngap_amg_level", L.ngap_amg_level(self._H, l, _ptr(rp), _ptr(col), _ptr(val, _dp), _ptr(cf), _ptr(prp),
                                                 _ptr(pcol), _ptr(pval, _dp)))
        out = {"n": n, "A": (rp, col[:nnz], val[:nnz])}
        if l < self.nlevels() - 1:
            out["cf"] = cf
            out["P"] = (prp, pcol[:pnnz], pval[:pnnz])
        return out

    def free(self):
        if self._H is not None and self._owned:
            lib().ngap_amg_free(self._H)
        self._H = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ------------------------------------------------------------------ the linsol type
class NgapUns:
    """hypre_uns(cfg, name, method, nst) with NGA2's public fields and lifecycle."""

    def __init__(self, cfg, name, method, nst):
        self.cfg = cfg
        self.name = str(name).strip()
        self.method = int(method)
        self.nst = int(nst)
        self.stc = np.zeros((self.nst, 3), dtype=np.int32)
        self.opr = np.zeros((self.nst,) + cfg.shape, dtype=np.float64, order="F")
        self.rhs = cfg.field()
        self.sol = cfg.field()
        self.maxit, self.rcvg, self.acvg = 0, 0.0, 0.0
        self.it, self.aerr, self.rerr = 0, 0.0, 0.0
        self.maxlevel = 0
        self.ind = None
        self.ncell = self.ncell_ = 0
        self.setup_done = False
        self.opts = {}
        self._h = None

    def init(self):
        L = lib()
        stc = np.asfortranarray(self.stc, dtype=np.int32)
        h = _vp()
        check("ngap_uns_create", L.ngap_uns_create(C.byref(h), C.byref(grid_of(self.cfg)), self.nst, _ptr(stc),
                                                   self.method))
        self._h = h
        self.stmap = {tuple(int(v) for v in self.stc[s]): s + 1 for s in range(self.nst)}
        opr = np.asfortranarray(self.opr)
        check("ngap_uns_init", L.ngap_uns_init(h, opr.ctypes.data))
        self.ind = np.zeros(self.cfg.shape, dtype=np.int32, order="F")
        check("ngap_uns_index", L.ngap_uns_index(h, _ptr(self.ind)))
        self.ncell = self.ncell_ = int(L.ngap_uns_rows(h))

    def setup(self):
        if self._h is None:
            raise NgapError("ngap_uns_setup", _lib.ERR_NOT_INIT)
        opr = np.asfortranarray(self.opr)
        o = default_opts(**self.opts)
        self.setup_done = False
        check("ngap_uns_setup", lib().ngap_uns_setup(self._h, opr.ctypes.data, int(self.maxit), float(self.rcvg),
                                                     C.byref(o)))
        self.setup_done = True

    def solve(self):
        if self._h is None or not self.setup_done:
            raise NgapError("ngap_uns_solve", _lib.ERR_NOT_SETUP)
        self.it, self.aerr, self.rerr = -1, float(np.finfo(float).max), float(np.finfo(float).max)
        rhs = np.asfortranarray(self.rhs, dtype=np.float64)
        if not self.sol.flags.f_contiguous:
            self.sol = np.asfortranarray(self.sol)
        it, rerr, aerr = C.c_int(), C.c_double(), C.c_double()
        check("ngap_uns_solve", lib().ngap_uns_solve(self._h, rhs.ctypes.data, self.sol.ctypes.data, C.byref(it),
                                                     C.byref(rerr), C.byref(aerr)))
        self.it, self.rerr, self.aerr = it.value, rerr.value, aerr.value
        self.cfg.sync(self.sol)

    def hierarchy(self):
        return Hierarchy(lib().ngap_uns_hierarchy(self._h), owned=False)

    def stats(self):
        s = UnsStats()
        check("ngap_uns_last_stats", lib().ngap_uns_last_stats(self._h, C.byref(s)))
        return s.as_dict()

    def destroy(self):
        if self._h is not None:
            lib().ngap_uns_destroy(self._h)
            self._h = None
        self.setup_done = False

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def _lines(self):
        return [f"Unstructured Hypre Linear Solver [{self.name}] for config [{self.cfg.name}]",
                f" >     method = {self.method}",
                f" >   it/maxit = {self.it}/{self.maxit}",
                f" >  aerr/acvg = {self.aerr:12.5E}/{self.acvg:12.5E}",
                f" >  rerr/rcvg = {self.rerr:12.5E}/{self.rcvg:12.5E}"]

    def print(self):
        print("\n".join(self._lines()))

    def print_short(self):
        print(f"Unstructured Hypre Linear Solver [{self.name:16s}] for config [{self.cfg.name:16s}] -> "
              f"it/maxit = {self.it:3d}/{self.maxit:3d} and rerr/rcvg = {self.rerr:12.5E}/{self.rcvg:12.5E}")


class DeviceUns:
    """The same solver on torch CUDA tensors (ghosted float64, Fortran order flattened)."""

    def __init__(self, cfg, opr, method=PCG_AMG, stc=None, maxit=200, rcvg=1e-10, **opts):
        from .incomp import STC7
        self.cfg = cfg
        stc = np.asfortranarray(STC7 if stc is None else stc, dtype=np.int32)
        opr = np.asfortranarray(opr, dtype=np.float64)
        L = lib()
        h = _vp()
        check("ngap_uns_create", L.ngap_uns_create(C.byref(h), C.byref(grid_of(cfg)), stc.shape[0], _ptr(stc), method))
        self._h = h
        check("ngap_uns_init", L.ngap_uns_init(h, opr.ctypes.data))
        o = default_opts(**opts)
        check("ngap_uns_setup", L.ngap_uns_setup(h, opr.ctypes.data, int(maxit), float(rcvg), C.byref(o)))
        self.rows = int(L.ngap_uns_rows(h))

    def solve(self, d_rhs, d_sol):
        import torch
        torch.cuda.synchronize()
        it, rerr, aerr = C.c_int(), C.c_double(), C.c_double()
        check("ngap_uns_solve_device", lib().ngap_uns_solve_device(self._h, d_rhs.data_ptr(), d_sol.data_ptr(),
                                                                   C.byref(it), C.byref(rerr), C.byref(aerr)))
        return it.value, rerr.value, aerr.value

    def vcycle(self, d_b, d_x):
        import torch
        torch.cuda.synchronize()
        check("ngap_uns_vcycle_device", lib().ngap_uns_vcycle_device(self._h, d_b.data_ptr(), d_x.data_ptr()))

    def stats(self):
        s = UnsStats()
        check("ngap_uns_last_stats", lib().ngap_uns_last_stats(self._h, C.byref(s)))
        return s.as_dict()

    def destroy(self):
        if self._h is not None:
            lib().ngap_uns_destroy(self._h)
            self._h = None
