"""ctypes binding of libngapsolve.so (the C ABI in include/ngapsolve.h).

The library is the product: if it is missing this module raises, there is no
fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libngapsolve.so")

# methods (hypre_str_class.f90:17-22)
PFMG, PCG_PFMG, GMRES_PFMG, SMG, PCG_SMG, GMRES_SMG = 1, 2, 3, 4, 5, 6
FP64, MIXED = 0, 1

OK = 0
ERR_ARG, ERR_STENCIL, ERR_NOT_INIT, ERR_NOT_SETUP, ERR_METHOD, ERR_UNSUPPORTED = -1, -2, -3, -4, -5, -6
ERR_NOT_CIRCULANT = -7
ERR_HIP, ERR_NONFINITE, ERR_BREAKDOWN, ERR_ALLOC, ERR_COARSE, ERR_STAGE, ERR_SINGULAR = 1, 2, 3, 4, 5, 6, 7
STAGE_GUESS, STAGE_PRECOND = 1, 2
STATUS_CONVERGED, STATUS_MAXIT, STATUS_STAGNATED = 0, 1, 2

K_SPMV, K_RBGS, K_RESTRICT, K_PROLONG, K_UPDATE, K_EXACT, K_SWEEP2, K_SWEEP2_ZERO = 0, 1, 2, 3, 4, 5, 6, 7

EXPORTS = [
    "ngap_default_opts", "ngap_create", "ngap_init", "ngap_setup", "ngap_setup_device", "ngap_solve",
    "ngap_solve_device", "ngap_set_tolerance", "ngap_set_stage", "ngap_residual_device", "ngap_apply_device", "ngap_destroy", "ngap_last_stats", "ngap_strerror",
]
TEST_EXPORTS = [
    "ngapt_nlevels", "ngapt_fail_alloc", "ngapt_level_info", "ngapt_level_coefs", "ngapt_apply", "ngapt_rbgs", "ngapt_resid_restrict",
    "ngapt_prolong_add", "ngapt_coarse_solve", "ngapt_vcycle", "ngapt_apply_exact", "ngapt_dot", "ngapt_time_kernel",
]


class Grid(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "nx", "ny", "nz", "imin_", "imax_", "jmin_", "jmax_", "kmin_", "kmax_",
        "imino_", "imaxo_", "jmino_", "jmaxo_", "kmino_", "kmaxo_", "xper", "yper", "zper")]


class Opts(C.Structure):
    _fields_ = [("precision", C.c_int), ("nu_pre", C.c_int), ("nu_post", C.c_int), ("coarsest_max", C.c_int),
                ("check_finite", C.c_int), ("verbose", C.c_int), ("graphs", C.c_int), ("kdim", C.c_int),
                ("skip_relax", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("status", C.c_int), ("it", C.c_int), ("nlevels", C.c_int), ("coarse_singular", C.c_int),
                ("precision", C.c_int), ("fallback", C.c_int), ("restarts", C.c_int), ("stage_calls", C.c_int),
                ("skip_relax", C.c_int), ("relaxed_levels", C.c_int),
                ("rerr", C.c_double), ("true_rerr", C.c_double), ("aerr", C.c_double),
                ("cnorm_ratio", C.c_double), ("asym", C.c_double), ("setup_ms", C.c_double),
                ("solve_ms", C.c_double), ("copy_ms", C.c_double), ("cells", C.c_longlong),
                ("device_bytes", C.c_longlong)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p
_L = None


def lib():
    """Load libngapsolve.so; raises if it has not been built."""
    global _L
    if _L is not None:
        return _L
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libngapsolve.so not found at {LIB_PATH}; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    L.ngap_default_opts.argtypes = [C.POINTER(Opts)]
    L.ngap_create.argtypes = [C.POINTER(_vp), C.POINTER(Grid), C.c_int, _ip, C.c_int]
    L.ngap_init.argtypes = [_vp]
    L.ngap_setup.argtypes = [_vp, _vp, C.c_int, C.c_double, C.c_int, C.POINTER(Opts)]
    L.ngap_setup_device.argtypes = [_vp, _vp, C.c_int, C.c_double, C.c_int, C.POINTER(Opts)]
    L.ngap_solve.argtypes = [_vp, _vp, _vp, _ip, _dp, _dp]
    L.ngap_solve_device.argtypes = [_vp, _vp, _vp, _ip, _dp, _dp]
    L.ngap_set_tolerance.argtypes = [_vp, C.c_int, C.c_double]
    L.ngap_set_stage.argtypes = [_vp, _vp, _vp, C.c_int, C.c_int]
    L.ngap_residual_device.argtypes = [_vp, _vp, _vp, _vp, _dp]
    L.ngap_apply_device.argtypes = [_vp, _vp, _vp]
    L.ngap_destroy.argtypes = [_vp]
    L.ngap_last_stats.argtypes = [_vp, C.POINTER(Stats)]
    L.ngap_strerror.argtypes = [C.c_int]
    L.ngap_strerror.restype = C.c_char_p
    L.ngapt_nlevels.argtypes = [_vp]
    L.ngapt_fail_alloc.argtypes = [_vp, C.c_int]
    L.ngapt_level_info.argtypes = [_vp, C.c_int, _ip, _ip]
    L.ngapt_level_coefs.argtypes = [_vp, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
    L.ngapt_apply.argtypes = [_vp, C.c_int, C.c_int, _dp, _dp]
    L.ngapt_rbgs.argtypes = [_vp, C.c_int, C.c_int, _dp, _dp, C.c_int]
    L.ngapt_resid_restrict.argtypes = [_vp, C.c_int, C.c_int, _dp, _dp, _dp]
    L.ngapt_prolong_add.argtypes = [_vp, C.c_int, C.c_int, _dp, _dp]
    L.ngapt_coarse_solve.argtypes = [_vp, C.c_int, _dp, _dp]
    L.ngapt_vcycle.argtypes = [_vp, C.c_int, _dp, _dp]
    L.ngapt_apply_exact.argtypes = [_vp, _dp, _dp]
    L.ngapt_dot.argtypes = [_vp, C.c_longlong, _dp, _dp, _dp]
    L.ngapt_time_kernel.argtypes = [_vp, C.c_int, C.c_int, C.c_int, _dp, _dp]
    _L = L
    return L


def strerror(code):
    return lib().ngap_strerror(int(code)).decode()


class NgapError(RuntimeError):
    def __init__(self, where, code):
        super().__init__(f"{where}: {strerror(code)} ({code})")
        self.code = code


def check(where, rc):
    if rc != 0:
        raise NgapError(where, rc)
    return rc


def default_opts(**kw):
    o = Opts()
    lib().ngap_default_opts(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, int(v))
    return o
