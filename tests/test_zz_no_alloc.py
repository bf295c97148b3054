"""The allocation-free mean removal (ngap_proj_shift_p_noalloc, include/
ngapproj.h) against the handle-based entry point it varies
(ngap_proj_shift_p): the same result bits, on two calls in a row, the
second on a smaller grid, through the C ABI."""
import ctypes as C

import numpy as np
import pytest
import torch

from nga2_amd import _lib, incomp, inputs
from nga2_amd.config import Config, tanh_stretched, uniform
from nga2_amd.linsol import grid_of
from nga2_amd.proj import DeviceProjection

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))).to(DEV)


def _grids():
    big = Config(uniform(30), tanh_stretched(22, 1.0, 2.0), uniform(17), no=2, xper=True, yper=False, zper=True)
    big.VF[:] = 1.0
    big.VF[6:12, 5:9, 4:8] = 0.0
    big.sync(big.VF)
    big.calc_fluid_vol()
    small = inputs.periodic_gaussian(n=8).cfg
    return [("walls_hole_30x22x17", big), ("periodic8", small)]


def test_noalloc_matches_shift_p():
    L = _lib.lib()
    L.ngap_proj_shift_p_noalloc.argtypes = [C.POINTER(_lib.Grid), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_void_p]
    for name, cfg in _grids():
        print(f"\n[noalloc {name}]", flush=True)
        met = incomp.Metrics(cfg)
        dp = DeviceProjection(met)
        rng = np.random.default_rng(len(name))
        P = rng.standard_normal(cfg.shape) + 2.0
        d_ref = _dev(P)
        mean_ref = dp.shift_p(d_ref)
        d_P = _dev(P)
        d_vol = _dev(cfg.vol)
        d_vf = _dev(cfg.VF)
        d_mean = torch.full((1,), 7.0, dtype=torch.float64, device=DEV)
        torch.cuda.synchronize()
        rc = L.ngap_proj_shift_p_noalloc(C.byref(grid_of(cfg)), C.c_void_p(d_vol.data_ptr()),
                                         C.c_void_p(d_vf.data_ptr()), C.c_void_p(d_P.data_ptr()),
                                         C.c_void_p(d_mean.data_ptr()), None)
        assert rc == 0, _lib.strerror(rc)
        got = d_P.cpu().numpy()
        ref = d_ref.cpu().numpy()
        assert np.array_equal(got, ref), np.abs(got - ref).max()
        assert float(d_mean.cpu()[0]) == mean_ref
        # and the host restatement, to the tolerance of its different summation order
        host = incomp.shift_p(cfg, P.copy()).ravel(order="F")
        assert np.abs(got - host).max() <= 1e-14 * np.abs(host).max()
        dp.destroy()
    # misuse is refused before anything is launched
    assert L.ngap_proj_shift_p_noalloc(None, None, None, None, None, None) == _lib.ERR_ARG
