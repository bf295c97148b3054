"""CPU checks of the boundary: the shared library loads and exports every
symbol the headers declare; host-side mirrors (config, assembly) hold the
invariants of NGA2's formulas. No kernel is launched here."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from nga2_amd import _lib, incomp, inputs
from nga2_amd.config import Config, tanh_stretched, uniform


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(ngapt?_\w+)\s*\(", txt, re.M)))


def test_library_exports_all_declared_symbols():
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared("ngapsolve.h") + _declared("ngapsolve_test.h")
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(_lib.EXPORTS) <= set(names)
    assert set(_lib.TEST_EXPORTS) <= set(names)


def test_net_library_exports_all_declared_symbols():
    from nga2_amd import unet

    txt = open(os.path.join(ROOT, "include", "ngapnet.h")).read()
    names = sorted(set(re.findall(r"^\s*int\s+(ngapnet_\w+)\s*\(", txt, re.M)))
    assert len(names) >= 18
    L = unet.lib()  # loads and sets every argtypes entry without touching the device
    for n in names:
        assert hasattr(L, n), n
        assert getattr(L, n).argtypes is not None, n
    # pure host query: voxel-tile groups of the weight-gradient split
    assert L.ngapnet_wgrad_parts(16, 4, 4, 32, 32) == 1
    assert L.ngapnet_wgrad_parts(256, 256, 256, 32, 32) == 256


def test_strerror_and_defaults_without_gpu():
    assert _lib.strerror(0) == "ok"
    assert "setup" in _lib.strerror(_lib.ERR_NOT_SETUP)
    o = _lib.default_opts()
    assert (o.precision, o.nu_pre, o.nu_post, o.check_finite, o.kdim) == (0, 0, 0, 1, 5)


def test_create_rejects_bad_arguments_without_gpu():
    """ngap_create validates before touching the device."""
    cfg = Config(uniform(8), uniform(8), uniform(8), no=1)
    from nga2_amd.linsol import grid_of

    g = grid_of(cfg)
    h = ctypes.c_void_p()
    stc = np.asfortranarray(incomp.STC7)
    L = _lib.lib()
    assert L.ngap_create(ctypes.byref(h), ctypes.byref(g), 0, stc.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), 2) \
        == _lib.ERR_ARG
    g.imax_ = 5  # a partial box: this build holds the whole grid on one GPU
    assert L.ngap_create(ctypes.byref(h), ctypes.byref(g), 7, stc.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), 2) \
        == _lib.ERR_ARG
    assert L.ngap_solve(None, None, None, None, None, None) == _lib.ERR_ARG


def test_metrics_match_sgrid_formulas():
    y = tanh_stretched(12, 1.0, 2.0)
    cfg = Config(uniform(8), y, uniform(4), no=2, xper=True, yper=False, zper=True)
    no = 2
    # ghost extension by the first/last interior spacing (sgrid_class.f90:166-189)
    assert np.isclose(cfg.y[0], y[0] - 2 * (y[1] - y[0]))
    assert np.isclose(cfg.y[-1], y[-1] + 2 * (y[-1] - y[-2]))
    assert np.allclose(cfg.dy[no:no + 12], np.diff(y))
    assert np.allclose(cfg.dymi[1:], 1.0 / np.diff(cfg.ym))
    assert np.allclose(cfg.vol[2, :, 3], cfg.dx[2] * cfg.dy * cfg.dz[3])


def test_operator_closed_form_uniform():
    """Uniform periodic grid: diagonal 6h, couplings -h (vol = h^3)."""
    n = 16
    P = inputs.periodic_gaussian(n=n)
    it = P.cfg.interior
    op = P.opr[(slice(None),) + it]
    h = 1.0 / n
    assert np.allclose(op[0], 6 * h, rtol=1e-13)
    assert np.allclose(op[1:], -h, rtol=1e-13)


def test_operator_invariants_channel_and_spheres():
    for P in (inputs.channel_spectral(n=16), inputs.spheres_two_phase(n=16, nsph=4, rmin=0.15, rmax=0.3)):
        cfg = P.cfg
        it = cfg.interior
        op = P.opr[(slice(None),) + it]
        # row sums vanish (periodic/Neumann), centre positive, couplings negative
        assert np.abs(op.sum(axis=0)).max() <= 1e-13 * op[0].max()
        assert np.all(op[0] > 0)
        assert np.all(op[1:] <= 0)
        # symmetry opr(2,i) == opr(3,i+1) (x periodic in both workloads)
        assert np.allclose(op[1], np.roll(op[2], -1, axis=0), rtol=1e-12, atol=0)
        # rhs mean removed
        assert abs(P.rhs[it].sum()) <= 1e-12 * np.abs(P.rhs[it]).sum()
    # channel: Neumann walls in y -> no coupling across the wall
    P = inputs.channel_spectral(n=16)
    op = P.opr[(slice(None),) + P.cfg.interior]
    assert np.all(op[4][:, 0, :] == 0.0) and np.all(op[3][:, -1, :] == 0.0)


def test_shift_p_and_sync():
    P = inputs.periodic_gaussian(n=8)
    cfg = P.cfg
    x = cfg.field()
    x[cfg.interior] = np.random.default_rng(0).standard_normal((8, 8, 8)) + 3.0
    incomp.shift_p(cfg, x)
    assert abs(cfg.integrate(x)) < 1e-12
    # periodic ghosts equal the wrapped interior
    assert np.array_equal(x[0, 1:-1, 1:-1], x[8, 1:-1, 1:-1])
    assert np.array_equal(x[9, 1:-1, 1:-1], x[1, 1:-1, 1:-1])


def test_divergence_of_gradient_is_operator():
    """-vol * div(grad p) from get_div(get_pgrad(p)) equals opr applied to p
    (incomp_class.f90:766-787 vs :1144-1184)."""
    P = inputs.channel_spectral(n=12)
    cfg = P.cfg
    met = incomp.Metrics(cfg)
    p = cfg.field()
    p[cfg.interior] = np.random.default_rng(2).standard_normal((12, 12, 12))
    cfg.sync(p)
    gx, gy, gz = incomp.get_pgrad(met, p)
    div = incomp.get_div(met, gx, gy, gz)
    lhs = (-cfg.vol * div)[cfg.interior]
    from oracle.orc import Oracle, interior

    o = Oracle.from_cfg(cfg, P.opr)
    Ap = o.apply_exact(interior(p, cfg))
    assert np.allclose(np.transpose(lhs, (2, 1, 0)).ravel(), Ap, rtol=1e-10, atol=1e-12 * np.abs(Ap).max())


def test_rhs_sequence_torch_matches_numpy():
    """RhsSequence.rhs_device (torch FFT) restates rhs (numpy FFT); run on the CPU here."""
    seq = inputs.RhsSequence(n=32, seed=7, kmax=8)
    for step in (0, 5):
        ref = seq.rhs(step).ravel(order="F")
        got = seq.rhs_device(step, "cpu").numpy()
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_temporal_predictor_recovers_a_rotation():
    """Increments of p_n = cos(w n) v + sin(w n) u obey d_n = 2 cos(w) d_{n-1} - d_{n-2}:
    the least-squares fit with m = 2 returns (2 cos w, -1) and extrapolates exactly."""
    import torch
    from nga2_amd import learned

    g = torch.Generator().manual_seed(0)
    v, u = torch.randn(500, generator=g, dtype=torch.float64), torch.randn(500, generator=g, dtype=torch.float64)
    w = 0.3
    sols = [np.cos(w * n) * v + np.sin(w * n) * u for n in range(30)]
    alpha = learned.fit_predictor(sols, 2)
    assert np.allclose(alpha, [2 * np.cos(w), -1.0], rtol=0, atol=1e-10)
    x = learned.extrapolate(sols[-3:], alpha)
    exact = np.cos(w * 30) * v + np.sin(w * 30) * u
    assert float(torch.max(torch.abs(x - exact))) <= 1e-9


def test_projection_exports():
    txt = open(os.path.join(ROOT, "include", "ngapproj.h")).read()
    names = sorted(set(re.findall(r"^\s*int\s+(ngap_proj_\w+)\s*\(", txt, re.M)))
    from nga2_amd import proj

    L = proj._bind(ctypes.CDLL(_lib.LIB_PATH))
    assert set(names) == set(proj.EXPORTS)
    for n in names:
        assert hasattr(L, n), n


def test_face_densities_arithmetic_mean():
    cfg = Config(uniform(6), uniform(5), uniform(4), no=1, xper=True, yper=True, zper=True)
    rng = np.random.default_rng(0)
    vf = rng.uniform(0, 1, (6, 5, 4))
    ru, rv, rw = inputs.face_densities(cfg, vf, 1000.0, 1.0)
    rho = vf * 1000.0 + (1 - vf) * 1.0
    # interior face between cells (i-1, i) along x, with the periodic image at i = 0
    assert np.isclose(ru[1 + 2, 1 + 3, 1 + 1], 0.5 * (rho[1, 3, 1] + rho[2, 3, 1]))
    assert np.isclose(ru[1 + 0, 1 + 0, 1 + 0], 0.5 * (rho[5, 0, 0] + rho[0, 0, 0]))
    assert np.isclose(rw[1 + 1, 1 + 2, 1 + 3], 0.5 * (rho[1, 2, 2] + rho[1, 2, 3]))


def test_fourier_exports_and_host_checks():
    """libngapfft.so loads, exports what ngapfft.h declares, and its create
    rejects bad grids before touching the device."""
    from nga2_amd import fourier
    from nga2_amd.linsol import grid_of

    txt = open(os.path.join(ROOT, "include", "ngapfft.h")).read()
    names = sorted(set(re.findall(r"^\s*int\s+(ngap_fft_\w+)\s*\(", txt, re.M)))
    assert set(names) == set(fourier.EXPORTS)
    L = fourier.lib()
    for n in names:
        assert hasattr(L, n), n
        assert getattr(L, n).argtypes is not None, n
    h = ctypes.c_void_p()
    stc = np.asfortranarray(incomp.STC7)
    sp = stc.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    cfg = Config(uniform(8), uniform(8), uniform(8), no=1, yper=False)  # walls in y
    assert L.ngap_fft_create(ctypes.byref(h), ctypes.byref(grid_of(cfg)), 7, sp) == _lib.ERR_UNSUPPORTED
    g = grid_of(Config(uniform(8), uniform(8), uniform(8), no=1))
    assert L.ngap_fft_create(ctypes.byref(h), ctypes.byref(g), 0, sp) == _lib.ERR_ARG
    g.imax_ = 5
    assert L.ngap_fft_create(ctypes.byref(h), ctypes.byref(g), 7, sp) == _lib.ERR_ARG
    assert L.ngap_fft_solve(None, None, None) == _lib.ERR_ARG
    assert "uniform" in _lib.strerror(_lib.ERR_NOT_CIRCULANT)
    assert "near zero" in _lib.strerror(_lib.ERR_SINGULAR)


def test_convection_diffusion_operator():
    """The implicit-solve test operator: with zero velocities each interior
    row sums to rho (diffusion conserves constants, walls included), and on a
    uniform grid the operator is symmetric; a velocity field adds couplings
    that are antisymmetric between neighbours (central convection). Per unit
    volume, as NGA2's implicit operator, so a stretched grid is not
    symmetric."""
    import scipy.sparse as sp
    from conftest import assemble_csr
    from nga2_amd.config import tanh_stretched

    for cfg, sym in ((Config(uniform(10), tanh_stretched(8), uniform(6), no=1, yper=False), False),
                     (Config(uniform(10), uniform(8), uniform(6), no=1, yper=False, zper=False), True)):
        met = incomp.Metrics(cfg)
        zero = [np.zeros(cfg.shape)] * 3
        opr = incomp.convection_diffusion(met, zero, 2.0, 0.1, 0.3)
        A = assemble_csr(cfg, opr)
        assert np.allclose(np.asarray(A.sum(axis=1)).ravel(), 2.0, rtol=0, atol=1e-12)
        asym = abs(A - A.T).max()
        assert (asym <= 1e-12 * abs(A).max()) == sym
        if not sym:
            continue
        rng = np.random.default_rng(0)
        faces = [np.asfortranarray(rng.standard_normal(cfg.shape)) for _ in range(3)]
        for f in faces:
            cfg.sync(f)
        C = assemble_csr(cfg, incomp.convection_diffusion(met, faces, 2.0, 0.1, 0.3)) - A
        Coff = C - sp.diags(C.diagonal())
        assert abs(Coff + Coff.T).max() <= 1e-12 * abs(Coff).max()
        assert abs(Coff).max() > 0.0


def test_uns_library_exports_all_declared_symbols():
    from nga2_amd import uns

    txt = open(os.path.join(ROOT, "include", "ngapuns.h")).read()
    names = sorted(set(re.findall(r"^\s*(?:int|long long|void|const ngap_amg \*)\s*(ngap_(?:uns|amg)_\w+)\s*\(", txt,
                                  re.M)))
    assert len(names) >= 19
    L = uns.lib()  # loads and sets every argtypes entry without touching the device
    for n in names:
        assert hasattr(L, n), n
        assert getattr(L, n).argtypes is not None, n
    assert set(uns.EXPORTS) == set(names)
