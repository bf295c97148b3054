"""CPU pins of the fourier3d oracle (oracle/fourier3d_ref.py), the checker of
libngapfft.so, and of the host-side checks of nga2_amd.fourier.

Each pin is chosen so that a plausible slip fails it:
- the transformed 7-point operator against the closed-form eigenvalues
  h sum_d (2 - 2 cos(2 pi k_d / n_d)) (a wrong offset placement or transform
  sign changes F);
- the 16^3 solution against the committed closed-form fixture;
- manufactured solutions x* on anisotropic, non-cubic and 2-D grids (a
  transposed axis or a wrong normalisation does not return x*);
- the reference's failure cases: non-uniform operator, vanishing modes,
  non-periodic or non-uniform grids.
"""
import os

import numpy as np
import pytest

from conftest import STC7
from nga2_amd import fourier, incomp, inputs
from nga2_amd.config import Config, uniform
from oracle import fourier3d_ref as F3
from oracle.orc import interior

HERE = os.path.dirname(os.path.abspath(__file__))


def _apply(cfg, opr, x):
    """A x over the interior with periodic wrap: sum_s opr(s) x(i + stc(s))."""
    it = cfg.interior
    X = x[it]
    O = opr[(slice(None),) + it]
    y = O[0] * X
    for s in range(1, 7):
        off = STC7[s]
        y = y + O[s] * np.roll(X, shift=tuple(-int(v) for v in off), axis=(0, 1, 2))
    return y


def _box(n, L, per=(True, True, True)):
    cfg = Config(uniform(n[0], L[0]), uniform(n[1], L[1]), uniform(n[2], L[2]), no=1, xper=per[0], yper=per[1],
                 zper=per[2])
    return cfg, incomp.laplacian(incomp.Metrics(cfg))


def test_transformed_operator_closed_form():
    P = inputs.periodic_gaussian(n=16)
    Finv = F3.setup(P.cfg, P.opr, STC7, fourier.circ_tol(P.cfg))
    F = 1.0 / Finv
    h = 1.0 / 16
    k = np.arange(16)
    lam1 = h * (2.0 - 2.0 * np.cos(2.0 * np.pi * k / 16))
    lam = lam1[:, None, None] + lam1[None, :, None] + lam1[None, None, :]
    lam[0, 0, 0] = 1.0
    assert np.abs(F - lam).max() <= 1e-15
    # anisotropic spacing: each direction carries vol / h_d^2
    cfg, opr = _box((12, 10, 6), (1.3, 0.7, 2.0))
    F = 1.0 / F3.setup(cfg, opr, STC7, fourier.circ_tol(cfg))
    hx, hy, hz = 1.3 / 12, 0.7 / 10, 2.0 / 6
    vol = hx * hy * hz
    lx = vol / hx ** 2 * (2 - 2 * np.cos(2 * np.pi * np.arange(12) / 12))
    ly = vol / hy ** 2 * (2 - 2 * np.cos(2 * np.pi * np.arange(10) / 10))
    lz = vol / hz ** 2 * (2 - 2 * np.cos(2 * np.pi * np.arange(6) / 6))
    lam = lx[:, None, None] + ly[None, :, None] + lz[None, None, :]
    lam[0, 0, 0] = 1.0
    assert np.abs(F - lam).max() <= 1e-13 * np.abs(lam).max()


def test_matches_closed_form_fixture():
    d = np.load(os.path.join(HERE, "golden", "fixtures.npz"))
    P = inputs.periodic_gaussian(n=16)
    assert np.array_equal(interior(P.rhs, P.cfg), d["periodic16_rhs"])
    Finv = F3.setup(P.cfg, P.opr, STC7, fourier.circ_tol(P.cfg))
    x = interior(F3.solve(P.cfg, Finv, P.rhs), P.cfg)
    ref = d["periodic16_sol"]
    assert np.abs(x - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("n,L,per", [((12, 10, 6), (1.3, 0.7, 2.0), (True, True, True)),
                                     ((16, 8, 1), (1.0, 0.5, 0.1), (True, True, False)),
                                     ((32, 1, 1), (2.0, 1.0, 1.0), (True, False, False)),
                                     ((8, 8, 8), (1.0, 1.0, 1.0), (True, True, True))])
def test_manufactured_solution(n, L, per):
    cfg, opr = _box(n, L, per)
    rng = np.random.default_rng(7)
    x = cfg.field()
    x[cfg.interior] = rng.standard_normal(n)
    x[cfg.interior] -= x[cfg.interior].mean()
    b = cfg.field()
    b[cfg.interior] = _apply(cfg, opr, x)
    Finv = F3.setup(cfg, opr, STC7, fourier.circ_tol(cfg))
    sol = F3.solve(cfg, Finv, b)
    got = sol[cfg.interior]
    assert np.abs(got - x[cfg.interior]).max() <= 1e-11 * np.abs(x).max()
    assert abs(got.mean()) <= 1e-13 * np.abs(x).max()
    # ghosts are the periodic images
    ref = sol.copy()
    cfg.sync(ref)
    assert np.array_equal(ref, sol)


def test_failure_cases():
    P = inputs.periodic_gaussian(n=8)
    tol = fourier.circ_tol(P.cfg)
    opr = P.opr.copy()
    opr[1, 3, 4, 5] += 10 * tol  # one coupling off by more than the tolerance
    with pytest.raises(F3.NotCirculant):
        F3.setup(P.cfg, opr, STC7, tol)
    opr = P.opr.copy()
    opr[1, 3, 4, 5] += 0.5 * tol  # within it: accepted
    F3.setup(P.cfg, opr, STC7, tol)
    # couplings in x only on a 3-D grid: every mode (0, ky, kz) vanishes
    opr = P.opr.copy()
    opr[3:] = 0.0
    opr[0] = -(opr[1] + opr[2])
    with pytest.raises(F3.Singular):
        F3.setup(P.cfg, opr, STC7, tol)
    # the constructor's grid checks
    with pytest.raises(ValueError):
        fourier.check_grid(inputs.channel_spectral(n=8).cfg)  # walls in y
    x = uniform(8)
    x[3] += 0.01
    with pytest.raises(ValueError):
        fourier.check_grid(Config(x, uniform(8), uniform(8), no=1))  # non-uniform x
    fourier.check_grid(Config(uniform(8), uniform(8), uniform(1), no=1, zper=False))  # one cell in z
    with pytest.raises(ValueError):
        fourier.check_grid(Config(uniform(1), uniform(1), uniform(1), no=1))  # no direction to transform


def test_min_meshsize_and_tolerance():
    cfg = Config(uniform(16, 2.0), uniform(8, 1.0), uniform(1, 0.5), no=1)
    # geometric mean over the directions with more than one cell (config_class.f90:112-125)
    assert fourier.min_meshsize(cfg) == pytest.approx((2.0 / 16 * 1.0 / 8) ** 0.5, rel=1e-15)
    assert fourier.circ_tol(cfg) == pytest.approx(6 * np.finfo(float).eps / (2.0 / 16 * 1.0 / 8) ** 2, rel=1e-14)
