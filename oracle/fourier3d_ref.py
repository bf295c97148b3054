"""CPU restatement of NGA2's fourier3d linsol (numpy FFT, float64).

Test infrastructure only: the tests, smoke() and the bench's CPU baseline leg
use it as the checker of libngapfft.so; the product never imports it.

It follows src/solver/fourier3d_class.f90:
- setup (:1103-1166): circulance check against the first interior cell's
  stencil (:1114-1126), the stencil placed at its offsets modulo the grid
  (:1128-1137), the forward transform over the directions with more than one
  cell (:988-1043, FFTW_FORWARD = numpy's fftn sign), zero mode set to 1
  (:1143-1144), near-zero check |F| < 1000 eps (:1146-1152), inversion and the
  NaN check (:1154-1160);
- solve (:1170-1202): forward transform of the interior rhs with its zero mode
  set to 0 (:1039-1040), product with 1/F (:1186-1187), backward transform
  divided by n in each transformed direction (:1047-1099, numpy's ifftn
  normalisation), real part into the interior of sol, then the ghost fill.

Arrays follow nga2_amd.config: opr is (nst, nxo, nyo, nzo), rhs/sol (nxo, nyo,
nzo), interior slices cfg.interior.
"""
import numpy as np

EPS = np.finfo(np.float64).eps


class NotCirculant(ValueError):
    """fourier3d_class.f90:1126: operator must be uniform in space"""


class Singular(ValueError):
    """fourier3d_class.f90:1148-1152, 1156-1157: transformed operator near zero or NaN"""


def _axes(shape):
    return [d for d in range(3) if shape[d] > 1]


def setup(cfg, opr, stc, circ_tol):
    """1/F on the full (nx, ny, nz) spectrum (complex)."""
    it = cfg.interior
    n = (cfg.nx, cfg.ny, cfg.nz)
    sub = opr[(slice(None),) + it]
    ref = sub[:, 0, 0, 0].copy()
    if np.any(np.abs(sub - ref[:, None, None, None]) > circ_tol):
        raise NotCirculant("operator must be uniform in space")
    K = np.zeros(n, dtype=np.complex128)
    for s in range(len(ref)):
        K[stc[s][0] % n[0], stc[s][1] % n[1], stc[s][2] % n[2]] += ref[s]
    F = np.fft.fftn(K, axes=_axes(n)) if _axes(n) else K.copy()
    F[0, 0, 0] = 1.0
    if np.count_nonzero(np.abs(F) < 1000.0 * EPS):
        raise Singular("elements of transformed operator near zero")
    with np.errstate(divide="ignore", invalid="ignore"):
        Finv = 1.0 / F
    if np.isnan(np.abs(Finv)).any():
        raise Singular("elements of transformed operator are NaN")
    return Finv


def solve(cfg, Finv, rhs, sol=None):
    """sol (ghosted) with the interior from the transform and periodic ghosts."""
    it = cfg.interior
    n = (cfg.nx, cfg.ny, cfg.nz)
    ax = _axes(n)
    G = np.fft.fftn(rhs[it].astype(np.complex128), axes=ax) if ax else rhs[it].astype(np.complex128)
    G[0, 0, 0] = 0.0
    G = G * Finv
    x = np.fft.ifftn(G, axes=ax) if ax else G
    out = np.array(sol, dtype=np.float64, copy=True) if sol is not None else cfg.field()
    out[it] = x.real
    cfg.sync(out)
    return out
