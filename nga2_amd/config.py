"""Serial mirror of NGA2's structured grid and config objects.

Only what the pressure solve touches is restated: vertex/centroid positions,
1-D metrics, ghost extension, cell volumes, the wall field VF, ghost-cell
synchronisation (periodic wrap on one rank) and the VF-weighted integral used
by ``shift_p``.

Reference:
  src/grid/sgrid_class.f90:151-159   index bounds (imin=1, imino=imin-no, ...)
  src/grid/sgrid_class.f90:166-189   ghost extension of the vertex mesh
  src/grid/sgrid_class.f90:192-257   xm, dx, dxm, dxi, dxmi
  src/config/config_class.f90:103-111  vol = dx*dy*dz
  src/config/config_class.f90:196-214  integrate (sum of vol*VF*A over interior)
  src/config/config_class.f90:176-192  calc_fluid_vol
  src/grid/pgrid_class.f90:439-518   rsync ghost fill (periodic wrap on 1 rank)

Arrays use Fortran memory order (``order='F'``) and 0-based numpy indices; the
numpy index ``a`` maps to the Fortran index ``imino + a``. Interior cells are
``no:no+nx``.
"""
from __future__ import annotations

import numpy as np


class Config:
    """One-rank config: a structured grid with ``no`` ghost layers."""

    def __init__(self, x, y, z, no=1, xper=True, yper=True, zper=True, name="config"):
        x = np.asarray(x, dtype=np.float64)
        y = np.asarray(y, dtype=np.float64)
        z = np.asarray(z, dtype=np.float64)
        if no < 0:
            raise ValueError("overlap size cannot be negative")
        self.name = name
        self.no = int(no)
        self.xper, self.yper, self.zper = bool(xper), bool(yper), bool(zper)
        self.nx, self.ny, self.nz = len(x) - 1, len(y) - 1, len(z) - 1
        # Fortran bounds (one rank: local == global)
        self.imin, self.jmin, self.kmin = 1, 1, 1
        self.imax, self.jmax, self.kmax = self.nx, self.ny, self.nz
        self.imino, self.jmino, self.kmino = 1 - no, 1 - no, 1 - no
        self.imaxo, self.jmaxo, self.kmaxo = self.nx + no, self.ny + no, self.nz + no
        self.nxo, self.nyo, self.nzo = self.nx + 2 * no, self.ny + 2 * no, self.nz + 2 * no
        self.x, self.xm, self.dx, self.dxi, self.dxm, self.dxmi = _metrics(x, no)
        self.y, self.ym, self.dy, self.dyi, self.dym, self.dymi = _metrics(y, no)
        self.z, self.zm, self.dz, self.dzi, self.dzm, self.dzmi = _metrics(z, no)
        for name_, arr in (("x", self.dx), ("y", self.dy), ("z", self.dz)):
            if np.min(arr) <= 0.0:
                raise ValueError(f"mesh in {name_} is not monotonically increasing")
        self.vol = np.asfortranarray(self.dx[:, None, None] * self.dy[None, :, None] * self.dz[None, None, :])
        self.VF = np.ones(self.shape, dtype=np.float64, order="F")
        self.fluid_vol = float(np.sum(self.vol[self.interior]))

    # ------------------------------------------------------------------ shapes
    @property
    def shape(self):
        return (self.nxo, self.nyo, self.nzo)

    @property
    def interior(self):
        no = self.no
        return (slice(no, no + self.nx), slice(no, no + self.ny), slice(no, no + self.nz))

    @property
    def periodic(self):
        return (self.xper, self.yper, self.zper)

    def field(self, value=0.0):
        return np.full(self.shape, value, dtype=np.float64, order="F")

    # ------------------------------------------------------------------ sync
    def sync(self, A):
        """Fill ghost cells: periodic wrap in periodic directions; in
        non-periodic directions the outer ghosts are left as they are, as the
        one-rank ``pgrid%rsync`` does (pgrid_class.f90:459-516)."""
        no = self.no
        if no == 0:
            return A
        for ax, (n, per) in enumerate(((self.nx, self.xper), (self.ny, self.yper), (self.nz, self.zper))):
            if not per:
                continue
            sl_lo_dst = [slice(None)] * 3
            sl_lo_src = [slice(None)] * 3
            sl_hi_dst = [slice(None)] * 3
            sl_hi_src = [slice(None)] * 3
            sl_lo_dst[ax] = slice(0, no)
            sl_lo_src[ax] = slice(n, n + no)
            sl_hi_dst[ax] = slice(n + no, n + 2 * no)
            sl_hi_src[ax] = slice(no, 2 * no)
            if n < no:
                # fewer interior cells than ghosts: wrap cell by cell
                for g in range(no):
                    d_lo = [slice(None)] * 3
                    d_hi = [slice(None)] * 3
                    s_lo = [slice(None)] * 3
                    s_hi = [slice(None)] * 3
                    d_lo[ax] = g
                    s_lo[ax] = no + ((g - no) % n)
                    d_hi[ax] = n + no + g
                    s_hi[ax] = no + (g % n)
                    A[tuple(d_lo)] = A[tuple(s_lo)]
                    A[tuple(d_hi)] = A[tuple(s_hi)]
                continue
            A[tuple(sl_lo_dst)] = A[tuple(sl_lo_src)]
            A[tuple(sl_hi_dst)] = A[tuple(sl_hi_src)]
        return A

    # ------------------------------------------------------------------ integrals
    def calc_fluid_vol(self):
        """config_class.f90:176-192"""
        self.fluid_vol = float(np.sum((self.vol * self.VF)[self.interior]))
        return self.fluid_vol

    def integrate(self, A):
        """config_class.f90:196-214: sum of vol*VF*A over interior cells."""
        return float(np.sum((self.vol * self.VF * A)[self.interior]))


def _metrics(v, no):
    """Ghost-extended vertex array and 1-D metrics (sgrid_class.f90:166-257).

    Returns arrays indexed 0..nxo(+1) for Fortran indices imino..imaxo(+1):
    x (nxo+1), xm (nxo), dx (nxo), dxi (nxo), dxm (nxo; entry 0 is NaN since
    Fortran dxm starts at imino+1), dxmi (nxo; entry 0 NaN).
    """
    n = len(v) - 1
    x = np.empty(n + 2 * no + 1)
    x[no:no + n + 1] = v
    h_lo = v[1] - v[0]
    h_hi = v[n] - v[n - 1]
    for g in range(no - 1, -1, -1):
        x[g] = x[g + 1] - h_lo
    for g in range(no + n + 1, n + 2 * no + 1):
        x[g] = x[g - 1] + h_hi
    xm = 0.5 * (x[:-1] + x[1:])
    dx = x[1:] - x[:-1]
    dxi = 1.0 / dx
    dxm = np.full(n + 2 * no, np.nan)
    dxm[1:] = xm[1:] - xm[:-1]
    dxmi = np.full(n + 2 * no, np.nan)
    dxmi[1:] = 1.0 / dxm[1:]
    return x, xm, dx, dxi, dxm, dxmi


def uniform(n, L=1.0, x0=0.0):
    """n+1 uniform vertices on [x0, x0+L] (examples/channel/src/geometry.f90:41-43 form)."""
    return np.array([x0 + i / n * L for i in range(n + 1)], dtype=np.float64)


def tanh_stretched(n, L=1.0, stretch=2.0):
    """y(j) = 0.5 L tanh(s (2 (j-1)/n - 1)) / tanh(s), j = 1..n+1
    (examples/channel/src/geometry.f90:36-38)."""
    j = np.arange(n + 1, dtype=np.float64)
    return 0.5 * L * np.tanh(stretch * (2.0 * j / n - 1.0)) / np.tanh(stretch)
