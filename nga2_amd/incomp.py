"""Caller-side mirror of NGA2's projection step around the pressure solve.

This is the part of ``incomp`` / ``tpns`` that produces what the pressure
solver consumes (the 7-point operator ``opr`` and the right-hand side) and
what it does with the result (``shift_p``, ``get_pgrad``). It is host code
used to build solver inputs exactly as NGA2 would hand them over.

Reference:
  src/constant_density/incomp_class.f90:173-242   P-cell and face masks
  src/constant_density/incomp_class.f90:341-396   divp_*, divu_x, divv_y, divw_z
  src/constant_density/incomp_class.f90:516-566   masked metrics (walls -> zero)
  src/constant_density/incomp_class.f90:711-738   lower-dimensional metrics
  src/constant_density/incomp_class.f90:758-787   stencil map and opr = -vol div grad
  src/two_phase/tpns_class.f90:1565-1593          update_laplacian, / rho at faces
  src/constant_density/incomp_class.f90:1144-1159 get_div
  src/constant_density/incomp_class.f90:1163-1184 get_pgrad
  src/constant_density/incomp_class.f90:1720-1740 shift_p (tpns_class.f90:2271-2294)
  src/two_phase/tpns_class.f90:1616-1691          add_surface_tension_jump (CSF)
  src/two_phase/tpns_class.f90:1695-1716          get_pgrad with the jump
  src/constant_density/incomp_class.f90:1577-1716 get_mfr, correct_mfr (tpns_class.f90:2131-2267)
  examples/channel/src/simulation.f90:318-328     rhs scaling, velocity correction
  examples/falling_drop/src/simulation.f90:302-311 the same with face densities
"""
from __future__ import annotations

import numpy as np

from .config import Config

# incomp_class.f90:758-764: stencil entry -> (i,j,k) shift
STC7 = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], dtype=np.int32)


class Metrics:
    """Divergence metrics of the pressure operator with wall masks applied."""

    def __init__(self, cfg: Config):
        self.cfg = cfg
        nxo, nyo, nzo = cfg.shape
        no = cfg.no
        # --- P-cell mask (incomp_class.f90:173-191)
        mask = np.zeros(cfg.shape, dtype=np.int32, order="F")
        if not cfg.xper:
            mask[:no, :, :] = 2
            mask[no + cfg.nx:, :, :] = 2
        if not cfg.yper:
            mask[:, :no, :] = 2
            mask[:, no + cfg.ny:, :] = 2
        if not cfg.zper:
            mask[:, :, :no] = 2
            mask[:, :, no + cfg.nz:] = 2
        mask[cfg.VF == 0.0] = 1
        cfg.sync(mask)
        self.mask = mask
        # --- face masks (incomp_class.f90:194-241)
        self.umask = self._face_mask(0, cfg.xper, cfg.nx)
        self.vmask = self._face_mask(1, cfg.yper, cfg.ny)
        self.wmask = self._face_mask(2, cfg.zper, cfg.nz)
        # --- 1-D metrics broadcast to 3-D and masked (incomp_class.f90:358-396, 516-566)
        live = (mask == 0)
        self.divp = []  # per direction: (coefficient on lower face, coefficient on upper face)
        self.divu = []  # per direction: (coefficient on cell below face, cell above face)
        for ax, (di, dmi, fmask, n) in enumerate(((cfg.dxi, cfg.dxmi, self.umask, cfg.nx),
                                                  (cfg.dyi, cfg.dymi, self.vmask, cfg.ny),
                                                  (cfg.dzi, cfg.dzmi, self.wmask, cfg.nz))):
            shp = [1, 1, 1]
            shp[ax] = -1
            d = np.broadcast_to(di.reshape(shp), cfg.shape)
            p0 = np.where(live, -d, 0.0)
            p1 = np.where(live, d, 0.0)
            dm = np.nan_to_num(dmi, nan=0.0)
            dmb = np.broadcast_to(dm.reshape(shp), cfg.shape)
            fl = (fmask == 0)
            u0 = np.where(fl, -dmb, 0.0)   # divu(-1,:) multiplies P(i-1)
            u1 = np.where(fl, dmb, 0.0)    # divu( 0,:) multiplies P(i)
            if n == 1:  # incomp_class.f90:711-738
                p0 = np.zeros_like(p0); p1 = np.zeros_like(p1)
                u0 = np.zeros_like(u0); u1 = np.zeros_like(u1)
            self.divp.append((np.asfortranarray(p0), np.asfortranarray(p1)))
            self.divu.append((np.asfortranarray(u0), np.asfortranarray(u1)))

    def _face_mask(self, ax, per, n):
        cfg = self.cfg
        no = cfg.no
        fm = np.zeros(cfg.shape, dtype=np.int32, order="F")
        sl = [slice(None)] * 3
        if not per:
            sl[ax] = no          # imin
            fm[tuple(sl)] = 2
            sl[ax] = no + n      # imax+1
            fm[tuple(sl)] = 2
        vf = cfg.VF
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        dst = [slice(None)] * 3
        lo[ax] = slice(0, -1)
        hi[ax] = slice(1, None)
        dst[ax] = slice(1, None)
        wall = np.minimum(vf[tuple(lo)], vf[tuple(hi)]) == 0.0
        sub = fm[tuple(dst)]
        sub[wall] = 1
        fm[tuple(dst)] = sub
        cfg.sync(fm)
        if not per:
            a = [slice(None)] * 3
            b = [slice(None)] * 3
            a[ax] = 0
            b[ax] = 1
            fm[tuple(a)] = fm[tuple(b)]
        return fm


def _shift(A, ax, s):
    """A shifted so that out[i] = A[i+s] along axis ax (values outside are 0)."""
    out = np.zeros_like(A)
    n = A.shape[ax]
    src = [slice(None)] * 3
    dst = [slice(None)] * 3
    if s > 0:
        src[ax] = slice(s, n)
        dst[ax] = slice(0, n - s)
    else:
        src[ax] = slice(0, n + s)
        dst[ax] = slice(-s, n)
    out[tuple(dst)] = A[tuple(src)]
    return out


def laplacian(met: Metrics, rho_faces=None):
    """Build ``opr(7, imino:imaxo, jmino:jmaxo, kmino:kmaxo)`` over interior cells.

    ``opr = -vol * div(grad(.)/rho_face)`` with stencil order
    centre, +x, -x, +y, -y, +z, -z (incomp_class.f90:766-787). With
    ``rho_faces=(rho_U, rho_V, rho_W)`` each face term is divided by the
    face density (tpns_class.f90:1569-1590). Ghost entries are zero.
    """
    cfg = met.cfg
    opr = np.zeros((7,) + cfg.shape, dtype=np.float64, order="F")
    it = cfg.interior
    terms_c = []
    offd = []
    for ax in range(3):
        p0, p1 = met.divp[ax]
        u0, u1 = met.divu[ax]
        # divu(-1, i+1), divu(0, i+1) at the upper face of cell i
        u0_up = _shift(u0, ax, 1)
        u1_up = _shift(u1, ax, 1)
        if rho_faces is not None:
            rf = rho_faces[ax]
            # the shifted face density is 0 past the last ghost layer; those
            # entries are outside the interior and never stored
            rf_up = _shift(rf, ax, 1)
            rf_up[rf_up == 0.0] = 1.0
            c_up = p1 * u0_up / rf_up
            c_lo = p0 * u1 / rf
            o_up = p1 * u1_up / rf_up
            o_lo = p0 * u0 / rf
        else:
            c_up = p1 * u0_up
            c_lo = p0 * u1
            o_up = p1 * u1_up
            o_lo = p0 * u0
        terms_c += [c_up, c_lo]
        offd += [o_up, o_lo]
    diag = terms_c[0] + terms_c[1] + terms_c[2] + terms_c[3] + terms_c[4] + terms_c[5]
    vol = cfg.vol
    opr[0][it] = (-diag * vol)[it]
    for s in range(6):
        opr[1 + s][it] = (-offd[s] * vol)[it]
    return opr


def get_div(met: Metrics, U, V, W):
    """div = sum(divp_x*U(i:i+1)) + sum(divp_y*V(j:j+1)) + sum(divp_z*W(k:k+1))
    over interior cells (incomp_class.f90:1144-1159), then synced."""
    cfg = met.cfg
    div = cfg.field()
    it = cfg.interior
    terms = []
    for ax, F in enumerate((U, V, W)):
        p0, p1 = met.divp[ax]
        terms.append(p0 * F + p1 * _shift(F, ax, 1))
    d = terms[0] + terms[1] + terms[2]
    div[it] = d[it]
    cfg.sync(div)
    return div


def get_pgrad(met: Metrics, P):
    """Pgrad_x(i) = divu_x(-1,i)*P(i-1) + divu_x(0,i)*P(i) etc.
    (incomp_class.f90:1163-1184)."""
    cfg = met.cfg
    it = cfg.interior
    out = []
    for ax in range(3):
        u0, u1 = met.divu[ax]
        g = cfg.field()
        val = u0 * _shift(P, ax, -1) + u1 * P
        g[it] = val[it]
        cfg.sync(g)
        out.append(g)
    return out


def shift_p(cfg: Config, P):
    """Remove the vol*VF-weighted mean over fluid cells (incomp_class.f90:1720-1740)."""
    tot = cfg.integrate(P) / cfg.fluid_vol
    it = cfg.interior
    sub = P[it]
    sub[cfg.VF[it] > 0.0] -= tot
    P[it] = sub
    cfg.sync(P)
    return P


def convection_diffusion(met: Metrics, faces, rho, visc, dt):
    """A nonsymmetric 7-point operator of the form NGA2's implicit velocity
    solve hands to its ``implicit`` linsol (incomp_class.f90:1756-1803:
    opr(1) = rho plus half a time step of the convective and viscous
    couplings, per unit volume), written on the pressure cells of ``met``:

        A x = rho x + dt/2 * [ div(u_f * avg(x)) - visc * div(grad x) ]

    with central face averages, the face velocities ``faces`` = (U, V, W)
    (zeroed on wall faces) and the masked metrics of ``met``. It is a test
    operator with the structure and scaling of the implicit one (diagonal
    rho + O(dt), couplings -dt visc/h^2 +- dt u/(4h)), not a restatement of
    the staggered U/V/W momentum operators. Returns opr(7, ...) over the
    interior cells, ghost entries 0, stencil order as ``laplacian``.
    """
    cfg = met.cfg
    opr = np.zeros((7,) + cfg.shape, dtype=np.float64, order="F")
    it = cfg.interior
    fmasks = (met.umask, met.vmask, met.wmask)
    diag = np.full(cfg.shape, float(rho))
    offd = []
    h = 0.5 * dt
    for ax in range(3):
        p0, p1 = met.divp[ax]
        u0, u1 = met.divu[ax]
        uf = np.where(fmasks[ax] == 0, np.asarray(faces[ax], dtype=np.float64), 0.0)
        uf_up = _shift(uf, ax, 1)
        u0_up = _shift(u0, ax, 1)
        u1_up = _shift(u1, ax, 1)
        # convection: p1 * u_up * (x + x_up)/2 + p0 * u_lo * (x_lo + x)/2
        diag = diag + h * (0.5 * p1 * uf_up + 0.5 * p0 * uf)
        conv_up = h * 0.5 * p1 * uf_up
        conv_lo = h * 0.5 * p0 * uf
        # diffusion: -visc * (p1 * grad_up + p0 * grad_lo)
        diag = diag - h * visc * (p1 * u0_up + p0 * u1)
        offd += [conv_up - h * visc * p1 * u1_up, conv_lo - h * visc * p0 * u0]
    opr[0][it] = diag[it]
    for s in range(6):
        opr[1 + s][it] = offd[s][it]
    return opr


def rhs_scaled(cfg: Config, div, rho, dt):
    """psolv%rhs = -cfg%vol*div*rho/dt over every cell: Fortran evaluates the
    product left to right and negates the result
    (examples/channel/src/simulation.f90:318; falling_drop:302 is rho = 1)."""
    return -(((np.asarray(cfg.vol) * div) * rho) / dt)


def correct_velocity(dt, rho, grads, vel, rho_faces=None, P=None, sol=None):
    """U -= dt*Px/rho (constant density, channel/src/simulation.f90:325-327)
    or /rho_U (face densities, falling_drop/src/simulation.f90:308-310) over
    every cell; P += sol (:324, :307). Returns the new (U, V, W) and P."""
    out = []
    for ax in range(3):
        r = rho if rho_faces is None else rho_faces[ax]
        out.append(vel[ax] - (dt * grads[ax]) / r)
    Pn = None if P is None else P + sol
    return out, Pn


def get_pgrad_jump(met: Metrics, P, dPj):
    """tpns get_pgrad: Pgrad_x = sum(divu_x*P(i-1:i)) - dPjx on interior
    cells, then synced (tpns_class.f90:1695-1716)."""
    cfg = met.cfg
    it = cfg.interior
    out = []
    for ax in range(3):
        u0, u1 = met.divu[ax]
        g = cfg.field()
        val = (u0 * _shift(P, ax, -1) + u1 * P) - dPj[ax]
        g[it] = val[it]
        cfg.sync(g)
        out.append(g)
    return out


def csf_jump(met: Metrics, sigma, dt, SD, curv, VF, rho_faces, Pj, div):
    """add_surface_tension_jump without a contact model (tpns_class.f90:
    1616-1691). Pj: the previous jump (three face arrays). Returns the new
    jump, its increment DPj (new - previous, every cell) and the new div
    (interior cells += dt * sum(divp * DPj / rho_face))."""
    cfg = met.cfg
    no = cfg.no
    vol = np.asarray(cfg.vol)
    n = (cfg.nx, cfg.ny, cfg.nz)
    region = tuple(slice(no, no + n[a] + 1) for a in range(3))
    Pn, D = [], []
    for ax in range(3):
        u0, u1 = met.divu[ax]
        SDm, cm, vm, VFm = (_shift(A, ax, -1) for A in (SD, curv, vol, VF))
        surf = SDm * vm + SD * vol
        num = (SDm * cm) * vm + (SD * curv) * vol
        cv = np.zeros_like(surf)
        pos = surf > 0.0
        cv[pos] = num[pos] / surf[pos]
        val = (sigma * cv) * (u0 * VFm + u1 * VF)
        new = np.array(Pj[ax], dtype=np.float64, order="F", copy=True)
        new[region] = val[region]
        Pn.append(new)
        D.append(new - Pj[ax])
    terms = []
    for ax in range(3):
        p0, p1 = met.divp[ax]
        r = rho_faces[ax]
        with np.errstate(divide="ignore", invalid="ignore"):  # the last slab is outside the interior
            terms.append((p0 * D[ax]) / r + (p1 * _shift(D[ax], ax, 1)) / _shift(r, ax, 1))
    it = cfg.interior
    dn = np.array(div, dtype=np.float64, order="F", copy=True)
    dn[it] = div[it] + dt * ((terms[0][it] + terms[1][it]) + terms[2][it])
    return Pn, D, dn


def correct_mfr(cfg: Config, bconds, rho, vel):
    """get_mfr + correct_mfr (incomp_class.f90:1577-1716; tpns with rho = 1).
    bconds: dicts with face (0/1/2), rdir (+1/-1), can_correct, n_in and ijk
    (Fortran (i, j, k) rows; the first n_in are interior cells). Sums run in
    the reference's sequential order. Returns (new velocities, mfr_error)."""
    d = (cfg.dx, cfg.dy, cfg.dz)
    lo = (cfg.imino, cfg.jmino, cfg.kmino)
    mfr, carea = [], 0.0
    for bc in bconds:
        f = bc["face"]
        a1, a2 = ((1, 2), (2, 0), (0, 1))[f]
        F = vel[f]
        m = 0.0
        area = 0.0
        for t in range(bc["n_in"]):
            ijk = [int(bc["ijk"][t][c]) - lo[c] for c in range(3)]
            w1, w2 = float(d[a1][ijk[a1]]), float(d[a2][ijk[a2]])
            m = m + (((bc["rdir"] * rho) * float(F[tuple(ijk)])) * w1) * w2
            area = area + w1 * w2
        mfr.append(m)
        carea = carea + area * (1.0 if bc["can_correct"] else 0.0)
    err = 0.0
    for m in mfr:
        err = err + m
    out = [np.array(v, dtype=np.float64, order="F", copy=True) for v in vel]
    tiny = 10.0 * np.finfo(np.float64).eps
    if abs(err) < tiny or abs(carea) < tiny:
        return out, err
    vc = -err / (rho * carea)
    for bc in bconds:
        if not bc["can_correct"]:
            continue
        for t in range(len(bc["ijk"])):
            ijk = tuple(int(bc["ijk"][t][c]) - lo[c] for c in range(3))
            out[bc["face"]][ijk] = out[bc["face"]][ijk] + bc["rdir"] * vc
    return out, err
