"""Seeded synthetic inputs of the pressure-solve workloads.

Every workload builds the config, operator and right-hand side exactly as
NGA2's callers hand them to ``linsol`` (see ``incomp.py``), from fixed numpy
``Generator(PCG64(seed))`` streams. NGA2's own generator seeds from
/dev/urandom (src/libraries/random.f90:19-32), so reference runs are not
reproducible; these recipes are.

Workloads (sizes are parameters; the named defaults are the benchmark sizes):
  periodic_gaussian   n^3 triply periodic unit cube, rho = 1, face velocities
                      i.i.d. N(0,1) with per-component mean removed,
                      rhs = -vol*div*rho/dt (examples/hit/src/simulation.f90:81-117),
                      dt = 1, then the rhs mean removed. Default n = 64, seed 1001.
  channel_spectral    n^3 box [0,1]x[-1/2,1/2]x[-1/2,1/2], periodic in x and z,
                      walls in y, y tanh-stretched with stretch 2
                      (examples/channel/src/geometry.f90:36-38). Velocity: a
                      solenoidal Fourier field |k| <= 32, E(k) ~ k^4 exp(-2(k/8)^2),
                      random phases, plus 1% white noise, sampled on faces, V = 0
                      on the wall faces; rhs = -vol*div (rho = dt = 1), mean
                      removed. Default n = 256, seed 2002.
  spheres_two_phase   n^3 triply periodic unit cube; VF from 50 spheres of
                      radius U[0.05,0.15] (centres U[0,1)^3, periodic images),
                      rho_cell = VF*1000 + (1-VF)*1, face density = arithmetic mean
                      of the two cells; opr per tpns%update_laplacian
                      (src/two_phase/tpns_class.f90:1565-1593); rhs as
                      periodic_gaussian with seed 1002 and rhs = -vol*div/dt
                      (examples/falling_drop/src/simulation.f90:302). Spheres seed 3003.
  rhs_sequence        n^3 periodic (default 128), a sequence of correlated
                      right-hand sides: a scalar Fourier field |k| <= 16 with
                      spectrum k^4 exp(-2(k/8)^2) whose mode phases advance by
                      omega_k per step, omega_k ~ U[-pi/8, pi/8]; rhs = -vol*field,
                      mean removed. Seed 4004.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import incomp
from .config import Config, tanh_stretched, uniform


@dataclass
class Problem:
    cfg: Config
    opr: np.ndarray          # (7, nxo, nyo, nzo) Fortran order
    rhs: np.ndarray          # (nxo, nyo, nzo) Fortran order
    name: str
    info: dict = field(default_factory=dict)


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def _stc7():
    return incomp.STC7.copy()


def _remove_mean_interior(cfg, a):
    it = cfg.interior
    a[it] -= a[it].mean()
    return a


def _gaussian_faces(cfg, rng):
    """U, V, W ~ N(0,1) on interior faces (x-fastest draw order), per-component
    mean removed, ghosts synced."""
    fields = []
    for _ in range(3):
        v = rng.standard_normal(cfg.nx * cfg.ny * cfg.nz).reshape((cfg.nx, cfg.ny, cfg.nz), order="F")
        v -= v.mean()
        F = cfg.field()
        F[cfg.interior] = v
        cfg.sync(F)
        fields.append(F)
    return fields


def periodic_gaussian(n=64, seed=1001, no=1):
    x = uniform(n)
    cfg = Config(x, x, x, no=no, xper=True, yper=True, zper=True, name="periodic_gaussian")
    met = incomp.Metrics(cfg)
    opr = incomp.laplacian(met)
    U, V, W = _gaussian_faces(cfg, _rng(seed))
    div = incomp.get_div(met, U, V, W)
    rho, dt = 1.0, 1.0
    rhs = -cfg.vol * div * rho / dt
    rhs = np.asfortranarray(rhs)
    _remove_mean_interior(cfg, rhs)
    return Problem(cfg, opr, rhs, f"periodic_gaussian_{n}", dict(seed=seed))


# ------------------------------------------------------------------ spectral fields

def _spectrum_amp(kk, kmax):
    """sqrt of the 3-D spectral density E(k)/(4 pi k^2) with E(k) = k^4 exp(-2 (k/8)^2)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        E = kk ** 4 * np.exp(-2.0 * (kk / 8.0) ** 2)
        amp = np.sqrt(E / (4.0 * np.pi * kk ** 2))
    amp[(kk == 0) | (kk > kmax)] = 0.0
    return amp


def _eval_series(coef, xs, ys, zs, L=(1.0, 1.0, 1.0)):
    """Real part of sum_k coef[k] exp(2 pi i (kx x/Lx + ky y/Ly + kz z/Lz)) at the
    tensor grid xs x ys x zs; coef has shape (m,m,m) with k = -K..K per axis."""
    m = coef.shape[0]
    K = (m - 1) // 2
    k = np.arange(-K, K + 1)
    Ex = np.exp(2j * np.pi * np.outer(xs / L[0], k))   # (nx, m)
    Ey = np.exp(2j * np.pi * np.outer(ys / L[1], k))
    Ez = np.exp(2j * np.pi * np.outer(zs / L[2], k))
    t = np.tensordot(coef, Ez, axes=([2], [1]))        # (m, m, nz)
    t = np.tensordot(t, Ey, axes=([1], [1]))           # (m, nz, ny)
    t = np.tensordot(t, Ex, axes=([0], [1]))           # (nz, ny, nx)
    return np.asfortranarray(np.transpose(t.real, (2, 1, 0)))


def channel_spectral(n=256, seed=2002, stretch=2.0, kmax=32, noise=0.01, no=1, ny=None, nz=None):
    ny = n if ny is None else ny
    nz = n if nz is None else nz
    x = uniform(n, 1.0, 0.0)
    y = tanh_stretched(ny, 1.0, stretch)
    z = uniform(nz, 1.0, -0.5)
    cfg = Config(x, y, z, no=no, xper=True, yper=False, zper=True, name="channel_spectral")
    # walls (examples/channel/src/geometry.f90:56-60)
    cfg.VF[:] = 0.0
    cfg.VF[cfg.interior] = 1.0
    cfg.sync(cfg.VF)
    cfg.calc_fluid_vol()
    met = incomp.Metrics(cfg)
    opr = incomp.laplacian(met)
    rng = _rng(seed)
    m = 2 * kmax + 1
    kr = np.arange(-kmax, kmax + 1, dtype=np.float64)
    KX, KY, KZ = np.meshgrid(kr, kr, kr, indexing="ij")
    kk = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    amp = _spectrum_amp(kk, kmax)
    a = (rng.standard_normal((3, m, m, m)) + 1j * rng.standard_normal((3, m, m, m))) * amp
    with np.errstate(divide="ignore", invalid="ignore"):
        kdot = (KX * a[0] + KY * a[1] + KZ * a[2]) / np.where(kk > 0, kk ** 2, 1.0)
    a[0] -= KX * kdot
    a[1] -= KY * kdot
    a[2] -= KZ * kdot
    no_ = cfg.no
    xf = cfg.x[no_:no_ + cfg.nx]      # U faces x(imin..imax)
    xc = cfg.xm[no_:no_ + cfg.nx]
    yf = cfg.y[no_:no_ + cfg.ny]      # V faces y(jmin..jmax)
    yc = cfg.ym[no_:no_ + cfg.ny]
    zf = cfg.z[no_:no_ + cfg.nz]
    zc = cfg.zm[no_:no_ + cfg.nz]
    u = _eval_series(a[0], xf, yc, zc)
    v = _eval_series(a[1], xc, yf, zc)
    w = _eval_series(a[2], xc, yc, zf)
    rms = np.sqrt((np.mean(u ** 2) + np.mean(v ** 2) + np.mean(w ** 2)) / 3.0)
    fields = []
    for comp in (u, v, w):
        comp = comp + noise * rms * rng.standard_normal(comp.size).reshape(comp.shape, order="F")
        F = cfg.field()
        F[cfg.interior] = comp
        fields.append(F)
    U, V, W = fields
    V[:, no_, :] = 0.0                 # bottom wall face y(jmin)
    V[:, no_ + cfg.ny, :] = 0.0        # top wall face y(jmax+1)
    for F in (U, V, W):
        cfg.sync(F)
    div = incomp.get_div(met, U, V, W)
    rhs = np.asfortranarray(-cfg.vol * div)
    _remove_mean_interior(cfg, rhs)
    return Problem(cfg, opr, rhs, f"channel_spectral_{n}", dict(seed=seed, stretch=stretch, kmax=kmax))


# ------------------------------------------------------------------ spheres

def sphere_set(nsph, rmin, rmax, seed):
    rng = _rng(seed)
    centres = rng.random((nsph, 3))
    radii = rmin + (rmax - rmin) * rng.random(nsph)
    return centres, radii


def sphere_vf(cfg, centres, radii, nsub=4, device=None):
    """Liquid volume fraction of the union of spheres (periodic images) on a
    uniform periodic unit-cube grid.

    Cells whose centre lies farther than the half-diagonal from every sphere
    surface are classified exactly (the level set r - |x - c| is 1-Lipschitz);
    cut cells are sampled at nsub^3 sub-cell midpoints, the leaf rule of a
    log2(nsub)-level recursive refinement (cf. cube_refine_vol,
    src/libraries/mms_geom.f90:554-626)."""
    n = cfg.nx
    h = 1.0 / n
    hd = 0.5 * np.sqrt(3.0) * h
    xc = (np.arange(n) + 0.5) * h
    G = np.full((n, n, n), -np.inf)
    owner_cut = []
    for s, (c, r) in enumerate(zip(centres, radii)):
        lo = np.floor((c - r - 2 * hd) / h).astype(int)
        hi = np.ceil((c + r + 2 * hd) / h).astype(int)
        ii = np.arange(lo[0], hi[0] + 1) % n
        jj = np.arange(lo[1], hi[1] + 1) % n
        kk = np.arange(lo[2], hi[2] + 1) % n
        dx = (xc[ii] - c[0] + 0.5) % 1.0 - 0.5
        dy = (xc[jj] - c[1] + 0.5) % 1.0 - 0.5
        dz = (xc[kk] - c[2] + 0.5) % 1.0 - 0.5
        d = np.sqrt(dx[:, None, None] ** 2 + dy[None, :, None] ** 2 + dz[None, None, :] ** 2)
        g = r - d
        sub = G[np.ix_(ii, jj, kk)]
        G[np.ix_(ii, jj, kk)] = np.maximum(sub, g)
    VF = np.where(G > hd, 1.0, 0.0)
    cut = np.nonzero(np.abs(G) <= hd)
    ci, cj, ck = cut
    M = len(ci)
    if M and nsub >= 8 and (nsub & (nsub - 1)) == 0:
        VF[ci, cj, ck] = _cut_fraction_tree(xc[ci], xc[cj], xc[ck], centres, radii, h, hd, nsub, device=device)
    elif M:
        off = (np.arange(nsub) + 0.5) / nsub - 0.5
        ox, oy, oz = np.meshgrid(off, off, off, indexing="ij")
        ox, oy, oz = (o.ravel() * h for o in (ox, oy, oz))
        inside = np.zeros((M, nsub ** 3), dtype=bool)
        ccx, ccy, ccz = xc[ci], xc[cj], xc[ck]
        for c, r in zip(centres, radii):
            dx = (ccx - c[0] + 0.5) % 1.0 - 0.5
            dy = (ccy - c[1] + 0.5) % 1.0 - 0.5
            dz = (ccz - c[2] + 0.5) % 1.0 - 0.5
            dc = np.sqrt(dx * dx + dy * dy + dz * dz)
            cand = np.nonzero(np.abs(r - dc) <= hd)[0]
            if len(cand) == 0:
                continue
            for a0 in range(0, len(cand), 65536):
                idx = cand[a0:a0 + 65536]
                px = dx[idx, None] + ox[None, :]
                py = dy[idx, None] + oy[None, :]
                pz = dz[idx, None] + oz[None, :]
                inside[idx] |= (px * px + py * py + pz * pz) < r * r
        VF[ci, cj, ck] = inside.mean(axis=1)
    return VF


def _cut_fraction_tree(ccx, ccy, ccz, centres, radii, h, hd, nsub, device=None):
    """Fraction of the nsub^3 leaf midpoints of each cut cell that lie inside
    the union of spheres, found by recursive octree refinement as
    cube_refine_vol does (src/libraries/mms_geom.f90:554-626): a sub-cube
    whose level-set value r - |x - c| at its centre exceeds its half-diagonal
    in magnitude is wholly inside or outside (the level set is 1-Lipschitz),
    so only cut sub-cubes are split, down to the leaves (log2(nsub) levels).
    The leaves are classified by their midpoint with the same arithmetic as
    the uniform sampling of sphere_vf, so the result equals it exactly.
    Cells are processed in chunks; device: a torch device to run on (the
    same float64 arithmetic), or None for numpy."""
    M = len(ccx)
    # candidate spheres of every cut cell: those whose surface passes within hd
    # of its centre (the others cannot change the union inside the cell),
    # padded to K per cell with a sphere of radius 0 far away (never inside,
    # never the maximum)
    pc, pdx, pdy, pdz, pr = [], [], [], [], []
    for c, r in zip(centres, radii):
        dx = (ccx - c[0] + 0.5) % 1.0 - 0.5
        dy = (ccy - c[1] + 0.5) % 1.0 - 0.5
        dz = (ccz - c[2] + 0.5) % 1.0 - 0.5
        dc = np.sqrt(dx * dx + dy * dy + dz * dz)
        idx = np.nonzero(np.abs(r - dc) <= hd)[0]
        pc.append(idx)
        pdx.append(dx[idx])
        pdy.append(dy[idx])
        pdz.append(dz[idx])
        pr.append(np.full(len(idx), r))
    pc = np.concatenate(pc)
    order = np.argsort(pc, kind="stable")
    pc = pc[order]
    per = np.bincount(pc, minlength=M)
    K = max(1, int(per.max()) if M else 1)
    slot = np.arange(len(pc)) - np.repeat(np.cumsum(per) - per, per)
    DX = np.full((M, K), 1e3)
    DY = np.full((M, K), 1e3)
    DZ = np.full((M, K), 1e3)
    R = np.zeros((M, K))
    for A, parts in ((DX, pdx), (DY, pdy), (DZ, pdz), (R, pr)):
        A[pc, slot] = np.concatenate(parts)[order]
    if device is not None:
        import torch

        return _tree_torch(torch, device, DX, DY, DZ, R, h, nsub)
    out = np.zeros(M)
    sh = np.array([[a, b, c] for c in (0, 1) for b in (0, 1) for a in (0, 1)], dtype=np.int64)
    chunk = max(1, (1 << 22) // (nsub * nsub * 8 * K))
    for c0 in range(0, M, chunk):
        c1 = min(M, c0 + chunk)
        dX, dY, dZ, rR = DX[c0:c1], DY[c0:c1], DZ[c0:c1], R[c0:c1]
        count = np.zeros(c1 - c0, dtype=np.int64)
        cell = np.arange(c1 - c0)
        mx = np.zeros(c1 - c0, dtype=np.int64)
        my = np.zeros(c1 - c0, dtype=np.int64)
        mz = np.zeros(c1 - c0, dtype=np.int64)
        w = nsub
        while len(cell):
            if w == 1:
                px = dX[cell] + (((mx + 0.5) / nsub - 0.5) * h)[:, None]
                py = dY[cell] + (((my + 0.5) / nsub - 0.5) * h)[:, None]
                pz = dZ[cell] + (((mz + 0.5) / nsub - 0.5) * h)[:, None]
                ins = np.any((px * px + py * py + pz * pz) < rR[cell] * rR[cell], axis=1)
                np.add.at(count, cell[ins], 1)
                break
            px = dX[cell] + (((mx + 0.5 * w) / nsub - 0.5) * h)[:, None]
            py = dY[cell] + (((my + 0.5 * w) / nsub - 0.5) * h)[:, None]
            pz = dZ[cell] + (((mz + 0.5 * w) / nsub - 0.5) * h)[:, None]
            g = np.max(rR[cell] - np.sqrt(px * px + py * py + pz * pz), axis=1)
            hw = 0.5 * np.sqrt(3.0) * h * w / nsub * (1.0 + 1e-9)
            ins = g > hw
            np.add.at(count, cell[ins], w ** 3)
            keep = np.abs(g) <= hw
            cell, mx, my, mz = cell[keep], mx[keep], my[keep], mz[keep]
            w //= 2
            n = len(cell)
            cell = np.repeat(cell, 8)
            mx = np.repeat(mx, 8) + np.tile(sh[:, 0] * w, n)
            my = np.repeat(my, 8) + np.tile(sh[:, 1] * w, n)
            mz = np.repeat(mz, 8) + np.tile(sh[:, 2] * w, n)
        out[c0:c1] = count / float(nsub ** 3)
    return out


def _tree_torch(torch, device, DX, DY, DZ, R, h, nsub):
    """_cut_fraction_tree's refinement in torch float64 on a device."""
    M, K = DX.shape
    dev = torch.device(device)
    tX, tY, tZ, tR = (torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (DX, DY, DZ, R))
    sh = torch.tensor([[a, b, c] for c in (0, 1) for b in (0, 1) for a in (0, 1)], dtype=torch.int64, device=dev)
    out = torch.zeros(M, dtype=torch.float64, device=dev)
    chunk = max(1, (1 << 24) // (nsub * nsub * 8 * K))
    for c0 in range(0, M, chunk):
        c1 = min(M, c0 + chunk)
        dX, dY, dZ, rR = tX[c0:c1], tY[c0:c1], tZ[c0:c1], tR[c0:c1]
        count = torch.zeros(c1 - c0, dtype=torch.int64, device=dev)
        cell = torch.arange(c1 - c0, device=dev)
        mx = torch.zeros(c1 - c0, dtype=torch.int64, device=dev)
        my = torch.zeros_like(mx)
        mz = torch.zeros_like(mx)
        w = nsub
        while cell.numel():
            if w == 1:
                px = dX[cell] + (((mx + 0.5) / nsub - 0.5) * h)[:, None]
                py = dY[cell] + (((my + 0.5) / nsub - 0.5) * h)[:, None]
                pz = dZ[cell] + (((mz + 0.5) / nsub - 0.5) * h)[:, None]
                ins = torch.any((px * px + py * py + pz * pz) < rR[cell] * rR[cell], dim=1)
                count.index_add_(0, cell[ins], torch.ones_like(cell[ins]))
                break
            px = dX[cell] + (((mx + 0.5 * w) / nsub - 0.5) * h)[:, None]
            py = dY[cell] + (((my + 0.5 * w) / nsub - 0.5) * h)[:, None]
            pz = dZ[cell] + (((mz + 0.5 * w) / nsub - 0.5) * h)[:, None]
            g = torch.max(rR[cell] - torch.sqrt(px * px + py * py + pz * pz), dim=1).values
            hw = 0.5 * np.sqrt(3.0) * h * w / nsub * (1.0 + 1e-9)
            ins = g > hw
            count.index_add_(0, cell[ins], torch.full_like(cell[ins], w ** 3))
            keep = torch.abs(g) <= hw
            cell, mx, my, mz = cell[keep], mx[keep], my[keep], mz[keep]
            w //= 2
            n = cell.numel()
            cell = cell.repeat_interleave(8)
            mx = mx.repeat_interleave(8) + (sh[:, 0] * w).repeat(n)
            my = my.repeat_interleave(8) + (sh[:, 1] * w).repeat(n)
            mz = mz.repeat_interleave(8) + (sh[:, 2] * w).repeat(n)
        out[c0:c1] = count.to(torch.float64) / float(nsub ** 3)
    return out.cpu().numpy()


def face_densities(cfg, vf, rho_l, rho_g):
    """Cell density VF rho_l + (1 - VF) rho_g (periodic ghosts synced), then the
    arithmetic mean of the two cells on each face: rho_U(i) = (rho(i-1) + rho(i)) / 2
    (tpns_class.f90:2610-2611 with even sub-cell volumes). Faces without a lower
    neighbour in the array keep 1."""
    rho = cfg.field()
    rho[cfg.interior] = vf * rho_l + (1.0 - vf) * rho_g
    cfg.sync(rho)
    rho_faces = []
    for ax in range(3):
        rf = cfg.field(1.0)
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax] = slice(0, -1)
        hi[ax] = slice(1, None)
        rf[tuple(hi)] = 0.5 * (rho[tuple(lo)] + rho[tuple(hi)])
        rho_faces.append(rf)
    return rho_faces


def spheres_two_phase(n=256, seed=3003, vel_seed=1002, nsph=50, rmin=0.05, rmax=0.15, rho_l=1000.0, rho_g=1.0,
                      nsub=4, no=1, device=None):
    """nsub = 16 is the four-level refinement of falling_drop (amr_ref_lvl = 4,
    examples/falling_drop/src/simulation.f90:91,112), the configuration-3
    recipe (CONFIG3_NSUB); small test grids keep nsub = 4. device: a torch
    device for the cut-cell refinement (same result)."""
    x = uniform(n)
    cfg = Config(x, x, x, no=no, xper=True, yper=True, zper=True, name="spheres_two_phase")
    met = incomp.Metrics(cfg)
    centres, radii = sphere_set(nsph, rmin, rmax, seed)
    vf = sphere_vf(cfg, centres, radii, nsub=nsub, device=device)
    rho_faces = face_densities(cfg, vf, rho_l, rho_g)
    opr = incomp.laplacian(met, rho_faces=rho_faces)
    U, V, W = _gaussian_faces(cfg, _rng(vel_seed))
    div = incomp.get_div(met, U, V, W)
    dt = 1.0
    rhs = np.asfortranarray(-cfg.vol * div / dt)
    _remove_mean_interior(cfg, rhs)
    return Problem(cfg, opr, rhs, f"spheres_two_phase_{n}",
                   dict(seed=seed, vel_seed=vel_seed, liquid_fraction=float(vf.mean()), nsph=nsph))


# ------------------------------------------------------------------ correlated sequence

class RhsSequence:
    """Correlated right-hand sides on an n^3 periodic grid (see module doc)."""

    def __init__(self, n=128, seed=4004, kmax=16, no=1):
        x = uniform(n)
        self.cfg = Config(x, x, x, no=no, xper=True, yper=True, zper=True, name="rhs_sequence")
        self.met = incomp.Metrics(self.cfg)
        self.opr = incomp.laplacian(self.met)
        self.n = n
        rng = _rng(seed)
        k = np.fft.fftfreq(n, d=1.0 / n)
        KX, KY, KZ = np.meshgrid(k, k, k, indexing="ij")
        kk = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
        amp = _spectrum_amp(kk, kmax)
        self.coef = (rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))) * amp
        self.omega = rng.uniform(-np.pi / 8, np.pi / 8, size=(n, n, n))
        self.vol = self.cfg.vol[self.cfg.interior]

    def field(self, step):
        c = self.coef * np.exp(1j * self.omega * step)
        f = np.fft.ifftn(c).real * self.n ** 3
        return f  # (nx, ny, nz) cell-centred

    def rhs(self, step):
        r = self.cfg.field()
        r[self.cfg.interior] = -self.vol * self.field(step)
        _remove_mean_interior(self.cfg, r)
        return r

    def rhs_device(self, step, device="cuda:0"):
        """rhs(step) computed with torch's FFT on the device: the ghosted
        array flattened in Fortran order (the layout of ngap_solve_device)."""
        import torch

        if getattr(self, "_dev", None) != str(device):
            self._dcoef = torch.from_numpy(self.coef).to(device)
            self._domega = torch.from_numpy(self.omega).to(device)
            self._dvol = torch.from_numpy(np.ascontiguousarray(self.vol)).to(device)
            self._dev = str(device)
        c = self._dcoef * torch.exp(1j * self._domega * float(step))
        f = torch.fft.ifftn(c).real * self.n ** 3
        a = -self._dvol * f  # [x, y, z]
        a = a - a.mean()
        cfg = self.cfg
        g = torch.zeros((cfg.nzo, cfg.nyo, cfg.nxo), dtype=torch.float64, device=device)
        no = cfg.no
        g[no:no + cfg.nz, no:no + cfg.ny, no:no + cfg.nx] = a.permute(2, 1, 0)
        return g.reshape(-1)


# the four-level sub-cell refinement of configuration 3's volume fraction
CONFIG3_NSUB = 16


def spheres_two_phase_config3(n=256, **kw):
    """Configuration 3 as specified: spheres_two_phase with the four-level
    refinement (nsub = 16), the cut cells refined on the GPU when torch sees
    one (the same result as on the host, see _cut_fraction_tree)."""
    device = None
    try:
        import torch

        if torch.cuda.is_available():
            device = "cuda:0"
    except ImportError:
        pass
    kw.setdefault("nsub", CONFIG3_NSUB)
    return spheres_two_phase(n=n, device=device, **kw)


WORKLOADS = {
    "periodic_gaussian": periodic_gaussian,
    "channel_spectral": channel_spectral,
    "spheres_two_phase": spheres_two_phase_config3,
}


def implicit_problem(cfg, seed=4004, rho=1.0, visc=0.02, dt=None, cfl=1.0):
    """A nonsymmetric implicit-solve test system on cfg (incomp.convection_diffusion):
    Gaussian face velocities (ghosts synced) scaled to |u|max = 1, dt from the convective
    CFL number (dt = cfl * h_min / |u|max), and a Gaussian right-hand side.
    Seeded with numpy PCG64(seed)."""
    rng = _rng(seed)
    met = incomp.Metrics(cfg)
    faces = [np.asfortranarray(rng.standard_normal(cfg.shape)) for _ in range(3)]
    for f in faces:
        cfg.sync(f)  # periodic images, as NGA2's velocity arrays hold them
    umax = max(float(np.abs(f[cfg.interior]).max()) for f in faces)
    faces = [f / umax for f in faces]
    h = min(float(np.min(cfg.dx[cfg.interior[0]])), float(np.min(cfg.dy[cfg.interior[1]])),
            float(np.min(cfg.dz[cfg.interior[2]])))
    if dt is None:
        dt = cfl * h
    opr = incomp.convection_diffusion(met, faces, rho, visc, dt)
    rhs = cfg.field()
    rhs[cfg.interior] = rng.standard_normal((cfg.nx, cfg.ny, cfg.nz))
    return Problem(cfg, opr, rhs, f"implicit_{cfg.nx}x{cfg.ny}x{cfg.nz}", dict(seed=seed, rho=rho, visc=visc, dt=dt))


def masked_spheres(n=128, seed=5005, nsph=40, rmin=0.04, rmax=0.12, no=1):
    """Immersed solids as masked cells, the case hypre_uns serves: a periodic
    unit cube n^3 with nsph solid spheres (centres and radii seeded as in
    sphere_set); a cell whose centre lies inside a sphere (periodic images)
    has VF = 0, so the masked metrics (incomp_class.f90:173-191, :516-566) give
    it opr = 0 and its fluid neighbours a homogeneous Neumann face. The
    right-hand side is N(0, 1) on the fluid cells with its mean removed over
    each connected fluid region (the operator's null space), zero elsewhere."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components

    from . import incomp

    cfg = Config(uniform(n), uniform(n), uniform(n), no=no)
    centres, radii = sphere_set(nsph, rmin, rmax, seed)
    h = 1.0 / n
    xc = (np.arange(n) + 0.5) * h
    solid = np.zeros((n, n, n), dtype=bool)
    for c, r in zip(centres, radii):
        dx = (xc - c[0] + 0.5) % 1.0 - 0.5
        dy = (xc - c[1] + 0.5) % 1.0 - 0.5
        dz = (xc - c[2] + 0.5) % 1.0 - 0.5
        solid |= dx[:, None, None] ** 2 + dy[None, :, None] ** 2 + dz[None, None, :] ** 2 < r * r
    it = cfg.interior
    cfg.VF[it] = np.where(solid, 0.0, 1.0)
    cfg.sync(cfg.VF)
    opr = incomp.laplacian(incomp.Metrics(cfg))
    rng = _rng(seed + 1)
    rhs = cfg.field()
    act = np.abs(opr[0][it]) > 10.0 * np.finfo(np.float64).eps
    vals = np.where(act, rng.standard_normal((n, n, n)), 0.0)
    # connected fluid regions through the nonzero couplings (periodic)
    idx = np.arange(n ** 3).reshape((n, n, n), order="F")
    rows, cols = [], []
    for s, ax in ((1, 0), (3, 1), (5, 2)):
        cpl = opr[s][it] != 0.0
        rows.append(idx[cpl & act])
        cols.append(np.roll(idx, -1, axis=ax)[cpl & act])
    G = sp.coo_matrix((np.ones(sum(len(r) for r in rows)), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n ** 3, n ** 3))
    ncomp, lab = connected_components(G, directed=False)
    lab = lab.reshape((n, n, n), order="F")
    for c in np.unique(lab[act]):
        m = act & (lab == c)
        vals[m] -= vals[m].mean()
    rhs[it] = vals
    return Problem(cfg, opr, rhs, f"masked_spheres_{n}", dict(seed=seed, nsph=nsph, rmin=rmin, rmax=rmax,
                                                              solid_cells=int(solid.sum()), regions=len(np.unique(lab[act]))))
