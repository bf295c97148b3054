"""CPU pins of the host restatement of the projection step around the solve
(nga2_amd/incomp.py), each against a property the mathematics fixes:

- rhs scaling + pressure gradient + velocity correction: the corrected
  velocity is discretely divergence-free (projection), for constant density
  and for face densities with ratio 1000 (channel/src/simulation.f90:315-328,
  falling_drop/src/simulation.f90:298-311);
- the CSF jump: with a constant curvature and SD > 0 everywhere, the face
  jump is sigma*kappa times the face gradient of VF (tpns_class.f90:1633-1663),
  and the divergence increment is dt times get_div of the jump increment when
  rho = 1 (:1682-1688);
- get_pgrad with the jump reduces to get_pgrad when the increment is 0;
- correct_mfr: after the correction the flow-rate mismatch is zero
  (incomp_class.f90:1669-1716), and nothing changes when it already is.
"""
import numpy as np
import scipy.sparse.linalg as spl

from conftest import assemble_csr, direct_solve_singular
from nga2_amd import incomp, inputs
from nga2_amd.config import Config, tanh_stretched, uniform


def _solve(cfg, opr, rhs):
    A = assemble_csr(cfg, opr)
    b = rhs[cfg.interior].ravel(order="F")
    x = direct_solve_singular(A, b - b.mean())
    P = cfg.field()
    P[cfg.interior] = x.reshape((cfg.nx, cfg.ny, cfg.nz), order="F")
    cfg.sync(P)
    return P


def _projection(cfg, rho_faces, rho):
    met = incomp.Metrics(cfg)
    rng = np.random.default_rng(11)
    # face velocities, zero on wall faces (the boundary condition NGA2 keeps there)
    U, V, W = (np.where(m == 0, rng.standard_normal(cfg.shape), 0.0) for m in (met.umask, met.vmask, met.wmask))
    for A in (U, V, W):
        cfg.sync(A)
    dt = 0.37
    opr = incomp.laplacian(met, rho_faces=rho_faces)
    div = incomp.get_div(met, U, V, W)
    rhs = incomp.rhs_scaled(cfg, div, rho if rho_faces is None else 1.0, dt)
    P = _solve(cfg, opr, rhs)
    grads = incomp.get_pgrad(met, P)
    (Un, Vn, Wn), _ = incomp.correct_velocity(dt, rho, grads, (U, V, W), rho_faces=rho_faces)
    for A in (Un, Vn, Wn):
        cfg.sync(A)
    d0 = np.abs(div[cfg.interior]).max()
    d1 = np.abs(incomp.get_div(met, Un, Vn, Wn)[cfg.interior]).max()
    return d0, d1


def test_projection_is_divergence_free_constant_density():
    cfg = Config(uniform(12), tanh_stretched(10, 1.0, 2.0), uniform(8), no=1, yper=False)
    d0, d1 = _projection(cfg, None, 1.7)
    assert d1 <= 1e-10 * d0, (d0, d1)


def test_projection_is_divergence_free_two_phase():
    P = inputs.spheres_two_phase(n=12, nsph=3, rmin=0.15, rmax=0.3)
    cfg = P.cfg
    centres, radii = inputs.sphere_set(3, 0.15, 0.3, 3003)
    rho_faces = inputs.face_densities(cfg, inputs.sphere_vf(cfg, centres, radii), 1000.0, 1.0)
    d0, d1 = _projection(cfg, rho_faces, 1.0)
    assert d1 <= 1e-9 * d0, (d0, d1)


def test_csf_jump_closed_form():
    cfg = Config(uniform(10), uniform(8), uniform(6), no=2)
    met = incomp.Metrics(cfg)
    rng = np.random.default_rng(5)
    VF = rng.uniform(0, 1, cfg.shape)
    SD = rng.uniform(0.5, 2.0, cfg.shape)
    kappa, sigma, dt = 3.5, 0.07, 0.2
    curv = np.full(cfg.shape, kappa)
    ones = tuple(np.ones(cfg.shape) for _ in range(3))
    Pj0 = [rng.standard_normal(cfg.shape) for _ in range(3)]
    div0 = rng.standard_normal(cfg.shape)
    Pj, D, div = incomp.csf_jump(met, sigma, dt, SD, curv, VF, ones, Pj0, div0)
    no = cfg.no
    dmi = (cfg.dxmi, cfg.dymi, cfg.dzmi)
    n = (cfg.nx, cfg.ny, cfg.nz)
    for ax in range(3):
        sl = tuple(slice(no, no + n[a] + 1) for a in range(3))
        grad = np.diff(VF, axis=ax, prepend=np.nan)
        shp = [1, 1, 1]
        shp[ax] = -1
        expect = sigma * kappa * grad * dmi[ax].reshape(shp)
        assert np.allclose(Pj[ax][sl], expect[sl], rtol=1e-12, atol=1e-14)
        # outside the recomputed faces the jump keeps its value and the increment is 0
        mask = np.ones(cfg.shape, bool)
        mask[sl] = False
        assert np.array_equal(Pj[ax][mask], Pj0[ax][mask])
        assert np.all(D[ax][mask] == 0.0)
        assert np.array_equal(D[ax], Pj[ax] - Pj0[ax])
    it = cfg.interior
    ref = incomp.get_div(met, *D)
    assert np.allclose(div[it], div0[it] + dt * ref[it], rtol=1e-13, atol=1e-13)


def test_pgrad_jump_reduces_to_pgrad():
    cfg = Config(uniform(9), tanh_stretched(7, 1.0, 2.0), uniform(5), no=1, yper=False)
    met = incomp.Metrics(cfg)
    P = np.random.default_rng(2).standard_normal(cfg.shape)
    zero = [np.zeros(cfg.shape)] * 3
    a = incomp.get_pgrad(met, P)
    b = incomp.get_pgrad_jump(met, P, zero)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    dpj = [np.full(cfg.shape, 0.25 * (ax + 1)) for ax in range(3)]
    c = incomp.get_pgrad_jump(met, P, dpj)
    for ax in range(3):
        assert np.allclose(c[ax][cfg.interior], a[ax][cfg.interior] - 0.25 * (ax + 1), rtol=0, atol=1e-14)


def _duct_bconds(cfg):
    """Inflow through x = imin (face x, rdir -1, fixed) and outflow through
    x = imax+1 (face x, rdir +1, correctable), as NGA2's dirichlet /
    clipped_neumann pair (examples/bctester/src/simulation.f90:104-105)."""
    out = []
    for i, rdir, cc in ((cfg.imin, -1, False), (cfg.imax + 1, 1, True)):
        inner = [(i, j, k) for k in range(cfg.kmin, cfg.kmax + 1) for j in range(cfg.jmin, cfg.jmax + 1)]
        over = [(i, j, k) for k in range(cfg.kmino, cfg.kmaxo + 1) for j in range(cfg.jmino, cfg.jmaxo + 1)
                if not (cfg.jmin <= j <= cfg.jmax and cfg.kmin <= k <= cfg.kmax)]
        out.append(dict(face=0, rdir=rdir, can_correct=cc, n_in=len(inner), ijk=np.array(inner + over)))
    return out


def test_correct_mfr_removes_mismatch():
    cfg = Config(uniform(8), tanh_stretched(6, 1.0, 2.0), uniform(5), no=1, xper=False, yper=False, zper=True)
    bcs = _duct_bconds(cfg)
    rng = np.random.default_rng(9)
    vel = [rng.standard_normal(cfg.shape) for _ in range(3)]
    rho = 2.5
    new, err = incomp.correct_mfr(cfg, bcs, rho, vel)
    assert abs(err) > 1e-3
    _, err2 = incomp.correct_mfr(cfg, bcs, rho, new)
    assert abs(err2) <= 1e-13 * max(1.0, abs(err))
    # only the correctable outlet's cells changed, by rdir * correction
    diff = new[0] - vel[0]
    cells = bcs[1]["ijk"]
    lo = (cfg.imino, cfg.jmino, cfg.kmino)
    idx = tuple(cells[:, c] - lo[c] for c in range(3))
    vc = diff[idx]
    assert np.allclose(vc, vc[0]) and abs(vc[0]) > 0
    mask = np.ones(cfg.shape, bool)
    mask[idx] = False
    assert np.all(diff[mask] == 0.0)
    assert np.array_equal(new[1], vel[1]) and np.array_equal(new[2], vel[2])
