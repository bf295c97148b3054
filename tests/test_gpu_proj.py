"""GPU parity of the device-side projection (include/ngapproj.h) with the
host restatement of NGA2's formulas (nga2_amd/incomp.py, nga2_amd/config.py).

The kernels evaluate each expression in the reference's order without FMA
contraction, so the following are compared bit for bit:
- opr, with and without face densities;
- the divergence;
- the pressure gradient;
- the ghost fill.
The mean removal of shift_p sums in a different (fixed) order, so it is held
to 1e-14 of max|P|.
"""
import numpy as np
import pytest
import torch

from nga2_amd import incomp, inputs
from nga2_amd.config import Config, tanh_stretched, uniform
from nga2_amd.proj import DeviceProjection

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))).to(DEV)


def _host(t, shape):
    return t.cpu().numpy().reshape(shape, order="F")


def _walls_with_hole(n=(14, 11, 9)):
    """Walls in y, periodic x and z, an internal VF=0 block (mask 1 cells and faces)."""
    cfg = Config(uniform(n[0]), tanh_stretched(n[1], 1.0, 2.0), uniform(n[2]), no=2, xper=True, yper=False,
                 zper=True)
    cfg.VF[:] = 1.0
    cfg.VF[6:9, 5:8, 4:6] = 0.0
    cfg.sync(cfg.VF)
    cfg.calc_fluid_vol()
    return cfg


def _cases():
    P = inputs.spheres_two_phase(n=24, nsph=6, rmin=0.1, rmax=0.25)
    rng = np.random.default_rng(3)
    # face densities of a density field with ratio 1000 (arithmetic faces)
    vf = rng.uniform(0, 1, (24, 24, 24))
    rho_faces = inputs.face_densities(P.cfg, vf, 1000.0, 1.0)
    return [
        ("periodic16", inputs.periodic_gaussian(n=16).cfg, None),
        ("channel16", inputs.channel_spectral(n=16).cfg, None),
        ("spheres24_rho", P.cfg, rho_faces),
        ("walls_hole", _walls_with_hole(), None),
        ("thin16x16x1", Config(uniform(16), uniform(16), uniform(1), no=1, xper=True, yper=True, zper=True), None),
    ]


@pytest.mark.parametrize("idx", range(5))
def test_projection_matches_host(idx):
    name, cfg, rho_faces = _cases()[idx]
    print(f"\n[proj {name}]", flush=True)
    met = incomp.Metrics(cfg)
    dp = DeviceProjection(met)
    shape = cfg.shape
    No = int(np.prod(shape))
    rng = np.random.default_rng(idx)

    # operator
    ref = incomp.laplacian(met, rho_faces=rho_faces)
    d_opr = torch.zeros(7 * No, dtype=torch.float64, device=DEV)
    d_rho = None if rho_faces is None else tuple(_dev(r) for r in rho_faces)
    dp.laplacian(d_opr, d_rho)
    got = d_opr.cpu().numpy().reshape((7,) + shape, order="F")
    assert np.array_equal(got, ref), np.abs(got - ref).max()

    # divergence
    U, V, W = (rng.standard_normal(shape) for _ in range(3))
    ref = incomp.get_div(met, U, V, W)
    d_div = torch.full((No,), 7.0, dtype=torch.float64, device=DEV)
    dp.div(_dev(U), _dev(V), _dev(W), d_div)
    assert np.array_equal(_host(d_div, shape), ref)

    # pressure gradient
    Pf = rng.standard_normal(shape)
    refs = incomp.get_pgrad(met, Pf)
    outs = [torch.full((No,), 7.0, dtype=torch.float64, device=DEV) for _ in range(3)]
    dp.pgrad(_dev(Pf), *outs)
    for ax in range(3):
        assert np.array_equal(_host(outs[ax], shape), refs[ax]), ax

    # ghost fill
    A = rng.standard_normal(shape)
    d_A = _dev(A)
    dp.sync(d_A)
    ref = A.copy()
    cfg.sync(ref)
    assert np.array_equal(_host(d_A, shape), ref)

    # mean removal
    Pm = rng.standard_normal(shape) + 3.0
    d_P = _dev(Pm)
    mean = dp.shift_p(d_P)
    ref = incomp.shift_p(cfg, Pm.copy())
    assert abs(mean - cfg.integrate(Pm) / cfg.fluid_vol) <= 1e-14 * abs(mean)
    assert np.abs(_host(d_P, shape) - ref).max() <= 1e-14 * np.abs(ref).max()
    dp.destroy()


def test_device_operator_feeds_the_solver():
    """Assemble the two-phase operator on the device, set up from it
    (ngap_setup_device) and solve: same iterations and solution as the host
    operator path."""
    print("\n[proj -> solver]", flush=True)
    from nga2_amd._lib import MIXED
    from nga2_amd.linsol import DeviceSolve

    P = inputs.spheres_two_phase(n=32, nsph=8, rmin=0.1, rmax=0.2)
    cfg = P.cfg
    centres, radii = inputs.sphere_set(8, 0.1, 0.2, 3003)
    rho_faces = inputs.face_densities(cfg, inputs.sphere_vf(cfg, centres, radii), 1000.0, 1.0)
    met = incomp.Metrics(cfg)
    ref_opr = incomp.laplacian(met, rho_faces=rho_faces)
    dp = DeviceProjection(met)
    d_opr = torch.zeros(7 * int(np.prod(cfg.shape)), dtype=torch.float64, device=DEV)
    dp.laplacian(d_opr, tuple(_dev(r) for r in rho_faces))
    d_rhs = _dev(P.rhs)
    sols = []
    for opr in (d_opr, _dev(ref_opr)):
        ds = DeviceSolve(cfg)
        ds.setup(opr, maxit=200, rcvg=1e-10, precision=MIXED)
        d_sol = torch.zeros_like(d_rhs)
        it, _ = ds.solve(d_rhs, d_sol)
        st = ds.stats()
        assert st["true_rerr"] <= 1e-10
        sols.append((it, d_sol.cpu().numpy()))
        ds.destroy()
    assert sols[0][0] == sols[1][0]
    assert np.array_equal(sols[0][1], sols[1][1])
    dp.destroy()


@pytest.mark.parametrize("idx", range(5))
def test_step_pieces_match_host(idx):
    """rhs scaling, velocity correction, the two-phase gradient, the CSF jump:
    bit for bit against the host restatement (same expression order, no FMA)."""
    name, cfg, rho_faces = _cases()[idx]
    print(f"\n[step pieces {name}]", flush=True)
    met = incomp.Metrics(cfg)
    dp = DeviceProjection(met)
    shape = cfg.shape
    No = int(np.prod(shape))
    rng = np.random.default_rng(100 + idx)
    if rho_faces is None:
        rho_faces = tuple(rng.uniform(1.0, 1000.0, shape) for _ in range(3))
    d_rf = tuple(_dev(r) for r in rho_faces)
    div = rng.standard_normal(shape)
    for rho, dt in ((1.0, 0.013), (998.2, 2.5e-3)):
        d_rhs = torch.full((No,), 7.0, dtype=torch.float64, device=DEV)
        dp.rhs(_dev(div), rho, dt, d_rhs)
        assert np.array_equal(_host(d_rhs, shape), incomp.rhs_scaled(cfg, div, rho, dt))
    grads = [rng.standard_normal(shape) for _ in range(3)]
    vel = [rng.standard_normal(shape) for _ in range(3)]
    P = rng.standard_normal(shape)
    sol = rng.standard_normal(shape)
    for faces in (None, rho_faces):
        dvel = [_dev(v) for v in vel]
        dP = _dev(P)
        dp.correct(0.01, 1.3, [_dev(g) for g in grads], dvel, rho_faces=None if faces is None else d_rf, P=dP,
                   sol=_dev(sol))
        ref, Pn = incomp.correct_velocity(0.01, 1.3, grads, vel, rho_faces=faces, P=P, sol=sol)
        for ax in range(3):
            assert np.array_equal(_host(dvel[ax], shape), ref[ax]), ax
        assert np.array_equal(_host(dP, shape), Pn)
    dPj = [rng.standard_normal(shape) for _ in range(3)]
    outs = [torch.full((No,), 7.0, dtype=torch.float64, device=DEV) for _ in range(3)]
    dp.pgrad_jump(_dev(P), [_dev(a) for a in dPj], outs)
    refs = incomp.get_pgrad_jump(met, P, dPj)
    for ax in range(3):
        assert np.array_equal(_host(outs[ax], shape), refs[ax]), ax
    # CSF: an SD that vanishes on part of the cells (mycurv = 0 branch)
    VF = rng.uniform(0, 1, shape)
    SD = np.where(rng.uniform(0, 1, shape) < 0.4, 0.0, rng.uniform(0.1, 3.0, shape))
    curv = rng.standard_normal(shape) * 5.0
    Pj0 = [rng.standard_normal(shape) for _ in range(3)]
    d_Pj = [_dev(a) for a in Pj0]
    d_D = [torch.full((No,), 7.0, dtype=torch.float64, device=DEV) for _ in range(3)]
    d_div = _dev(div)
    dp.csf_jump(0.072, 1e-3, _dev(SD), _dev(curv), _dev(VF), d_rf, d_Pj, d_D, d_div)
    Pj, D, dn = incomp.csf_jump(met, 0.072, 1e-3, SD, curv, VF, rho_faces, Pj0, div)
    for ax in range(3):
        assert np.array_equal(_host(d_Pj[ax], shape), Pj[ax]), ax
        assert np.array_equal(_host(d_D[ax], shape), D[ax]), ax
    assert np.array_equal(_host(d_div, shape), dn)
    dp.destroy()


def test_correct_mfr_matches_host():
    """correct_mfr through an inflow/outflow pair: the mismatch to 1e-13 of the
    host's sequential sum, the corrected velocities to 1e-13, and a second
    call finds no mismatch left."""
    from test_projection_host import _duct_bconds

    print("\n[correct_mfr]", flush=True)
    cfg = Config(uniform(20), tanh_stretched(12, 1.0, 2.0), uniform(9), no=2, xper=False, yper=False, zper=True)
    met = incomp.Metrics(cfg)
    dp = DeviceProjection(met)
    bcs = _duct_bconds(cfg)
    dp.set_bconds(bcs)
    rng = np.random.default_rng(21)
    vel = [rng.standard_normal(cfg.shape) for _ in range(3)]
    for rho in (1.0, 3.7):
        dvel = [_dev(v) for v in vel]
        err = dp.correct_mfr(rho, dvel)
        ref, ref_err = incomp.correct_mfr(cfg, bcs, rho, vel)
        assert abs(err - ref_err) <= 1e-13 * max(1.0, abs(ref_err)), (err, ref_err)
        for ax in range(3):
            assert np.abs(_host(dvel[ax], cfg.shape) - ref[ax]).max() <= 1e-13, ax
        err2 = dp.correct_mfr(rho, dvel)
        assert abs(err2) <= 1e-12 * abs(ref_err)
    dp.set_bconds([])
    assert dp.correct_mfr(1.0, [_dev(v) for v in vel]) == 0.0
    dp.destroy()
