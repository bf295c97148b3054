"""GPU tests of method gmres_pfmg (flexible GMRES(kdim) right-preconditioned
by the V-cycle of the operator's symmetric part) through the C ABI, on the
nonsymmetric implicit-solve test systems of inputs.implicit_problem.

The system is nonsingular, so its solution is unique: every solve is held to
a true residual <= 1e-10 and to 1e-8 relative of scipy's sparse direct
solution (small grids). Iteration counts are this build's own (HYPRE's are
not available here).
"""
import numpy as np
import pytest
import torch

from conftest import STC7, assemble_csr
from nga2_amd import inputs
from nga2_amd._lib import ERR_ARG, FP64, GMRES_PFMG, MIXED, PCG_PFMG, NgapError
from nga2_amd.config import Config, tanh_stretched, uniform
from nga2_amd.linsol import DeviceSolve, NgapStr
from oracle.orc import interior

pytestmark = pytest.mark.gpu
DEV = "cuda:0"

CASES = {
    "periodic16": lambda: Config(uniform(16), uniform(16), uniform(16), no=1),
    "channel16": lambda: Config(uniform(16), tanh_stretched(16), uniform(16), no=1, yper=False),
    "walls13x10x7": lambda: Config(uniform(13), uniform(10), uniform(7), no=1, xper=False, yper=False, zper=False),
    "plane16x16x1": lambda: Config(uniform(16), uniform(16), uniform(1), no=1, zper=False),
}


def _direct(cfg, P):
    import scipy.sparse.linalg as spl

    A = assemble_csr(cfg, P.opr)
    return spl.spsolve(A.tocsc(), interior(P.rhs, cfg))


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("prec", [FP64, MIXED])
def test_matches_direct_solve(name, prec):
    print(f"\n[gmres {name} prec={prec}]", flush=True)
    cfg = CASES[name]()
    P = inputs.implicit_problem(cfg, visc=0.05)
    xr = _direct(cfg, P)
    ps = NgapStr(cfg, "Velocity", GMRES_PFMG, 7)
    ps.stc[:] = STC7
    ps.maxit, ps.rcvg, ps.precision = 200, 1e-10, prec
    ps.opr[:] = P.opr
    ps.init()
    ps.setup()
    ps.rhs[:] = P.rhs
    ps.sol[:] = 0.0
    ps.solve()
    st = ps.stats()
    x = interior(ps.sol, cfg)
    err = np.linalg.norm(x - xr) / np.linalg.norm(xr)
    print(f"[gmres {name}] it={ps.it} restarts={st['restarts']} true={st['true_rerr']:.2e} err={err:.2e}", flush=True)
    assert st["true_rerr"] <= 1e-10 and st["status"] == 0
    assert err <= 1e-8
    assert ps.it >= 1
    ps.destroy()


@pytest.mark.parametrize("kdim", [1, 3, 12])
def test_restart_lengths(kdim):
    """kdim = 1 restarts every step; 12 splits the basis kernels into chunks."""
    print(f"\n[gmres kdim={kdim}]", flush=True)
    cfg = CASES["channel16"]()
    P = inputs.implicit_problem(cfg, visc=0.05)
    xr = _direct(cfg, P)
    ds = DeviceSolve(cfg, method=GMRES_PFMG)
    d_opr = torch.from_numpy(np.ascontiguousarray(P.opr.ravel(order="F"))).to(DEV)
    d_rhs = torch.from_numpy(np.ascontiguousarray(P.rhs.ravel(order="F"))).to(DEV)
    ds.setup(d_opr, maxit=500, rcvg=1e-10, kdim=kdim)
    d_x = torch.zeros_like(d_rhs)
    it, _ = ds.solve(d_rhs, d_x)
    st = ds.stats()
    x = interior(d_x.cpu().numpy().reshape(cfg.shape, order="F"), cfg)
    err = np.linalg.norm(x - xr) / np.linalg.norm(xr)
    print(f"[gmres kdim={kdim}] it={it} restarts={st['restarts']} err={err:.2e}", flush=True)
    assert st["true_rerr"] <= 1e-10 and err <= 1e-8
    if kdim == 1:
        assert st["restarts"] >= 1
    ds.destroy()


def test_symmetric_system_matches_pcg():
    """With zero velocities the operator is SPD: GMRES and PCG reach the same solution."""
    print("\n[gmres vs pcg]", flush=True)
    cfg = CASES["walls13x10x7"]()
    P = inputs.implicit_problem(cfg, visc=0.1)
    from nga2_amd import incomp

    P.opr[:] = incomp.convection_diffusion(incomp.Metrics(cfg), [np.zeros(cfg.shape)] * 3, 1.0, 0.1, 0.2)
    sols = []
    for method in (PCG_PFMG, GMRES_PFMG):
        ps = NgapStr(cfg, "S", method, 7)
        ps.stc[:] = STC7
        ps.maxit, ps.rcvg = 200, 1e-12
        ps.opr[:] = P.opr
        ps.init()
        ps.setup()
        ps.rhs[:] = P.rhs
        ps.sol[:] = 0.0
        ps.solve()
        sols.append(interior(ps.sol, cfg))
        assert ps.stats()["true_rerr"] <= 1e-12
        ps.destroy()
    assert np.linalg.norm(sols[0] - sols[1]) <= 1e-10 * np.linalg.norm(sols[1])


def test_bad_restart_length():
    cfg = CASES["periodic16"]()
    P = inputs.implicit_problem(cfg)
    ds = DeviceSolve(cfg, method=GMRES_PFMG)
    d_opr = torch.from_numpy(np.ascontiguousarray(P.opr.ravel(order="F"))).to(DEV)
    for k in (0, 31):
        with pytest.raises(NgapError) as e:
            ds.setup(d_opr, kdim=k)
        assert e.value.code == ERR_ARG
    ds.destroy()


def test_full_size_channel():
    """256^3 stretched channel (walls in y): FP64 and mixed GMRES converge to
    1e-10 and agree."""
    print("\n[gmres 256^3 channel]", flush=True)
    import time

    cfg = Config(uniform(256), tanh_stretched(256), uniform(256), no=1, yper=False)
    P = inputs.implicit_problem(cfg, visc=0.02)
    d_opr = torch.from_numpy(np.ascontiguousarray(P.opr.ravel(order="F"))).to(DEV)
    d_rhs = torch.from_numpy(np.ascontiguousarray(P.rhs.ravel(order="F"))).to(DEV)
    xs = []
    for prec in (FP64, MIXED):
        ds = DeviceSolve(cfg, method=GMRES_PFMG)
        ds.setup(d_opr, maxit=200, rcvg=1e-10, precision=prec)
        d_x = torch.zeros_like(d_rhs)
        ds.solve(d_rhs, d_x)
        d_x.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        it, _ = ds.solve(d_rhs, d_x)
        ms = (time.perf_counter() - t) * 1e3
        st = ds.stats()
        print(f"[gmres 256^3 prec={prec}] it={it} restarts={st['restarts']} true={st['true_rerr']:.2e} {ms:.1f} ms",
              flush=True)
        assert st["true_rerr"] <= 1e-10
        xs.append(d_x.cpu().numpy())
        ds.destroy()
    assert np.linalg.norm(xs[0] - xs[1]) <= 1e-8 * np.linalg.norm(xs[0])


@pytest.mark.parametrize("name", ["channel16", "walls13x10x7"])
def test_setup_again_matches_fresh_setup(name):
    """gmres_pfmg on one device handle set up again for a sequence of
    implicit operators (viscosities and velocity fields change, so the
    symmetric part's strengths move): each setup's hierarchy and solve are
    those of a fresh handle, bit for bit, and meet the direct solution."""
    print(f"\n[gmres setup again {name}]", flush=True)
    cfg = CASES[name]()
    to_dev = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a).ravel(order="F"))).to(DEV)
    seq = [inputs.implicit_problem(cfg, seed=s, visc=v) for s, v in ((1, 0.05), (2, 0.002), (3, 0.5), (4, 0.05))]
    ds = DeviceSolve(cfg, method=GMRES_PFMG)
    for q, P in enumerate(seq):
        d_opr, d_rhs = to_dev(P.opr), to_dev(P.rhs)
        ds.setup(d_opr, maxit=200, rcvg=1e-10)
        fresh = DeviceSolve(cfg, method=GMRES_PFMG)
        fresh.setup(d_opr, maxit=200, rcvg=1e-10)
        ha, hb = ds.hooks(), fresh.hooks()
        assert ha.nlevels == hb.nlevels
        assert [ha.level_info(l) for l in range(ha.nlevels)] == [hb.level_info(l) for l in range(hb.nlevels)]
        for l in range(ha.nlevels):
            ca, cb = ha.level_coefs(l), hb.level_coefs(l)
            for k in ca:
                assert np.array_equal(ca[k], cb[k]), (q, l, k)
        xa, xb = torch.zeros_like(d_rhs), torch.zeros_like(d_rhs)
        ita, _ = ds.solve(d_rhs, xa)
        itb, _ = fresh.solve(d_rhs, xb)
        assert ita == itb and torch.equal(xa, xb), q
        st = ds.stats()
        x = interior(xa.cpu().numpy().reshape(P.rhs.shape, order="F"), cfg)
        xr = _direct(cfg, P)
        print(f"  setup {q}: it {ita} true {st['true_rerr']:.2e}", flush=True)
        assert st["true_rerr"] <= 1e-10 and np.linalg.norm(x - xr) <= 1e-8 * np.linalg.norm(xr)
        fresh.destroy()
    ds.destroy()
