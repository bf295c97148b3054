"""GPU parity of hypre_uns (libngapuns.so) with the numpy oracle
(oracle/amg_ref.py) through the C ABI.

The hierarchy is built by the library's host setup, which the CPU tests hold
equal to the oracle's entry by entry; here the device cycle and the Krylov
drivers are compared with the oracle's on the same hierarchy: one V-cycle to
1e-12 of max|x|, every method's solution to 1e-8 (the operators are singular,
so solutions are compared with their fluid mean removed) with the oracle's
iteration count +-1. At 64^3 the masked solve is checked by its true residual
and against the structured PCG+PFMG solver (libngapsolve.so) on the same
system, an independent algorithm.
"""
import numpy as np
import pytest
import torch

from conftest import STC7
from nga2_amd import inputs, uns
from nga2_amd._lib import ERR_METHOD, ERR_NOT_SETUP, NgapError
from oracle import amg_ref as R

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
MC = 40  # max_coarse: several levels even on these small grids

CASES = {
    "masked12": lambda: inputs.masked_spheres(n=12, nsph=4, rmin=0.15, rmax=0.25),
    "channel8": lambda: inputs.channel_spectral(n=8),
    "periodic8": lambda: inputs.periodic_gaussian(n=8),
}


def _oracle(P):
    ind, n = R.umap(P.cfg, P.opr)
    rp, c, v = R.assemble(P.cfg, ind, P.opr, STC7)
    A = R.CSR(rp, c, v, n)
    return ind, A, R.build(A, max_coarse=MC)


def _solver(P, method, maxit=300, rcvg=1e-10, **opts):
    s = uns.NgapUns(P.cfg, "Pressure", method, 7)
    s.stc[:] = STC7
    s.opr[:] = P.opr
    s.init()
    s.maxit, s.rcvg = maxit, rcvg
    s.opts = dict(max_coarse=MC, **opts)
    return s


@pytest.mark.parametrize("name", list(CASES))
def test_vcycle_matches_oracle(name):
    print(f"\n[uns vcycle {name}]", flush=True)
    P = CASES[name]()
    ind, A, H = _oracle(P)
    ds = uns.DeviceUns(P.cfg, P.opr, method=uns.PCG_AMG, max_coarse=MC)
    assert ds.rows == A.n
    st = ds.stats()
    assert st["nlevels"] == len(H.A)
    rng = np.random.default_rng(11)
    b = rng.standard_normal(A.n)
    ref = R.vcycle(H, b)
    d_b = torch.from_numpy(b).to(DEV)
    d_x = torch.zeros_like(d_b)
    ds.vcycle(d_b, d_x)
    got = d_x.cpu().numpy()
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), np.abs(got - ref).max() / np.abs(ref).max()
    ds.destroy()


@pytest.mark.parametrize("method", [uns.AMG, uns.PCG_AMG, uns.GMRES_AMG, uns.BICGSTAB_AMG, uns.PCG, uns.GMRES])
@pytest.mark.parametrize("name", list(CASES))
def test_methods_match_oracle(name, method):
    print(f"\n[uns {uns.METHOD_NAMES[method]} {name}]", flush=True)
    P = CASES[name]()
    cfg = P.cfg
    ind, A, H = _oracle(P)
    it_ = cfg.interior
    act = ind[it_] >= 0
    rng = np.random.default_rng(5)
    sol0 = np.asfortranarray(rng.standard_normal(cfg.shape))
    cfg.sync(sol0)
    b = P.rhs[it_].ravel(order="F")[act.ravel(order="F")]
    x0 = sol0[it_].ravel(order="F")[act.ravel(order="F")]
    xr, itr, rr = R.solve(method, A, H, b, x0, 300, 1e-10)
    s = _solver(P, method)
    s.setup()
    s.rhs[:] = P.rhs
    s.sol[:] = sol0
    s.solve()
    assert s.rerr < 1e-10 and abs(s.it - itr) <= 1, (s.it, itr, s.rerr)
    got = s.sol[it_].ravel(order="F")[act.ravel(order="F")]
    d, e = got - got.mean(), xr - xr.mean()
    assert np.linalg.norm(d - e) <= 1e-8 * np.linalg.norm(e), np.linalg.norm(d - e) / np.linalg.norm(e)
    # cells outside the system keep the caller's values; ghosts are synced
    inter = s.sol[it_]
    assert np.array_equal(inter[~act], sol0[it_][~act])
    chk = s.sol.copy()
    cfg.sync(chk)
    assert np.array_equal(chk, s.sol)
    st = s.stats()
    assert st["it"] == s.it and st["status"] == 0
    s.destroy()


@pytest.mark.parametrize("method", [uns.AMG, uns.GMRES_AMG, uns.BICGSTAB_AMG, uns.GMRES])
def test_nonsymmetric_methods_match_oracle(method):
    """The implicit convection-diffusion operator on the walled channel grid
    (nonsymmetric, nonsingular): the solution equals the oracle's to 1e-8 with
    its iteration count +-1."""
    print(f"\n[uns nonsymmetric {uns.METHOD_NAMES[method]}]", flush=True)
    from nga2_amd.config import Config, tanh_stretched, uniform

    cfg = Config(uniform(8), tanh_stretched(8), uniform(8), no=1, yper=False, name="implicit")
    P = inputs.implicit_problem(cfg, visc=0.02)
    ind, A, H = _oracle(P)
    b = P.rhs[cfg.interior].ravel(order="F")
    xr, itr, _ = R.solve(method, A, H, b, np.zeros(A.n), 300, 1e-10)
    s = _solver(P, method)
    s.setup()
    s.rhs[:] = P.rhs
    s.sol[:] = 0.0
    s.solve()
    got = s.sol[cfg.interior].ravel(order="F")
    assert s.rerr < 1e-10 and abs(s.it - itr) <= 1, (s.it, itr, s.rerr)
    assert np.linalg.norm(got - xr) <= 1e-8 * np.linalg.norm(xr)
    s.destroy()


def test_mixed_vcycle_matches_oracle():
    """The FP32 cycle (precision NGAP_MIXED) against the oracle's FP64 cycle on
    the same hierarchy, to FP32 rounding."""
    print("\n[uns mixed vcycle]", flush=True)
    P = CASES["masked12"]()
    ind, A, H = _oracle(P)
    ds = uns.DeviceUns(P.cfg, P.opr, method=uns.PCG_AMG, max_coarse=MC, precision=1)
    b = np.random.default_rng(12).standard_normal(A.n)
    ref = R.vcycle(H, b)
    d_x = torch.zeros(A.n, dtype=torch.float64, device=DEV)
    ds.vcycle(torch.from_numpy(b).to(DEV), d_x)
    got = d_x.cpu().numpy()
    assert np.abs(got - ref).max() <= 2e-5 * np.abs(ref).max(), np.abs(got - ref).max() / np.abs(ref).max()
    ds.destroy()


@pytest.mark.parametrize("method", [uns.AMG, uns.PCG_AMG, uns.GMRES_AMG, uns.BICGSTAB_AMG])
@pytest.mark.parametrize("name", ["masked12", "channel8"])
def test_mixed_precision_methods(name, method):
    """FP32 cycle under the FP64 iteration: the same tolerance on the true
    residual, the FP64 oracle's solution to 1e-8, at most three more
    iterations than the oracle's FP64 count."""
    print(f"\n[uns mixed {uns.METHOD_NAMES[method]} {name}]", flush=True)
    P = CASES[name]()
    cfg = P.cfg
    ind, A, H = _oracle(P)
    it_ = cfg.interior
    act = ind[it_] >= 0
    b = P.rhs[it_].ravel(order="F")[act.ravel(order="F")]
    xr, itr, _ = R.solve(method, A, H, b, np.zeros(A.n), 300, 1e-10)
    s = _solver(P, method, precision=1)
    s.setup()
    s.rhs[:] = P.rhs
    s.sol[:] = 0.0
    s.solve()
    assert s.rerr < 1e-10 and s.it <= itr + 3, (s.it, itr, s.rerr)
    got = s.sol[it_].ravel(order="F")[act.ravel(order="F")]
    d, e = got - got.mean(), xr - xr.mean()
    assert np.linalg.norm(d - e) <= 1e-8 * np.linalg.norm(e)
    s.destroy()


@pytest.mark.parametrize("precision", [0, 1])
def test_single_level_hierarchy(precision):
    """216 rows (<= max_coarse 512): the hierarchy is the fine level alone and
    the cycle is the dense (pseudo-)inverse, so PCG converges at once."""
    print(f"\n[uns single level, precision {precision}]", flush=True)
    P = inputs.periodic_gaussian(n=6)
    s = uns.NgapUns(P.cfg, "Pressure", uns.PCG_AMG, 7)
    s.stc[:] = STC7
    s.opr[:] = P.opr
    s.init()
    s.maxit, s.rcvg, s.opts = 50, 1e-10, dict(precision=precision)
    s.setup()
    assert s.stats()["nlevels"] == 1 and s.ncell == 216
    s.rhs[:] = P.rhs
    s.sol[:] = 0.0
    s.solve()
    assert s.rerr < 1e-10 and s.it <= (2 if precision == 0 else 4), (s.it, s.rerr)
    s.destroy()


def test_errors_and_zero_rhs():
    print("\n[uns errors]", flush=True)
    P = CASES["masked12"]()
    s = _solver(P, uns.PCG_AMG)
    with pytest.raises(NgapError) as e:
        s.solve()
    assert e.value.code == ERR_NOT_SETUP
    s.setup()
    s.rhs[:] = 0.0
    s.sol[:] = 1.0
    s.solve()
    assert s.it == 0 and s.rerr == 0.0
    assert np.all(s.sol[P.cfg.interior][s.ind[P.cfg.interior] >= 0] == 0.0)
    s.destroy()
    for m in (uns.PCG_PARASAIL, uns.GMRES_PILUT):
        t = _solver(P, m)
        with pytest.raises(NgapError) as e:
            t.setup()
        assert e.value.code == ERR_METHOD
        t.destroy()
    t = uns.NgapUns(P.cfg, "Pressure", 9, 7)
    t.stc[:] = STC7
    with pytest.raises(NgapError) as e:
        t.init()
    assert e.value.code == ERR_METHOD


def test_masked64_against_structured():
    """64^3 with 40 solid spheres, default hierarchy (coarsest <= 512 rows)."""
    print("\n[uns masked64]", flush=True)
    from nga2_amd.linsol import NgapStr, PCG_PFMG

    P = inputs.masked_spheres(n=64)
    cfg = P.cfg
    assert P.info["regions"] == 1 and P.info["solid_cells"] > 0
    s = uns.NgapUns(cfg, "Pressure", uns.PCG_AMG, 7)
    s.stc[:] = STC7
    s.opr[:] = P.opr
    s.init()
    s.maxit, s.rcvg = 200, 1e-10
    s.setup()
    s.rhs[:] = P.rhs
    s.sol[:] = 0.0
    s.solve()
    st = s.stats()
    print(f"[uns masked64] rows {st['rows']}, levels {st['nlevels']}, complexity {st['op_complexity']:.2f}, "
          f"it {s.it}, rerr {s.rerr:.2e}, setup {st['setup_ms']:.1f} ms, solve {st['solve_ms']:.2f} ms", flush=True)
    assert s.rerr < 1e-10
    it_ = cfg.interior
    act = s.ind[it_] >= 0
    # the structured solver on the same operator (masked cells: identity rows)
    p = NgapStr(cfg, "Pressure", PCG_PFMG, 7)
    p.stc[:] = STC7
    p.init()
    p.opr[:] = P.opr
    p.maxit, p.rcvg = 200, 1e-11
    p.setup()
    p.rhs[:] = P.rhs
    p.sol[:] = 0.0
    p.solve()
    a = s.sol[it_][act]
    b = p.sol[it_][act]
    a, b = a - a.mean(), b - b.mean()
    err = np.linalg.norm(a - b) / np.linalg.norm(b)
    print(f"[uns masked64] structured PCG+PFMG it {p.it}, rel diff {err:.2e}", flush=True)
    assert err <= 1e-8
    # the FP32 cycle reaches the same solution
    m = uns.NgapUns(cfg, "Pressure", uns.PCG_AMG, 7)
    m.stc[:] = STC7
    m.opr[:] = P.opr
    m.init()
    m.maxit, m.rcvg, m.opts = 200, 1e-10, dict(precision=1)
    m.setup()
    m.rhs[:] = P.rhs
    m.sol[:] = 0.0
    m.solve()
    c = m.sol[it_][act]
    c = c - c.mean()
    print(f"[uns masked64] mixed it {m.it} rerr {m.rerr:.2e}", flush=True)
    assert m.rerr < 1e-10 and np.linalg.norm(c - b) <= 1e-8 * np.linalg.norm(b)
    m.destroy()
    p.destroy()
    s.destroy()
