"""GPU parity: every kernel of libngapsolve.so against the CPU oracle, through
the C ABI, on small and odd-shaped problems.

Tolerances: FP64 kernels are element-wise identical algorithms issued in the
same order, so they agree to a few ulp (<= 1e-12 relative to the field's max);
FP32 kernels to 1e-5 relative; solves to 1e-8 relative L2 after mean removal
(the bar of the FP64 oracle comparison) with a true residual <= 1e-10.
"""
import numpy as np
import pytest

from conftest import STC7, rel_err_mean_free, rel_err_nullspace
from nga2_amd import incomp, inputs
from nga2_amd._lib import MIXED, PCG_PFMG
from nga2_amd.config import Config, uniform
from nga2_amd.linsol import NgapStr
from oracle.orc import Oracle, interior

pytestmark = pytest.mark.gpu


def _box(nx, ny, nz, per, walls=True, L=(1.0, 1.0, 1.0)):
    cfg = Config(uniform(nx, L[0]), uniform(ny, L[1]), uniform(nz, L[2]), no=1, xper=per[0], yper=per[1],
                 zper=per[2])
    if walls:
        cfg.VF[:] = 0.0
        cfg.VF[cfg.interior] = 1.0
        cfg.sync(cfg.VF)
        cfg.calc_fluid_vol()
    met = incomp.Metrics(cfg)
    opr = incomp.laplacian(met)
    rng = np.random.default_rng(nx * 100 + ny * 10 + nz)
    rhs = cfg.field()
    rhs[cfg.interior] = rng.standard_normal((nx, ny, nz))
    rhs[cfg.interior] -= rhs[cfg.interior].mean()
    return inputs.Problem(cfg, opr, rhs, f"box{nx}x{ny}x{nz}")


def _masked_box(nx, ny, nz, per, block):
    """A box with a solid block (VF = 0): its cells get opr(1) = 0 and become
    identity rows with rhs = sol = 0 (hypre_str_class.f90:174-175,339-345);
    the fluid cells around it see homogeneous Neumann faces."""
    cfg = Config(uniform(nx), uniform(ny), uniform(nz), no=1, xper=per[0], yper=per[1], zper=per[2])
    cfg.VF[:] = 0.0
    cfg.VF[cfg.interior] = 1.0
    (i0, i1), (j0, j1), (k0, k1) = block
    cfg.VF[1 + i0:1 + i1, 1 + j0:1 + j1, 1 + k0:1 + k1] = 0.0
    cfg.sync(cfg.VF)
    cfg.calc_fluid_vol()
    opr = incomp.laplacian(incomp.Metrics(cfg))
    act = opr[0][cfg.interior] != 0.0
    rng = np.random.default_rng(nx * 7 + ny)
    vals = np.where(act, rng.standard_normal((nx, ny, nz)), 0.0)
    vals[act] -= vals[act].mean()
    rhs = cfg.field()
    rhs[cfg.interior] = vals
    return inputs.Problem(cfg, opr, rhs, f"masked{nx}x{ny}x{nz}")


CASES = {
    "periodic16": lambda: inputs.periodic_gaussian(n=16),
    "channel16": lambda: inputs.channel_spectral(n=16),
    "spheres24": lambda: inputs.spheres_two_phase(n=24, nsph=6, rmin=0.1, rmax=0.25),
    "walls13x10x7": lambda: _box(13, 10, 7, (False, False, False)),
    "mixed12x9x8": lambda: _box(12, 9, 8, (True, False, True)),
    "thin16x16x1": lambda: _box(16, 16, 1, (True, True, True), walls=False),
    "rect40x6x18": lambda: _box(40, 6, 18, (True, False, False)),
    # levels of 2^20 cells with rows of 128/256 cells: the row-wave kernels
    # (one x-row per wavefront), with walls in y, and in x (row ends)
    "rows256x16x256": lambda: _box(256, 16, 256, (True, False, True)),
    "rows128x64x128": lambda: _box(128, 64, 128, (False, True, True)),
    # rows that end inside a wavefront (260 and 200 cells: the last wave of
    # a row holds 1 or 50 active lanes in FP32), walled and periodic in x;
    # cells of equal size in all directions
    "rows260x12x20": lambda: _box(260, 12, 20, (False, False, False), L=(1.0, 12 / 260, 20 / 260)),
    "rows200x10x16": lambda: _box(200, 10, 16, (True, False, True), L=(1.0, 10 / 200, 16 / 200)),
    # a line of cells along z (x and y one and three cells wide)
    "line1x3x300": lambda: _box(1, 3, 300, (True, False, False)),
    # identity rows (opr(1) == 0): a solid block inside a box with walls in
    # y, and solid spheres immersed in a periodic cube
    "masked_block16x11x8": lambda: _masked_box(16, 11, 8, (True, False, True), ((5, 9), (3, 7), (2, 5))),
    "masked_spheres16": lambda: inputs.masked_spheres(n=16, nsph=3, rmin=0.15, rmax=0.3),
}


def _solver(P, precision=0, maxit=200, rcvg=1e-10, **opts):
    ps = NgapStr(P.cfg, "Pressure", PCG_PFMG, 7)
    ps.stc[:] = STC7
    ps.maxit, ps.rcvg, ps.precision = maxit, rcvg, precision
    ps.opts.update(opts)
    ps.opr[:] = P.opr
    ps.init()
    ps.setup()
    return ps


def _close(a, b, tol):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() <= tol * scale


@pytest.fixture(scope="module", params=list(CASES))
def case(request):
    print(f"\n[case] {request.param}", flush=True)
    P = CASES[request.param]()
    o = Oracle.from_cfg(P.cfg, P.opr)
    # V(1,1) in both precisions, as the oracle's V-cycle
    ps = _solver(P, precision=MIXED, nu_pre=1, nu_post=1)
    yield request.param, P, o, ps
    ps.destroy()


def test_hierarchy_matches_oracle(case):
    name, P, o, ps = case
    hk = ps.hooks()
    assert hk.nlevels == o.nlevels
    for l in range(o.nlevels):
        assert hk.level_info(l) == o.level_info(l), l
        cg, co = hk.level_coefs(l), o.level_coefs(l)
        for k in ("c", "ax", "ay", "az", "pm", "pp"):
            if l == 0 and k in ("pm", "pp"):
                continue
            assert _close(cg[k], co[k], 1e-13), (l, k)


def test_level_kernels_fp64(case):
    name, P, o, ps = case
    hk = ps.hooks()
    rng = np.random.default_rng(0)
    for l in range(o.nlevels):
        n = o.level_size(l)
        x = rng.standard_normal(n)
        f = rng.standard_normal(n)
        assert _close(hk.apply(l, x), o.apply(l, x), 1e-13), ("apply", l)
        for order in (0, 1):
            assert _close(hk.rbgs(l, f, x, order), o.rbgs(l, f, x, order), 1e-12), ("rbgs", l, order)
            # first sweep from a zero iterate (stencil form of the pre-smoother)
            zero = np.zeros(n)
            assert _close(hk.rbgs(l, f, x, 2 | order), o.rbgs(l, f, zero, order), 1e-12), ("rbgs0", l, order)
        if l + 1 < o.nlevels:
            assert _close(hk.resid_restrict(l, f, x), o.resid_restrict(l, f, x), 1e-12), ("restrict", l)
            vc = rng.standard_normal(o.level_size(l + 1))
            assert _close(hk.prolong_add(l, vc, x), o.prolong_add(l, vc, x), 1e-13), ("prolong", l)
    fc = rng.standard_normal(o.level_size(o.nlevels - 1))
    if o.coarse_singular:
        fc -= fc.mean()
    assert _close(hk.coarse_solve(fc), o.coarse_solve(fc), 1e-9)
    r = rng.standard_normal(o.N)
    r -= r.mean()
    assert _close(hk.vcycle(r), o.vcycle(r), 1e-9)
    x = rng.standard_normal(o.N)
    assert _close(hk.apply_exact(x), o.apply_exact(x), 1e-13)
    assert abs(hk.dot(x, r) - x @ r) <= 1e-12 * np.linalg.norm(x) * np.linalg.norm(r)


@pytest.mark.parametrize("prec", [0, 1])
def test_fused_sweep_pair_is_two_sweeps(case, prec):
    """The post-smoothing pair (two sweeps in one launch, 3-cell recompute
    ring, four half-steps pipelined over planes) equals two single sweeps bit
    for bit, on every level and in both colour orders."""
    name, P, o, ps = case
    hk = ps.hooks()
    rng = np.random.default_rng(7)
    for l in range(o.nlevels):
        n = o.level_size(l)
        x = rng.standard_normal(n)
        f = rng.standard_normal(n)
        for order in (0, 1):
            two = hk.rbgs(l, f, hk.rbgs(l, f, x, order, prec=prec), order, prec=prec)
            fused = hk.rbgs(l, f, x, 4 | order, prec=prec)
            assert np.array_equal(fused, two), (name, l, order, np.abs(fused - two).max())


@pytest.mark.parametrize("prec", [0, 1])
def test_fused_pair_from_zero(case, prec):
    """The pre-smoothing pair (two sweeps from a zero iterate in one launch,
    the V-cycle's choice on levels of >= 2^24 cells) equals the zero-iterate
    sweep followed by a sweep of the same order, bit for bit, on every level."""
    name, P, o, ps = case
    hk = ps.hooks()
    rng = np.random.default_rng(10)
    for l in range(o.nlevels):
        n = o.level_size(l)
        x = rng.standard_normal(n)
        f = rng.standard_normal(n)
        for order in (0, 1):
            two = hk.rbgs(l, f, hk.rbgs(l, f, x, 2 | order, prec=prec), order, prec=prec)
            fused = hk.rbgs(l, f, x, 6 | order, prec=prec)
            assert np.array_equal(fused, two), (name, l, order, np.abs(fused - two).max())


@pytest.mark.parametrize("prec", [0, 1])
def test_row_sweep_matches_tiled(case, prec):
    """On levels whose rows fit one workgroup row (<= 128 cells) the sweep runs
    whole-row tiles with a ring in y only; it equals the 64-wide tiled sweep
    (ring in x and y) bit for bit, from an iterate and from zero, on every
    level and in both colour orders."""
    name, P, o, ps = case
    hk = ps.hooks()
    rng = np.random.default_rng(12)
    for l in range(o.nlevels):
        n = o.level_size(l)
        x = rng.standard_normal(n)
        f = rng.standard_normal(n)
        for order in (0, 1):
            for zero in (0, 16):
                row = hk.rbgs(l, f, x, zero | order, prec=prec)
                tiled = hk.rbgs(l, f, x, 8 | zero | order, prec=prec)
                assert np.array_equal(row, tiled), (name, l, order, zero, np.abs(row - tiled).max())


@pytest.mark.parametrize("prec", [0, 1])
def test_zero_sweep_matches_stencil(case, prec):
    """The first pre-smoothing sweep from a zero iterate runs as the sweep
    kernel on levels of >= 2^20 cells with rows of <= 128 cells and as the
    stencil form elsewhere; the two give the same bits on every level."""
    name, P, o, ps = case
    hk = ps.hooks()
    rng = np.random.default_rng(13)
    for l in range(o.nlevels):
        n = o.level_size(l)
        x = rng.standard_normal(n)
        f = rng.standard_normal(n)
        for order in (0, 1):
            sweep = hk.rbgs(l, f, x, 16 | order, prec=prec)
            stencil = hk.rbgs(l, f, x, 2 | order, prec=prec)
            assert np.array_equal(sweep, stencil), (name, l, order, np.abs(sweep - stencil).max())


@pytest.mark.parametrize("prec", [0, 1])
def test_vector_transfers_match_scalar(case, prec):
    """Restriction and prolongation with CPL cells per thread and 16-byte
    accesses (y/z coarsening, rows a multiple of CPL) equal the
    one-cell-per-thread kernels bit for bit, on every level."""
    name, P, o, ps = case
    hk = ps.hooks()
    rng = np.random.default_rng(11)
    for l in range(o.nlevels - 1):
        n, nc = o.level_size(l), o.level_size(l + 1)
        x, f, vc = rng.standard_normal(n), rng.standard_normal(n), rng.standard_normal(nc)
        a, b = hk.resid_restrict(l, f, x, prec=prec), hk.resid_restrict(l, f, x, prec=prec | 2)
        assert np.array_equal(a, b), (name, "restrict", l, np.abs(a - b).max())
        a, b = hk.prolong_add(l, vc, x, prec=prec), hk.prolong_add(l, vc, x, prec=prec | 2)
        assert np.array_equal(a, b), (name, "prolong", l, np.abs(a - b).max())
    if prec:
        # FP32 prolongation against the oracle (the FP64 one is checked above)
        for l in range(o.nlevels - 1):
            n, nc = o.level_size(l), o.level_size(l + 1)
            x, vc = rng.standard_normal(n), rng.standard_normal(nc)
            assert _close(hk.prolong_add(l, vc, x, prec=1), o.prolong_add(l, vc, x), 2e-6), ("prolong", l)


def test_level_kernels_fp32(case):
    name, P, o, ps = case
    hk = ps.hooks()
    rng = np.random.default_rng(1)
    for l in range(o.nlevels):
        n = o.level_size(l)
        x = rng.standard_normal(n)
        f = rng.standard_normal(n)
        assert _close(hk.apply(l, x, prec=1), o.apply(l, x), 2e-5), ("apply", l)
        for order in (0, 1):
            assert _close(hk.rbgs(l, f, x, order, prec=1), o.rbgs(l, f, x, order), 1e-4), ("rbgs", l, order)
            zero = np.zeros(n)
            assert _close(hk.rbgs(l, f, x, 2 | order, prec=1), o.rbgs(l, f, zero, order), 1e-4), ("rbgs0", l, order)
        if l + 1 < o.nlevels:
            assert _close(hk.resid_restrict(l, f, x, prec=1), o.resid_restrict(l, f, x), 2e-5), ("restrict", l)
    r = rng.standard_normal(o.N)
    r -= r.mean()
    assert _close(hk.vcycle(r, prec=1), o.vcycle(r), 1e-3)


@pytest.mark.parametrize("precision", [0, MIXED])
def test_solve_matches_oracle(case, precision):
    name, P, o, _ = case
    b = interior(P.rhs, P.cfg)
    xr, st_r = o.solve(b, rcvg=1e-13, maxit=500)
    _, st_10 = o.solve(b, rcvg=1e-10, maxit=500)
    ps = _solver(P, precision=precision)
    ps.rhs[:] = P.rhs
    ps.sol[:] = 0.0
    ps.solve()
    st = ps.stats()
    x = interior(ps.sol, P.cfg)
    print(f"[{name} prec={precision}] it={ps.it} oracle_it={st_10['it']} true={st['true_rerr']:.2e}", flush=True)
    assert st["status"] == 0
    assert st["true_rerr"] <= 1e-10
    # one free constant per connected fluid region (identity rows are 0 in both)
    assert rel_err_nullspace(x, xr, P.cfg, P.opr) <= 1e-8
    if precision == 0:
        assert abs(ps.it - st_10["it"]) <= 1
    # ghosts synced like cfg%sync (hypre_str_class.f90:380)
    chk = ps.sol.copy()
    P.cfg.sync(chk)
    assert np.array_equal(chk, ps.sol)
    ps.destroy()


@pytest.mark.parametrize("name", list(CASES))
def test_skip_relax_matches_oracle(name):
    """Skipped relaxation (opts.skip_relax = 1): the GPU V-cycle in both
    precisions against the oracle's with the same skip rule, and the solves
    against the oracle's solution."""
    print(f"\n[skip_relax {name}]", flush=True)
    P = CASES[name]()
    o = Oracle.from_cfg(P.cfg, P.opr, skip_relax=1)
    ps = _solver(P, precision=MIXED, nu_pre=1, nu_post=1, skip_relax=1)
    hk = ps.hooks()
    rng = np.random.default_rng(3)
    r = rng.standard_normal(o.N)
    r -= r.mean()
    zo = o.vcycle(r)
    assert _close(hk.vcycle(r), zo, 1e-9)
    assert _close(hk.vcycle(r, prec=1), zo, 1e-3)
    # the skip changes the cycle wherever it skips a level
    o_all = Oracle.from_cfg(P.cfg, P.opr, skip_relax=0)
    if 0 in o.relax_levels():
        assert not _close(o_all.vcycle(r), zo, 1e-6)
    o_all.close()
    ps.destroy()
    b = interior(P.rhs, P.cfg)
    xr, _ = o.solve(b, rcvg=1e-13, maxit=500)
    _, st_10 = o.solve(b, rcvg=1e-10, maxit=500)
    for precision in (0, MIXED):
        ps = _solver(P, precision=precision, skip_relax=1)
        ps.rhs[:] = P.rhs
        ps.sol[:] = 0.0
        ps.solve()
        st = ps.stats()
        x = interior(ps.sol, P.cfg)
        print(f"  prec={precision} it={ps.it} oracle_it={st_10['it']} relax={o.relax_levels()}", flush=True)
        assert st["status"] == 0 and st["true_rerr"] <= 1e-10
        assert rel_err_nullspace(x, xr, P.cfg, P.opr) <= 1e-8
        if precision == 0:
            assert abs(ps.it - st_10["it"]) <= 1
        ps.destroy()
    o.close()


def _scaled(P, d, f):
    """P's operator with its couplings along direction d scaled by f and the
    centre moved so that every row sum stays (identity rows stay identity)."""
    opr = P.opr.copy()
    s_p, s_m = 1 + 2 * d, 2 + 2 * d
    act = opr[0] != 0.0
    opr[0] = np.where(act, opr[0] - (f - 1.0) * (opr[s_p] + opr[s_m]), 0.0)
    opr[s_p] = np.where(act, f * opr[s_p], opr[s_p])
    opr[s_m] = np.where(act, f * opr[s_m], opr[s_m])
    return opr


@pytest.mark.parametrize("name", ["channel16", "rect40x6x18", "masked_block16x11x8", "rows128x64x128"])
@pytest.mark.parametrize("precision", [0, MIXED])
def test_setup_again_matches_fresh_setup(name, precision):
    """A setup on a handle that has been set up before first rebuilds the
    previous hierarchy with its coarsening directions and checks them against
    the new strengths afterwards. Through a sequence of operators whose
    strongest direction changes, each setup gives the hierarchy (shapes,
    directions, coefficients) and the solve of a fresh handle, bit for bit."""
    print(f"\n[setup again: {name} prec={precision}]", flush=True)
    import torch

    from nga2_amd.linsol import DeviceSolve

    P = CASES[name]()
    seq = [P.opr, _scaled(P, 0, 30.0), _scaled(P, 0, 30.0), _scaled(P, 2, 30.0), P.opr]
    ps = _solver(P, precision=precision)
    # the same sequence through the device entry points (k_level0 also
    # writes the solver's copy of the operator there)
    dev = torch.device("cuda:0")
    to_dev = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a).ravel(order="F"))).to(dev)
    ds = DeviceSolve(P.cfg)
    ds.setup(to_dev(seq[0]), maxit=200, rcvg=1e-10, precision=precision)
    d_rhs = to_dev(P.rhs)
    d_sol = torch.zeros_like(d_rhs)
    dirs = []
    for q, opr in enumerate(seq[1:], 1):
        ps.opr[:] = opr
        ps.setup()
        ds.setup(to_dev(opr), maxit=200, rcvg=1e-10, precision=precision)
        fresh = NgapStr(P.cfg, "Pressure", PCG_PFMG, 7)
        fresh.stc[:] = STC7
        fresh.maxit, fresh.rcvg, fresh.precision = 200, 1e-10, precision
        fresh.opr[:] = opr
        fresh.init()
        fresh.setup()
        ha, hb, hd = ps.hooks(), fresh.hooks(), ds.hooks()
        assert ha.nlevels == hb.nlevels == hd.nlevels, q
        info = [ha.level_info(l) for l in range(ha.nlevels)]
        assert info == [hb.level_info(l) for l in range(hb.nlevels)], q
        assert info == [hd.level_info(l) for l in range(hd.nlevels)], q
        dirs.append(tuple(i[1] for i in info))
        for l in range(ha.nlevels):
            ca, cb, cd = ha.level_coefs(l), hb.level_coefs(l), hd.level_coefs(l)
            for k in ca:
                assert np.array_equal(ca[k], cb[k]) and np.array_equal(cd[k], cb[k]), (q, l, k)
        for s_ in (ps, fresh):
            s_.rhs[:] = P.rhs
            s_.sol[:] = 0.0
            s_.solve()
        d_sol.zero_()
        it_d, _ = ds.solve(d_rhs, d_sol)
        assert ps.it == fresh.it and np.array_equal(ps.sol, fresh.sol), q
        x_d = d_sol.cpu().numpy().reshape(P.rhs.shape, order="F")
        assert it_d == fresh.it and np.array_equal(interior(x_d, P.cfg), interior(fresh.sol, P.cfg)), q
        assert ps.stats()["true_rerr"] <= 1e-10 and ds.stats()["true_rerr"] <= 1e-10
        fresh.destroy()
    print(f"  coarsening directions per setup: {dirs}", flush=True)
    # the sequence does change the hierarchy's directions
    assert len(set(dirs)) >= 2
    ps.destroy()


@pytest.mark.parametrize("name", ["periodic16", "channel16", "walls13x10x7", "masked_spheres16"])
@pytest.mark.parametrize("coarsest", [16, 128, 512, 1024])
def test_coarsest_inverse_matches_oracle(name, coarsest):
    """The coarsest level's dense (pseudo-)inverse: built on the GPU in one
    workgroup for levels of <= 128 cells, on the host above that. Both apply
    like the oracle's coarse solve, singular (periodic) levels included, and
    the full solve still meets the oracle."""
    P = CASES[name]()
    o = Oracle.from_cfg(P.cfg, P.opr, coarsest_max=coarsest)
    ps = _solver(P, coarsest_max=coarsest)
    hk = ps.hooks()
    assert hk.nlevels == o.nlevels
    n = o.level_size(o.nlevels - 1)
    print(f"\n[coarsest {name} max {coarsest}] coarsest level {n} cells, singular {o.coarse_singular}", flush=True)
    rng = np.random.default_rng(coarsest)
    for _ in range(3):
        fc = rng.standard_normal(n)
        if o.coarse_singular:
            fc -= fc.mean()
        assert _close(hk.coarse_solve(fc), o.coarse_solve(fc), 1e-9)
    assert ps.stats()["coarse_singular"] == int(o.coarse_singular)
    ps.rhs[:] = P.rhs
    ps.sol[:] = 0.0
    ps.solve()
    assert ps.stats()["status"] == 0 and ps.stats()["true_rerr"] <= 1e-10
    ps.destroy()
    o.close()
