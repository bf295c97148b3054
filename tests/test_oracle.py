"""Pins of the CPU oracle (oracle/pcg_pfmg.c) — no GPU needed.

Each check is chosen so that a plausible mistake (dropped term, wrong sign or
index, transposed transfer) fails: closed-form FFT solutions, sparse direct
solves on the configs' geometry, manufactured solutions, textbook CG on dense
SPD matrices, and the algebraic identities of the multigrid pieces.
"""
import numpy as np
import pytest

from conftest import assemble_csr, closed_form_periodic, direct_solve_singular, rel_err_mean_free
from nga2_amd import inputs
from oracle.orc import Oracle, interior, pcg_dense


@pytest.fixture(scope="module")
def periodic16():
    P = inputs.periodic_gaussian(n=16)
    return P, Oracle.from_cfg(P.cfg, P.opr)


@pytest.mark.parametrize("n", [8, 16, 32])
def test_closed_form_periodic(n):
    P = inputs.periodic_gaussian(n=n)
    o = Oracle.from_cfg(P.cfg, P.opr)
    b = interior(P.rhs, P.cfg)
    x, st = o.solve(b, rcvg=1e-12, maxit=100)
    assert st["true_rerr"] <= 1e-12
    xr = closed_form_periodic(b, n, 1.0 / n)
    assert rel_err_mean_free(x, xr) < 1e-10
    assert st["it"] <= 20


def test_direct_channel_stretched():
    P = inputs.channel_spectral(n=16)
    A = assemble_csr(P.cfg, P.opr)
    b = interior(P.rhs, P.cfg)
    xd = direct_solve_singular(A, b)
    o = Oracle.from_cfg(P.cfg, P.opr)
    x, st = o.solve(b, rcvg=1e-12, maxit=200)
    assert st["true_rerr"] <= 1e-12
    assert rel_err_mean_free(x, xd) < 1e-10


def test_direct_spheres_density_ratio():
    P = inputs.spheres_two_phase(n=24, nsph=6, rmin=0.1, rmax=0.25)
    A = assemble_csr(P.cfg, P.opr)
    b = interior(P.rhs, P.cfg)
    xd = direct_solve_singular(A, b)
    o = Oracle.from_cfg(P.cfg, P.opr)
    x, st = o.solve(b, rcvg=1e-12, maxit=200)
    assert st["true_rerr"] <= 1e-12
    assert rel_err_mean_free(x, xd) < 1e-10


def test_golden_fixtures():
    """Committed fixtures (tests/golden/make_fixtures.py): closed-form and
    direct solutions at 16^3 / 24^3."""
    import os

    path = os.path.join(os.path.dirname(__file__), "golden", "fixtures.npz")
    d = np.load(path, allow_pickle=False)
    for name in ("periodic16", "channel16", "spheres24"):
        P = _rebuild(name)
        b = interior(P.rhs, P.cfg)
        assert np.allclose(b, d[f"{name}_rhs"], rtol=0, atol=1e-13 * np.abs(b).max())
        o = Oracle.from_cfg(P.cfg, P.opr)
        x, st = o.solve(b, rcvg=1e-12, maxit=200)
        assert rel_err_mean_free(x, d[f"{name}_sol"]) < 1e-10


def _rebuild(name):
    if name == "periodic16":
        return inputs.periodic_gaussian(n=16)
    if name == "channel16":
        return inputs.channel_spectral(n=16)
    return inputs.spheres_two_phase(n=24, nsph=6, rmin=0.1, rmax=0.25)


def test_manufactured_neumann_odd_sizes():
    """x* random, b = A x*, the solve recovers x* (mean free) on a box with
    walls in every direction and odd, unequal sizes."""
    from nga2_amd import incomp
    from nga2_amd.config import Config, uniform

    cfg = Config(uniform(13), uniform(10), uniform(7), no=1, xper=False, yper=False, zper=False)
    cfg.VF[:] = 0.0
    cfg.VF[cfg.interior] = 1.0
    met = incomp.Metrics(cfg)
    opr = incomp.laplacian(met)
    o = Oracle.from_cfg(cfg, opr)
    rng = np.random.default_rng(5)
    xs = rng.standard_normal(13 * 10 * 7)
    xs -= xs.mean()
    b = o.apply_exact(xs)
    x, st = o.solve(b, rcvg=1e-13, maxit=300)
    assert rel_err_mean_free(x, xs) < 1e-9
    assert o.coarse_singular


def test_manufactured_identity_rows():
    """Cells with opr(1)==0 become identity rows with rhs = sol = 0
    (hypre_str_class.f90:174-175, 339-345)."""
    P = inputs.periodic_gaussian(n=8)
    opr = P.opr.copy()
    it = P.cfg.interior
    # knock out a 2x2x2 block: centre 0 and neighbours decoupled from it
    blk = (slice(3, 5), slice(3, 5), slice(3, 5))
    sub = opr[(slice(None),) + it]
    sub[(slice(None),) + blk] = 0.0
    opr[(slice(None),) + it] = sub
    o = Oracle.from_cfg(P.cfg, opr)
    rng = np.random.default_rng(1)
    xs = rng.standard_normal(512)
    mask = np.zeros((8, 8, 8), bool)
    mask[blk] = True
    dead = np.transpose(mask, (2, 1, 0)).ravel()
    xs[dead] = 0.0
    b = o.apply_exact(xs)
    x, st = o.solve(b, rcvg=1e-13, maxit=300)
    assert np.all(x[dead] == 0.0)
    # the participating part is determined up to what the exact operator fixes
    assert np.linalg.norm(o.apply_exact(x) - b) <= 1e-12 * np.linalg.norm(b)


def test_pcg_dense_textbook():
    rng = np.random.default_rng(7)
    n = 60
    Q = rng.standard_normal((n, n))
    A = Q @ Q.T + n * np.eye(n)
    b = rng.standard_normal(n)
    xe = np.linalg.solve(A, b)
    for Minv in (None, np.diag(1.0 / np.diag(A))):
        for flex in (False, True):
            x, st = pcg_dense(A, b, Minv=Minv, rcvg=1e-13, maxit=500, flexible=flex)
            assert np.linalg.norm(x - xe) / np.linalg.norm(xe) < 1e-11
            assert st["it"] <= n
    # with the exact inverse as preconditioner CG converges in one step
    x, st = pcg_dense(A, b, Minv=np.linalg.inv(A), rcvg=1e-12, maxit=10)
    assert st["it"] == 1


def test_pcg_matches_reference_recurrence():
    """Five iterations of the oracle's PCG against a direct numpy transcription
    of preconditioned CG (Hestenes-Stiefel) with the same preconditioner."""
    rng = np.random.default_rng(3)
    n = 40
    Q = rng.standard_normal((n, n))
    A = Q @ Q.T + 0.5 * n * np.eye(n)
    M = np.diag(1.0 / (np.diag(A) + rng.random(n)))
    b = rng.standard_normal(n)
    x = np.zeros(n)
    r = b.copy()
    z = M @ r
    p = z.copy()
    rho = r @ z
    for _ in range(5):
        q = A @ p
        a = rho / (p @ q)
        x += a * p
        r -= a * q
        z = M @ r
        rn = r @ z
        p = z + (rn / rho) * p
        rho = rn
    xo, st = pcg_dense(A, b, Minv=M, rcvg=1e-300, maxit=5)
    assert st["it"] == 5
    assert np.allclose(xo, x, rtol=1e-12, atol=1e-14)


def test_restriction_is_prolongation_transpose(periodic16):
    P, o = periodic16
    for l in range(o.nlevels - 1):
        n = o.level_size(l)
        nc = o.level_size(l + 1)
        Pm = np.zeros((n, nc))
        for J in range(nc):
            e = np.zeros(nc)
            e[J] = 1.0
            Pm[:, J] = o.prolong_add(l, e, np.zeros(n))
        R = np.zeros((nc, n))
        zero = np.zeros(n)
        for i in range(n):
            f = np.zeros(n)
            f[i] = 1.0
            R[:, i] = o.resid_restrict(l, f, zero)
        assert np.allclose(R, Pm.T, atol=1e-15)
        # interpolation reproduces constants (zero row sums)
        assert np.allclose(Pm.sum(axis=1), 1.0)
        if l == 0:
            break


@pytest.mark.parametrize("which", ["periodic", "channel", "spheres"])
def test_vcycle_symmetric_positive(which):
    P = {"periodic": lambda: inputs.periodic_gaussian(n=16),
         "channel": lambda: inputs.channel_spectral(n=16),
         "spheres": lambda: inputs.spheres_two_phase(n=16, nsph=4, rmin=0.15, rmax=0.3)}[which]()
    o = Oracle.from_cfg(P.cfg, P.opr)
    rng = np.random.default_rng(11)
    N = o.N
    u, v = rng.standard_normal(N), rng.standard_normal(N)
    u -= u.mean()
    v -= v.mean()
    Mu, Mv = o.vcycle(u), o.vcycle(v)
    assert abs(Mu @ v - u @ Mv) <= 1e-12 * np.linalg.norm(Mu) * np.linalg.norm(v)
    assert u @ Mu > 0 and v @ Mv > 0


@pytest.mark.parametrize("which", ["periodic", "channel", "spheres"])
def test_coarse_operators_symmetric_consistent(which):
    P = {"periodic": lambda: inputs.periodic_gaussian(n=16),
         "channel": lambda: inputs.channel_spectral(n=16),
         "spheres": lambda: inputs.spheres_two_phase(n=16, nsph=4, rmin=0.15, rmax=0.3)}[which]()
    o = Oracle.from_cfg(P.cfg, P.opr)
    for l in range(o.nlevels):
        N = o.level_size(l)
        ones = np.ones(N)
        # zero row sums (constant null space) at every level, M-matrix signs
        assert np.abs(o.apply(l, ones)).max() <= 1e-12 * np.abs(o.level_coefs(l)["c"]).max()
        c = o.level_coefs(l)
        assert np.all(c["c"] > 0)
        for a in ("ax", "ay", "az"):
            assert np.all(c[a] <= 0)
        # symmetry of the level operator: <A u, v> = <u, A v>
        rng = np.random.default_rng(l)
        u, v = rng.standard_normal(N), rng.standard_normal(N)
        assert abs(o.apply(l, u) @ v - u @ o.apply(l, v)) <= 1e-12 * np.linalg.norm(u) * np.linalg.norm(v) * c["c"].max()


def test_vcycle_iteration_contracts():
    """The V-cycle as a stationary iteration x <- x + M(b - A x) contracts the
    A-norm error on the stretched channel and on the density-ratio problem."""
    for P in (inputs.channel_spectral(n=16), inputs.spheres_two_phase(n=16, nsph=4, rmin=0.15, rmax=0.3)):
        o = Oracle.from_cfg(P.cfg, P.opr)
        b = interior(P.rhs, P.cfg)
        xs, _ = o.solve(b, rcvg=1e-14, maxit=300)
        x = np.zeros_like(b)
        errs = []
        for _ in range(6):
            x = x + o.vcycle(b - o.apply(0, x))
            e = x - xs
            e -= e.mean()
            errs.append(np.sqrt(e @ o.apply(0, e)))
        rates = [errs[i + 1] / errs[i] for i in range(len(errs) - 1)]
        assert max(rates) < 0.9, rates


def test_coarsening_rule():
    """Semicoarsening picks the strongest direction; on the uniform cube it
    cycles x, y, z; it stops at <= coarsest_max cells."""
    P = inputs.periodic_gaussian(n=32)
    o = Oracle.from_cfg(P.cfg, P.opr)
    dirs = [o.level_info(l)[1] for l in range(o.nlevels)]
    assert dirs[:6] == [0, 1, 2, 0, 1, 2]
    assert o.level_size(o.nlevels - 1) <= 128
    # the stretched channel coarsens y first (strongest coupling near walls)
    Pc = inputs.channel_spectral(n=16)
    oc = Oracle.from_cfg(Pc.cfg, Pc.opr)
    assert oc.level_info(0)[1] == 1
    # maxlevel caps the hierarchy
    o3 = Oracle.from_cfg(P.cfg, P.opr, maxlevel=3)
    assert o3.nlevels == 3


def test_odd_periodic_rejected():
    from nga2_amd import incomp
    from nga2_amd.config import Config, uniform

    cfg = Config(uniform(9), uniform(8), uniform(8), no=1)
    opr = incomp.laplacian(incomp.Metrics(cfg))
    with pytest.raises(RuntimeError):
        Oracle.from_cfg(cfg, opr)


def test_skip_relax_rule_and_cycle():
    """Skipped relaxation: on the uniform periodic cube semicoarsening x, y, z
    takes the direction strengths (1,1,1) -> (1/2,2,2) -> (1,1,4) -> (2,2,2), so
    exactly every third level is smoothed (and the coarsest is solved). With
    the skip the V-cycle stays symmetric positive definite, equals the cycle
    that smooths every level when nothing is skipped (a two-level hierarchy),
    and still contracts as a stationary iteration on the two-phase problem."""
    P = inputs.periodic_gaussian(n=32)
    o = Oracle.from_cfg(P.cfg, P.opr, skip_relax=1)
    rl = o.relax_levels()
    nl = o.nlevels
    assert rl == [1 if (l % 3 == 0 or l == nl - 1) else 0 for l in range(nl)], rl
    # the strengths behind the rule, from the level coefficients
    for l in range(1, nl - 1):
        c = o.level_coefs(l)
        s = sorted(np.abs(c[a]).sum() for a in ("ax", "ay", "az"))
        assert (s[2] <= 2 * (1 + 1e-6) * s[0]) == bool(rl[l]), (l, s)
    rng = np.random.default_rng(5)
    u, v = rng.standard_normal(o.N), rng.standard_normal(o.N)
    u -= u.mean()
    v -= v.mean()
    Mu, Mv = o.vcycle(u), o.vcycle(v)
    assert abs(Mu @ v - u @ Mv) <= 1e-12 * np.linalg.norm(Mu) * np.linalg.norm(v)
    assert u @ Mu > 0 and v @ Mv > 0
    # two levels: nothing can be skipped, the cycles agree bit for bit
    o2s = Oracle.from_cfg(P.cfg, P.opr, maxlevel=2, skip_relax=1)
    o2 = Oracle.from_cfg(P.cfg, P.opr, maxlevel=2)
    assert np.array_equal(o2s.vcycle(u), o2.vcycle(u))
    # contraction with the skip on the density-ratio problem
    Ps = inputs.spheres_two_phase(n=16, nsph=4, rmin=0.15, rmax=0.3)
    os_ = Oracle.from_cfg(Ps.cfg, Ps.opr, skip_relax=1)
    b = interior(Ps.rhs, Ps.cfg)
    xs, _ = os_.solve(b, rcvg=1e-14, maxit=300)
    x = np.zeros_like(b)
    errs = []
    for _ in range(6):
        x = x + os_.vcycle(b - os_.apply(0, x))
        e = x - xs
        e -= e.mean()
        errs.append(np.sqrt(e @ os_.apply(0, e)))
    assert max(errs[i + 1] / errs[i] for i in range(5)) < 0.95


def test_skip_relax_gate():
    """The skip applies only to a locally isotropic fine operator: the
    tanh-stretched channel (wall cells with y couplings many times the x ones)
    smooths every level, and would lose many iterations to skipping; the
    two-phase spheres (a few interface cells) and the solid spheres
    (one-sided couplings at the walls, ratio 2) skip from level 1 on."""
    Pc = inputs.channel_spectral(n=64)
    oc = Oracle.from_cfg(Pc.cfg, Pc.opr, skip_relax=1, nu_pre=2, nu_post=2)
    assert all(r == 1 for r in oc.relax_levels())
    b = interior(Pc.rhs, Pc.cfg)
    _, st_all = oc.solve(b, rcvg=1e-10, maxit=500)
    oc.set_relax_levels([1 if (l == 0 or l == oc.nlevels - 1 or l % 3 == 0) else 0 for l in range(oc.nlevels)])
    _, st_skip = oc.solve(b, rcvg=1e-10, maxit=500)
    assert st_skip["it"] >= st_all["it"] + 4, (st_all, st_skip)
    for P in (inputs.spheres_two_phase(n=32, nsph=6, rmin=0.1, rmax=0.2),
              inputs.masked_spheres(n=32, nsph=4, rmin=0.1, rmax=0.2)):
        o = Oracle.from_cfg(P.cfg, P.opr, skip_relax=1)
        rl = o.relax_levels()
        assert rl[0] == 1 and rl[1] == 0 and rl[-1] == 1, rl
    # skip_relax=0 smooths every level
    o0 = Oracle.from_cfg(P.cfg, P.opr, skip_relax=0)
    assert all(r == 1 for r in o0.relax_levels())
