"""CPU pins of the hypre_uns oracle (oracle/amg_ref.py), the checker of
libngapuns.so, and of the library's host-side setup against it.

HYPRE is absent, so nothing the reference's tests hold pins the multigrid;
each piece is pinned by what the mathematics fixes, chosen so that a
plausible slip fails a check:
- the numbering and the CSR rows against an independent construction from
  the (i, j, k) offsets (a transposed axis, a dropped periodic wrap or a kept
  masked column changes them);
- strength and PMIS by their defining properties (no two C points strongly
  connected, every F point with a strong C dependency unless nothing depends
  on it);
- extended+i interpolation: identity on C points, weights 1/2, 1/2 between two
  C points of a 1-D Laplacian, constants reproduced on zero-row-sum rows (a
  wrong sign or a dropped lumped term breaks the row sum);
- A_c against the dense product P^T A P;
- the V-cycle symmetric and positive definite on the range (PCG needs both);
- every provided method's solution against scipy's direct solve;
and the library's host setup (numbering, CSR, C/F split, P, every A_c) is
compared with the oracle entry by entry.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from conftest import STC7
from nga2_amd import _lib, inputs, uns
from nga2_amd.config import Config, uniform
from oracle import amg_ref as R


def _problem(name):
    if name == "masked12":
        return inputs.masked_spheres(n=12, nsph=4, rmin=0.15, rmax=0.25)
    if name == "channel8":
        return inputs.channel_spectral(n=8)
    if name == "implicit8":
        from nga2_amd.config import tanh_stretched

        cfg = Config(uniform(8), tanh_stretched(8), uniform(8), no=1, yper=False, name="implicit")
        return inputs.implicit_problem(cfg, visc=0.02)
    return inputs.periodic_gaussian(n=8)


def _csr(P):
    ind, n = R.umap(P.cfg, P.opr)
    rp, c, v = R.assemble(P.cfg, ind, P.opr, STC7)
    return ind, R.CSR(rp, c, v, n)


def test_umap_and_assembly_against_offsets():
    P = _problem("masked12")
    cfg, opr = P.cfg, P.opr
    ind, A = _csr(P)
    it = cfg.interior
    act = np.abs(opr[0][it]) > 10 * np.finfo(float).eps
    assert A.n == act.sum() and 0 < A.n < cfg.nx * cfg.ny * cfg.nz
    # numbering: i fastest over the participating cells, -1 elsewhere
    num = -np.ones(act.shape, dtype=np.int64)
    num.T[act.T] = np.arange(A.n)
    assert np.array_equal(ind[it], num)
    # ghosts hold the periodic images
    chk = ind.copy()
    cfg.sync(chk)
    assert np.array_equal(chk, ind)
    # rows from the offsets, masked neighbours dropped
    n = cfg.nx
    D = np.zeros((A.n, A.n))
    O = opr[(slice(None),) + it]
    for (i, j, k) in zip(*np.nonzero(act)):
        r = num[i, j, k]
        for s, (dx, dy, dz) in enumerate(STC7):
            c = num[(i + dx) % n, (j + dy) % n, (k + dz) % n]
            if c >= 0:
                D[r, c] += O[s, i, j, k]
    assert np.array_equal(A.dense(), D)
    # the masked operator is symmetric and its rows sum to zero (Neumann at solid faces)
    assert np.abs(D - D.T).max() <= 1e-15 * np.abs(D).max()
    assert np.abs(D.sum(axis=1)).max() <= 1e-14 * np.abs(D).max()


def test_repeated_columns_merge():
    """A periodic direction of two cells makes +x and -x the same neighbour."""
    cfg = Config(uniform(2), uniform(3), uniform(3), no=1)
    from nga2_amd import incomp

    opr = incomp.laplacian(incomp.Metrics(cfg))
    ind, A = _csr(inputs.Problem(cfg, opr, cfg.field(), "tiny"))
    for i in range(A.n):
        c, _ = A.row(i)
        assert len(set(c.tolist())) == len(c)
    assert np.abs(A.dense().sum(axis=1)).max() <= 1e-14 * np.abs(A.val).max()


@pytest.mark.parametrize("name", ["masked12", "channel8", "periodic8"])
def test_strength_and_pmis_properties(name):
    _, A = _csr(_problem(name))
    S = R.strength(A)
    dep, inf = R.strong_lists(A, S)
    # strength by its definition
    for i in range(A.n):
        c, v = A.row(i)
        off = v[c != i]
        if len(off) == 0 or off.min() >= 0:
            assert not dep[i]
            continue
        assert set(dep[i]) == {int(c[q]) for q in range(len(c)) if c[q] != i and v[q] < 0.25 * off.min()}
    cf = R.pmis(A, S)
    assert set(np.unique(cf)) <= {R.C_PT, R.F_PT}
    for i in range(A.n):
        if cf[i] == R.C_PT:
            assert not any(cf[j] == R.C_PT for j in dep[i] + inf[i])
        elif inf[i]:
            # an F point some point depends on became F through a C dependency
            assert any(cf[j] == R.C_PT for j in dep[i])


def test_ext_i_one_dimensional():
    """1-D Laplacian: an F point between two C points gets weights 1/2, 1/2;
    every row of P sums to one (constants are reproduced)."""
    n = 32
    rp, col, val = [0], [], []
    for i in range(n):
        for c, v in ((i, 2.0), ((i - 1) % n, -1.0), ((i + 1) % n, -1.0)):
            col.append(c)
            val.append(v)
        rp.append(len(col))
    A = R.CSR(rp, col, val, n)
    S = R.strength(A)
    cf = R.pmis(A, S)
    P = R.interp_ext_i(A, S, cf)
    cmap = np.cumsum(cf == R.C_PT) - 1
    hit = 0
    for i in range(n):
        c, v = P.row(i)
        assert abs(v.sum() - 1.0) <= 1e-15
        if cf[i] == R.C_PT:
            assert list(c) == [cmap[i]] and list(v) == [1.0]
        elif cf[(i - 1) % n] == R.C_PT and cf[(i + 1) % n] == R.C_PT:
            assert sorted(c.tolist()) == sorted([cmap[(i - 1) % n], cmap[(i + 1) % n]])
            assert np.allclose(v, 0.5, rtol=0, atol=1e-15)
            hit += 1
    assert hit > 0


@pytest.mark.parametrize("name", ["masked12", "channel8"])
def test_interpolation_rows_and_galerkin(name):
    _, A = _csr(_problem(name))
    H = R.build(A, max_coarse=40)
    assert len(H.A) >= 3
    for l in range(len(H.P)):
        Al, Pl, cf = H.A[l], H.P[l], H.cf[l]
        Pd = Pl.dense()
        # C rows are identity rows
        cm = np.cumsum(cf == R.C_PT) - 1
        for i in np.nonzero(cf == R.C_PT)[0]:
            row = np.zeros(Pl.ncol)
            row[cm[i]] = 1.0
            assert np.array_equal(Pd[i], row)
        # rows of a zero-row-sum operator: interpolation reproduces constants
        Ad = Al.dense()
        zero = np.abs(Ad.sum(axis=1)) <= 1e-12 * np.abs(np.diag(Ad))
        nonempty = np.abs(Pd).sum(axis=1) > 0
        sel = zero & nonempty
        assert sel.sum() > 0
        assert np.abs(Pd[sel].sum(axis=1) - 1.0).max() <= 1e-12
        assert (np.count_nonzero(Pd, axis=1) <= 4).all()
        # Galerkin operator
        Ac = Pd.T @ Ad @ Pd
        assert np.abs(H.A[l + 1].dense() - Ac).max() <= 1e-13 * np.abs(Ac).max()


def test_vcycle_symmetric_positive():
    P = _problem("masked12")
    _, A = _csr(P)
    H = R.build(A, max_coarse=40)
    rng = np.random.default_rng(3)
    u, v = rng.standard_normal(A.n), rng.standard_normal(A.n)
    Mu, Mv = R.vcycle(H, u), R.vcycle(H, v)
    assert abs(u @ Mv - v @ Mu) <= 1e-12 * np.linalg.norm(u) * np.linalg.norm(Mv)
    assert u @ Mu > 0 and v @ Mv > 0
    # the cycle as a stationary iteration e <- (I - M A) e contracts in the
    # energy norm: measured factor over 10 cycles from a random error
    Ad = A.dense()
    e = rng.standard_normal(A.n)
    e -= e.mean()
    en = lambda x: np.sqrt(x @ Ad @ x)
    e0 = en(e)
    for _ in range(10):
        e = e - R.vcycle(H, Ad @ e)
    rho = (en(e) / e0) ** 0.1
    assert rho < 0.7, rho


@pytest.mark.parametrize("method", [R.AMG, R.PCG_AMG, R.GMRES_AMG, R.BICGSTAB_AMG, R.PCG, R.GMRES])
def test_methods_against_direct_solve(method):
    P = _problem("masked12")
    ind, A = _csr(P)
    H = R.build(A, max_coarse=40)
    b = P.rhs[P.cfg.interior].ravel(order="F")[(np.abs(P.opr[0][P.cfg.interior]) > 1e-14).ravel(order="F")]
    assert abs(b.sum()) <= 1e-12 * np.abs(b).sum()
    As = sp.csr_matrix(A.dense())
    A2 = As.tolil()
    A2[0, :] = 0.0
    A2[0, 0] = 1.0
    b2 = b.copy()
    b2[0] = 0.0
    xd = spl.spsolve(A2.tocsr(), b2)
    xd -= xd.mean()
    x, it, rerr = R.solve(method, A, H, b, np.zeros(A.n), 400, 1e-10)
    assert rerr < 1e-10 and 0 < it < 400
    x = x - x.mean()
    assert np.linalg.norm(x - xd) <= 1e-7 * np.linalg.norm(xd)
    with pytest.raises(ValueError):
        R.solve(R.PCG_PARASAIL, A, H, b, np.zeros(A.n), 10, 1e-10)


@pytest.mark.parametrize("method", [R.AMG, R.GMRES_AMG, R.BICGSTAB_AMG, R.GMRES])
def test_nonsymmetric_methods_against_direct_solve(method):
    """The implicit convection-diffusion operator (nonsymmetric, nonsingular):
    the methods hypre_uns offers for it reach scipy's direct solution."""
    P = _problem("implicit8")
    ind, A = _csr(P)
    D = A.dense()
    assert np.abs(D - D.T).max() > 0.1 * np.abs(D).max()
    H = R.build(A, max_coarse=40)
    b = P.rhs[P.cfg.interior].ravel(order="F")
    xd = spl.spsolve(sp.csr_matrix(D), b)
    x, it, rerr = R.solve(method, A, H, b, np.zeros(A.n), 300, 1e-10)
    assert rerr < 1e-10 and np.linalg.norm(x - xd) <= 1e-8 * np.linalg.norm(xd)


@pytest.mark.parametrize("name", ["masked12", "channel8", "periodic8", "implicit8"])
def test_library_host_setup_matches_oracle(name):
    P = _problem(name)
    cfg, opr = P.cfg, P.opr
    ind_o, A = _csr(P)
    ind_c, n = uns.umap_host(cfg, opr)
    assert n == A.n and np.array_equal(ind_c, ind_o)
    rp, col, val = uns.assemble_host(cfg, STC7, ind_c, opr)
    assert np.array_equal(rp, A.rowptr) and np.array_equal(col, A.col) and np.array_equal(val, A.val)
    H = R.build(A, max_coarse=40)
    Hc = uns.Hierarchy.build(rp, col, val, max_coarse=40)
    assert Hc.nlevels() == len(H.A)
    for l in range(len(H.A)):
        lv = Hc.level(l)
        lrp, lcol, lval = lv["A"]
        assert np.array_equal(lrp, H.A[l].rowptr) and np.array_equal(lcol, H.A[l].col)
        assert np.array_equal(lval, H.A[l].val)  # same operations in the same order
        if l < len(H.P):
            assert np.array_equal(lv["cf"], H.cf[l])
            prp, pcol, pval = lv["P"]
            assert np.array_equal(prp, H.P[l].rowptr) and np.array_equal(pcol, H.P[l].col)
            assert np.array_equal(pval, H.P[l].val)
    Hc.free()


def test_library_rejects_bad_input():
    with pytest.raises(uns.NgapError):
        uns.Hierarchy.build(np.array([0, 1]), np.array([3]), np.array([1.0]))  # column out of range
    o = uns.default_opts()
    assert (o.strong_threshold, o.max_row_sum, o.pmax, o.max_coarse, o.max_levels, o.nu, o.kdim) == \
        (0.25, 0.9, 4, 512, 25, 1, 5)


def test_library_rejects_bad_sizes():
    """Malformed row pointers, knobs out of range, and a hierarchy whose
    coarsening stops above the dense-coarsest limit (4096 rows) are refused
    with a code, not computed."""
    lap = sp.diags([-np.ones(9), 2 * np.ones(10), -np.ones(9)], [-1, 0, 1], format="csr")
    rp, col, val = lap.indptr, lap.indices, lap.data
    with pytest.raises(uns.NgapError) as e:
        uns.Hierarchy.build(rp + 1, col, val)  # rowptr[0] != 0
    assert e.value.code == _lib.ERR_ARG
    bad = rp.copy()
    bad[5] = bad[4] - 1  # decreasing row pointer
    with pytest.raises(uns.NgapError):
        uns.Hierarchy.build(bad, col, val)
    for knob in ({"max_coarse": 4097}, {"max_coarse": 0}, {"strong_threshold": 1.5}, {"max_row_sum": 0.0}):
        with pytest.raises(uns.NgapError) as e:
            uns.Hierarchy.build(rp, col, val, **knob)
        assert e.value.code == _lib.ERR_ARG, knob
    H = uns.Hierarchy.build(rp, col, val, max_coarse=4096)  # the limit itself is accepted
    H.free()
    # a diagonal matrix has no strong connections: no coarse points, so the
    # coarsest level is the whole matrix, here above the dense limit
    n = 5000
    with pytest.raises(uns.NgapError) as e:
        uns.Hierarchy.build(np.arange(n + 1), np.arange(n), np.full(n, 2.0))
    assert e.value.code == _lib.ERR_COARSE
