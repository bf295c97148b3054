import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

STC7 = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], dtype=np.int32)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libngapsolve.so on cuda:0)")


@pytest.fixture(scope="session", autouse=True)
def _torch_hip_first(request):
    """torch ships its own HIP runtime; it must initialise the device before
    libngapsolve.so (system ROCm) does, or torch reports no GPU afterwards."""
    if any(item.get_closest_marker("gpu") for item in request.session.items):
        import torch

        if torch.cuda.is_available():
            torch.cuda.init()
            torch.cuda.set_device(0)
    yield


def closed_form_periodic(b_flat, n, h):
    """Exact solution of the uniform periodic 7-point problem -vol*lap x = b
    (vol = h^3) by FFT, zero mean (the fourier3d method,
    src/solver/fourier3d_class.f90:1103-1196)."""
    k = np.arange(n)
    lam1 = h * (2.0 - 2.0 * np.cos(2.0 * np.pi * k / n))
    lam = lam1[:, None, None] + lam1[None, :, None] + lam1[None, None, :]
    B = np.fft.fftn(b_flat.reshape(n, n, n))
    lam[0, 0, 0] = 1.0
    X = B / lam
    X[0, 0, 0] = 0.0
    return np.fft.ifftn(X).real.ravel()


def assemble_csr(cfg, opr):
    """Sparse matrix of the exact 7-point operator over interior cells
    (identity rows where the centre is 0; off-grid couplings dropped)."""
    import scipy.sparse as sp
    from oracle.orc import interior_opr

    op = interior_opr(opr, cfg).copy()
    N = cfg.nx * cfg.ny * cfg.nz
    part = np.abs(op[..., 0]) > 0
    op[~part] = 0.0
    op[..., 0][~part] = 1.0
    idx = np.arange(N).reshape(cfg.nz, cfg.ny, cfg.nx)
    k, j, i = np.meshgrid(np.arange(cfg.nz), np.arange(cfg.ny), np.arange(cfg.nx), indexing="ij")
    rows, cols, vals = [], [], []
    per = cfg.periodic
    for s, (dx, dy, dz) in enumerate(STC7):
        ii, jj, kk = i + dx, j + dy, k + dz
        ok = np.ones_like(ii, dtype=bool)
        for a, n, p in ((ii, cfg.nx, per[0]), (jj, cfg.ny, per[1]), (kk, cfg.nz, per[2])):
            if p:
                a %= n
            else:
                ok &= (a >= 0) & (a < n)
        rows.append(idx[ok])
        cols.append(idx[kk[ok], jj[ok], ii[ok]])
        vals.append(op[..., s][ok])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))


def direct_solve_singular(A, b):
    """Direct solve of a singular (constant null space) system: pin one DOF,
    then remove the mean."""
    import scipy.sparse.linalg as spl

    A2 = A.tolil()
    A2[0, :] = 0.0
    A2[0, 0] = 1.0
    b2 = b.copy()
    b2[0] = 0.0
    x = spl.spsolve(A2.tocsr(), b2)
    return x - x.mean()


def rel_err_mean_free(x, xr):
    x = x - x.mean()
    xr = xr - xr.mean()
    return np.linalg.norm(x - xr) / np.linalg.norm(xr)


def rel_err_nullspace(x, xr, cfg, opr):
    """Relative L2 difference after removing, from both solutions, their mean
    over each connected region of participating cells (opr(1) != 0, linked by
    nonzero couplings, periodic). The singular system fixes the solution only
    up to one constant per region; with a single region and no identity rows
    this is rel_err_mean_free."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components

    from oracle.orc import interior_opr

    op = interior_opr(opr, cfg)  # [k, j, i, stencil]
    act = (op[..., 0] != 0.0).ravel()
    N = cfg.nx * cfg.ny * cfg.nz
    idx = np.arange(N).reshape(cfg.nz, cfg.ny, cfg.nx)
    rows, cols = [], []
    for s, d in ((1, 0), (3, 1), (5, 2)):
        ax = 2 - d  # array axis of direction d
        cpl = (op[..., s] != 0.0)
        if not cfg.periodic[d]:
            last = [slice(None)] * 3
            last[ax] = -1
            cpl[tuple(last)] = False
        rows.append(idx[cpl])
        cols.append(np.roll(idx, -1, axis=ax)[cpl])
    G = sp.coo_matrix((np.ones(sum(len(r) for r in rows)), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(N, N))
    _, lab = connected_components(G, directed=False)
    x = np.asarray(x, dtype=np.float64).ravel().copy()
    xr = np.asarray(xr, dtype=np.float64).ravel().copy()
    for c in np.unique(lab[act]):
        m = act & (lab == c)
        x[m] -= x[m].mean()
        xr[m] -= xr[m].mean()
    x[~act] = 0.0
    xr[~act] = 0.0
    return np.linalg.norm(x - xr) / np.linalg.norm(xr)
