"""CPU checks of the drifting-spheres inputs: the torch restatements of the
sphere VF and of the face densities give the numpy generators' values bit for
bit (run with torch on the CPU here; the GPU uses the same code)."""
import numpy as np
import pytest
import torch

from nga2_amd import drift, inputs, learned
from nga2_amd.config import Config, uniform


def test_sphere_vf_torch_matches_numpy():
    n = 24
    x = uniform(n)
    cfg = Config(x, x, x, no=1, xper=True, yper=True, zper=True)
    centres, radii = inputs.sphere_set(7, 0.05, 0.2, 11)
    centres[0] = [0.01, 0.99, 0.5]  # crosses the periodic boundary
    ref = inputs.sphere_vf(cfg, centres, radii)
    got = drift.sphere_vf_device(n, centres, radii, device="cpu").numpy()
    assert np.array_equal(got, ref)
    assert 0.0 < ref.mean() < 1.0 and ((ref > 0) & (ref < 1)).any()


def test_face_densities_torch_matches_numpy():
    n = 12
    x = uniform(n)
    cfg = Config(x, x, x, no=1, xper=True, yper=True, zper=True)
    vf = np.random.default_rng(2).uniform(0, 1, (n, n, n))
    ref = inputs.face_densities(cfg, vf, 1000.0, 1.0)
    got = drift.face_densities_device(cfg, torch.from_numpy(vf), 1000.0, 1.0)
    for r, g in zip(ref, got):
        assert np.array_equal(g.numpy(), r.ravel(order="F"))


def test_drift_positions_wrap_and_speed():
    c, r = inputs.sphere_set(200, 0.02, 0.08, 5005)
    rng = np.random.Generator(np.random.PCG64(5006))
    v = rng.standard_normal((200, 3))
    v = v / np.sqrt((v * v).sum(axis=1, keepdims=True).mean()) * (0.5 / 256)
    p = (c + 40 * v) % 1.0
    assert (p >= 0).all() and (p < 1).all()
    # rms drift of 0.5 cell per step
    assert abs(np.sqrt((v * v).sum(axis=1).mean()) * 256 - 0.5) < 1e-12


def _uniform_vf_reference(n, centres, radii, nsub):
    """Brute force: every cut cell sampled at all nsub^3 leaf midpoints."""
    from nga2_amd.config import Config, uniform

    h = 1.0 / n
    xc = (np.arange(n) + 0.5) * h
    X, Y, Z = np.meshgrid(xc, xc, xc, indexing="ij")
    off = ((np.arange(nsub) + 0.5) / nsub - 0.5) * h
    vf = np.zeros((n, n, n))
    for a in off:
        for b in off:
            for c_ in off:
                ins = np.zeros((n, n, n), dtype=bool)
                for c, r in zip(centres, radii):
                    dx = (X - c[0] + 0.5) % 1.0 - 0.5 + a
                    dy = (Y - c[1] + 0.5) % 1.0 - 0.5 + b
                    dz = (Z - c[2] + 0.5) % 1.0 - 0.5 + c_
                    ins |= dx * dx + dy * dy + dz * dz < r * r
                vf += ins
    return vf / nsub ** 3


@pytest.mark.parametrize("nsub", [8, 16])
def test_octree_vf_equals_uniform_leaf_sampling(nsub):
    """The recursive refinement of cut cells (4 levels for nsub = 16, the
    falling_drop setting amr_ref_lvl = 4) gives the same VF as sampling all
    nsub^3 leaf midpoints of every cell, including overlapping spheres and
    periodic images."""
    from nga2_amd.config import Config, uniform

    n = 12
    x = uniform(n)
    cfg = Config(x, x, x, no=1)
    centres = np.array([[0.3, 0.4, 0.5], [0.45, 0.5, 0.5], [0.95, 0.05, 0.5], [0.7, 0.8, 0.1]])
    radii = np.array([0.2, 0.15, 0.12, 0.09])
    got = inputs.sphere_vf(cfg, centres, radii, nsub=nsub)
    ref = _uniform_vf_reference(n, centres, radii, nsub)
    assert np.max(np.abs(got - ref)) <= 0.5 / nsub ** 3
    # and the VF sums to the spheres' volume to the resolution of the leaves
    assert abs(got.mean() - ref.mean()) < 1e-12


def _periodic_operator(n, seed):
    """A random symmetric periodic 7-point operator with positive diagonal,
    as (7, n, n, n) [s, k, j, i] in the stencil order (0, +x, -x, +y, -y, +z, -z)."""
    rng = np.random.default_rng(seed)
    a = [-rng.uniform(0.001, 1.0, (n, n, n)) for _ in range(3)]  # coupling of a cell to its +d neighbour
    o = np.zeros((7, n, n, n))
    for d, ax in enumerate((2, 1, 0)):
        o[1 + 2 * d] = a[d]
        o[2 + 2 * d] = np.roll(a[d], 1, axis=ax)
    o[0] = -o[1:].sum(axis=0) + 0.1
    return o


def test_crop_apply_is_the_principal_submatrix():
    # crop_apply equals the full periodic stencil on every crop cell whose
    # neighbours are all inside the crop, and drops the outside couplings
    n, m, off = 12, 7, (9, 2, 5)
    o = _periodic_operator(n, 3)
    x = np.random.default_rng(4).standard_normal((n, n, n))
    full = o[0] * x
    for s, (ax, sh) in enumerate(((2, -1), (2, 1), (1, -1), (1, 1), (0, -1), (0, 1))):
        full += o[1 + s] * np.roll(x, sh, axis=ax)
    ix = [(np.arange(m) + q) % n for q in off]
    crop = lambda a: a[np.ix_(ix[2], ix[1], ix[0])]  # noqa: E731
    oc = torch.from_numpy(np.stack([crop(o[s]) for s in range(7)]))
    yc = torch.from_numpy(crop(x))
    got = learned.crop_apply(oc, yc).numpy()
    assert np.allclose(got[1:-1, 1:-1, 1:-1], crop(full)[1:-1, 1:-1, 1:-1], rtol=1e-13, atol=1e-13)
    # a field that vanishes inside the crop sees no coupling from outside
    only_out = x.copy()
    only_out[np.ix_(ix[2], ix[1], ix[0])] = 0.0
    assert np.abs(learned.crop_apply(oc, torch.from_numpy(crop(only_out))).numpy()).max() == 0.0


def test_crop_loss_gradient_matches_autograd():
    # the training gradient of the window residual loss against torch
    # autograd (a sign, scale or transpose slip fails it)
    m = 10
    o = torch.from_numpy(_periodic_operator(m, 8)).clone()
    for s, (ax, sl) in enumerate(((2, -1), (2, 0), (1, -1), (1, 0), (0, -1), (0, 0))):
        idx = [slice(None)] * 3
        idx[ax] = sl
        o[1 + s][tuple(idx)] = 0.0  # the window's outside couplings are absent
    g = torch.Generator().manual_seed(1)
    w = torch.randn((m, m, m), generator=g, dtype=torch.float64)
    y = torch.randn((m, m, m), generator=g, dtype=torch.float64, requires_grad=True)
    mask = torch.zeros((m, m, m), dtype=torch.float64)
    mask[2:-2, 2:-2, 2:-2] = 1.0
    kappa, c = 0.7, 3.1
    loss, grad = learned.crop_loss_grad(w, o, c, y.detach(), kappa, mask)
    res = mask * (w - c * kappa * learned.crop_apply(o, y))
    L = (res * res).sum() / mask.sum()
    L.backward()
    Lv = float(L.detach())
    assert abs(loss - Lv) <= 1e-12 * abs(Lv)
    assert torch.allclose(grad, y.grad, rtol=1e-10, atol=1e-14)


def test_crop_loss_is_the_guess_residual():
    # with e = s kappa y on the full periodic grid (a window that is the whole
    # grid, no border, wrap couplings included through a periodic operator
    # built as the window's own), the loss is ||b - A e||^2 / ||b||^2 * N
    n = 8
    o = torch.from_numpy(_periodic_operator(n, 5))
    g = torch.Generator().manual_seed(2)
    b = torch.randn((n, n, n), generator=g, dtype=torch.float64)
    y = torch.randn((n, n, n), generator=g, dtype=torch.float64)
    dinv = 1.0 / o[0]
    s = float(torch.sqrt(((dinv * b) ** 2).mean()))
    t = float(torch.sqrt((b * b).mean()))
    kappa = 0.3
    e = s * kappa * y
    mask = torch.ones((n, n, n), dtype=torch.float64)
    loss, _ = learned.crop_loss_grad(b / t, o, s / t, y, kappa, mask)
    r = b - learned.crop_apply(o, e)
    assert abs(loss - float((r * r).sum() / (b * b).sum())) <= 1e-12
