"""CPU checks of the drifting-spheres inputs: the torch restatements of the
sphere VF and of the face densities give the numpy generators' values bit for
bit (run with torch on the CPU here; the GPU uses the same code)."""
import numpy as np
import torch

from nga2_amd import drift, inputs
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
