"""GPU test of the drifting-spheres two-phase sequence (64^3, 40 spheres):
per step the VF, the face densities and the operator are built on the device
(ngap_proj_laplacian) and solved three ways (FP64 from zero, FP64 from the
previous solution, mixed from the learned temporal prediction). Every solve
meets the true-residual tolerance, and the solutions agree to 1e-8. The
device-assembled operator equals the host assembly bit for bit."""
import numpy as np
import pytest
import torch

from nga2_amd import drift, incomp, inputs

pytestmark = pytest.mark.gpu


def test_device_assembly_equals_host_assembly():
    print("\n[drift assembly 32^3]", flush=True)
    seq = drift.DriftingSpheres(n=32, nsph=10, rmin=0.05, rmax=0.15)
    vf = seq.operator(3).cpu().numpy()
    cfg = seq.cfg
    ref_vf = inputs.sphere_vf(cfg, seq.positions(3), seq.radii)
    assert np.array_equal(vf, ref_vf)
    ref = incomp.laplacian(seq.met, rho_faces=inputs.face_densities(cfg, ref_vf, 1000.0, 1.0))
    got = seq.opr.cpu().numpy().reshape((7,) + cfg.shape, order="F")
    assert np.array_equal(got, ref)
    seq.proj.destroy()


def test_drift_sequence_solves():
    print("\n[drift sequence 64^3]", flush=True)
    res = drift.run(n=64, steps=22, fit_steps=16, m=4, nsph=40, drift_cells=0.5, log=lambda m: None)
    print("  " + str({k: res[k] for k in ("zero_fp64", "prev_fp64", "learned_mixed")}), flush=True)
    for k in ("zero_fp64", "prev_fp64", "learned_mixed"):
        assert res[k]["max_true_rerr"] <= 1e-10, k
    assert res["max_diff_mixed"] <= 1e-8
