"""GPU tests of the drifting-spheres two-phase sequence (configuration 5):
per step the VF, the face densities, the operator and a fresh right-hand side
are built on the device, then solved plain FP64, mixed, and mixed with the
U-Net guess and flexible-PCG stage. Every solve meets the true-residual
tolerance, the solutions agree to 1e-8, and one step agrees with the FP64 CPU
oracle. The device-assembled operator equals the host assembly bit for bit,
and the device right-hand side equals the host recipe on the same faces."""
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


def test_step_rhs_equals_host_recipe():
    # the device RHS of a step against incomp.get_div on the same faces
    print("\n[drift rhs 24^3]", flush=True)
    seq = drift.DriftingSpheres(n=24, nsph=5, rmin=0.05, rmax=0.15)
    cfg = seq.cfg
    got = seq.rhs_step(7).cpu().numpy().reshape(cfg.shape, order="F")
    U, V, W = (f.cpu().numpy().reshape(cfg.shape, order="F") for f in seq._faces)
    for F in (U, V, W):  # the recipe's per-component mean removal and ghost sync
        assert abs(F[cfg.interior].mean()) < 1e-14
        G = F.copy()
        cfg.sync(G)
        assert np.array_equal(G, F)
    ref = -cfg.vol * incomp.get_div(seq.met, U, V, W)
    ref = ref[cfg.interior] - ref[cfg.interior].mean()
    assert np.allclose(got[cfg.interior], ref, rtol=0, atol=1e-15 * np.abs(ref).max())
    assert abs(got[cfg.interior].sum()) < 1e-12 * np.abs(ref).sum()
    # a different step draws different faces
    other = seq.rhs_step(8).cpu().numpy().reshape(cfg.shape, order="F")
    assert np.linalg.norm(other[cfg.interior] - got[cfg.interior]) > 0.5 * np.linalg.norm(ref)
    seq.proj.destroy()


def test_drift_sequence_small_against_oracle():
    # 128^3, 40 spheres: the three modes and one step against the FP64 CPU
    # oracle (the U-Net needs nx % 128 == 0, ny and nz % 32 == 0)
    from oracle.orc import Oracle, interior

    print("\n[drift sequence 128^3]", flush=True)
    errs = {}

    def check(step, seq, x64, xmx):
        c = seq.cfg
        opr_h = seq.opr.cpu().numpy().reshape((7,) + c.shape, order="F")
        b = interior(seq.rhs.cpu().numpy().reshape(c.shape, order="F"), c)
        o = Oracle.from_cfg(c, opr_h)
        xr, _ = o.solve(b, rcvg=1e-12, maxit=500)
        o.close()
        xr = xr - xr.mean()
        errs[step] = [float(np.linalg.norm(x.cpu().numpy() - xr) / np.linalg.norm(xr)) for x in (x64, xmx)]

    res = drift.run(n=128, steps=6, train_steps=3, nsph=40, crop=128, crops_per_step=2, epochs=3,
                    check=check, check_steps=(5,), log=lambda m: None)
    print("  " + str({k: res[k] for k in ("zero_fp64", "mixed", "unet_mixed")}), flush=True)
    print(f"  oracle errors {errs}", flush=True)
    for k in ("zero_fp64", "mixed", "unet_mixed"):
        assert res[k]["max_true_rerr"] <= 1e-10, k
    assert res["max_diff_mixed"] <= 1e-8 and res["max_diff_unet"] <= 1e-8
    assert res["unet_mixed"]["stage_calls"] > 0
    assert max(errs[5]) <= 1e-8, errs


def test_drift_sequence_512():
    # configuration 5 at its size: 512^3, 200 spheres r 0.02-0.08, density
    # ratio 1000; two training steps for the U-Net, two evaluated steps
    print("\n[drift sequence 512^3, 200 spheres]", flush=True)
    res = drift.run(n=512, steps=4, train_steps=2, crop=128, crops_per_step=3, epochs=2,
                    log=lambda m: print("  " + m, flush=True))
    print("  " + str({k: res[k] for k in ("zero_fp64", "mixed", "unet_mixed")}), flush=True)
    for k in ("zero_fp64", "mixed", "unet_mixed"):
        assert res[k]["max_true_rerr"] <= 1e-10, k
    assert res["max_diff_mixed"] <= 1e-8, res["max_diff_mixed"]
    assert res["max_diff_unet"] <= 1e-8, res["max_diff_unet"]
    assert res["unet_mixed"]["stage_calls"] > 0
