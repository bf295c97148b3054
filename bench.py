"""Pressure-solve benchmark: solves/s to 1e-10 relative residual at 256^3.

One step = one pressure solve of the 256^3 channel workload (x/z periodic,
y walls, tanh-stretched y, stretch 2; seeded spectral RHS) from a zero initial
guess, with opr, rhs and sol already resident in HBM (device entry points of
libngapsolve.so). The headline mode is the FP64 flexible PCG with the FP32
V-cycle; the plain FP64 method is timed beside it. Prints one JSON line.

Outside the timed region the line also carries:
- the 256^3 two-phase spheres solve (density ratio 1000), setup and solve,
  mixed against plain FP64;
- the learned-guess sequence (128^3 periodic, 200 correlated right-hand
  sides): train on steps 0-149, then iterations and wall time on steps
  150-199 against the previous-step guess (nga2_amd.learned);
- the drifting-spheres two-phase sequence (512^3, 200 spheres, 100 steps):
  operator and a new right-hand side built on the device every step, setup +
  solve per step of plain FP64, mixed, and mixed with the U-Net guess and
  flexible-PCG stage (nga2_amd.drift), one step against the FP64 CPU oracle;
- fourier3d, the FFT direct solve, at 256^3 triply periodic (nga2_amd.fourier),
  against the PCG solve of the same system;
- gmres_pfmg on a nonsymmetric implicit-solve system on the 256^3 channel
  grid, FP64 and with the FP32 V-cycle;
- hypre_uns's pcg_amg (algebraic multigrid over the CSR matrix of the
  participating cells) on 128^3 immersed solids, against PCG+PFMG on the same
  system.

  python bench.py [--steps K] [--warmup W] [--n 256] [--workload channel_spectral]
                  [--no-cpu] [--cpu-threads T] [--roofline-reps R]
                  [--no-two-phase] [--no-sequence] [--seq-n 128] [--seq-m 8] [--seq-epochs 6]
                  [--no-drift] [--drift-n 512] [--drift-steps 100] [--drift-train 12] [--no-fourier] [--no-implicit]
                  [--no-uns] [--uns-n 128]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "pressure solves/sec to 1e-10 rel. residual at 256^3; HBM GB/s vs roofline"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--workload", default="channel_spectral")
    ap.add_argument("--rcvg", type=float, default=1e-10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-1core", action="store_true", help="skip the one-core CPU oracle solve")
    ap.add_argument("--no-host-path", action="store_true", help="skip the host-array (ngap_solve) leg")
    ap.add_argument("--roofline-reps", type=int, default=20)
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--no-two-phase", action="store_true", help="skip the 256^3 two-phase spheres leg")
    ap.add_argument("--no-two-phase-unet", action="store_true", help="two-phase leg without the U-Net stage")
    ap.add_argument("--no-device-step", action="store_true", help="skip the device-resident projection step leg")
    ap.add_argument("--no-config1", action="store_true", help="skip the 64^3 configuration-1 leg")
    ap.add_argument("--no-sequence", action="store_true", help="skip the learned-guess sequence leg")
    ap.add_argument("--no-drift", action="store_true", help="skip the drifting-spheres two-phase sequence leg")
    ap.add_argument("--no-fourier", action="store_true", help="skip the fourier3d (FFT direct solve) leg")
    ap.add_argument("--no-implicit", action="store_true", help="skip the gmres_pfmg implicit-solve leg")
    ap.add_argument("--no-uns", action="store_true", help="skip the hypre_uns (pcg_amg) masked-solids leg")
    ap.add_argument("--uns-n", type=int, default=128)
    ap.add_argument("--drift-n", type=int, default=512)
    ap.add_argument("--drift-steps", type=int, default=100)
    ap.add_argument("--drift-train", type=int, default=12, help="drift steps whose solves train the U-Net")
    ap.add_argument("--no-drift-unet", action="store_true", help="drift leg without the U-Net path")
    ap.add_argument("--seq-n", type=int, default=128)
    ap.add_argument("--seq-m", type=int, default=8)
    ap.add_argument("--seq-epochs", type=int, default=6)
    ap.add_argument("--kernel-only", choices=sorted(KERNEL_IDS), default=None,
                    help="time one fine-level kernel (for rocprofv3 summaries and counters) and exit")
    return ap.parse_args()


def cpu_model():
    """The host CPU's model name (lscpu, else /proc/cpuinfo) and socket/core counts."""
    import subprocess

    out = {}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                out[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    if "Model name" not in out:
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        out["Model name"] = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
    out["os.cpu_count"] = os.cpu_count()
    return out


def profiled_traffic(name):
    """HBM bytes per launch recorded in a committed rocprofv3 summary
    (FETCH_SIZE and WRITE_SIZE counters scaled by a copy of known size)."""
    try:
        with open(os.path.join(ROOT, "profiles", name)) as fh:
            for line in fh:
                if line.startswith("traffic_bytes:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


KERNEL_IDS = {"rbgs": (1, 1), "spmv": (0, 0), "restrict": (2, 1), "prolong": (3, 1),  # (id, precision)
              "sweep2": (6, 1), "sweep2_zero": (7, 1), "rbgs64": (1, 0)}


def kernel_only(args):
    """Set up the workload and launch one fine-level kernel `--roofline-reps`
    times (HIP events on the solver's stream), plus a plain device copy of one
    FP32 fine-level array as the counter calibration kernel."""
    import torch

    from nga2_amd import inputs
    from nga2_amd._lib import MIXED
    from nga2_amd.linsol import DeviceSolve

    dev = torch.device("cuda:0")
    torch.cuda.set_device(0)
    P = inputs.WORKLOADS[args.workload](n=args.n)
    d_opr = torch.from_numpy(np.ascontiguousarray(P.opr.ravel(order="F"))).to(dev)
    ds = DeviceSolve(P.cfg)
    ds.setup(d_opr, maxit=500, rcvg=args.rcvg, precision=MIXED)
    kid, prec = KERNEL_IDS[args.kernel_only]
    ms, nbytes = ds.hooks().time_kernel(kid, prec=prec, reps=args.roofline_reps)
    n = P.cfg.nx * P.cfg.ny * P.cfg.nz
    src = torch.ones(n, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    for _ in range(10):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"kernel": args.kernel_only, "avg_ms": ms, "algorithmic_bytes": nbytes,
                      "achieved_GBs": nbytes / (ms * 1e-3) / 1e9, "copy_calibration_bytes": 2 * 4 * n,
                      "reps": args.roofline_reps}), flush=True)
    ds.destroy()
    return 0


def host_path_leg(args, P, device_ms):
    """The solve through the host entry point (ngap_solve), as the Fortran
    drop-in calls it: rhs and the guess in pageable host arrays, sol copied
    back (hypre_str_solve, src/solver/hypre_str_class.f90:336-377). Same
    workload and precision as the headline; 1 warm-up, median of 5."""
    from nga2_amd._lib import MIXED, PCG_PFMG
    from nga2_amd.linsol import NgapStr

    cfg = P.cfg
    ps = NgapStr(cfg, "Pressure", PCG_PFMG, 7)
    ps.stc[:] = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]])
    ps.maxit, ps.rcvg, ps.precision = 500, args.rcvg, MIXED
    ps.opr[:] = P.opr
    ps.init()
    ps.setup()
    walls, copies, its = [], [], []
    for rep in range(6):
        ps.rhs[:] = P.rhs
        ps.sol[:] = 0.0
        t0 = time.perf_counter()
        ps.solve()
        w = (time.perf_counter() - t0) * 1e3
        st = ps.stats()
        if rep:
            walls.append(w)
            copies.append(st["copy_ms"])
            its.append(ps.it)
    ps.destroy()
    wall = float(np.median(walls))
    gb = 3 * cfg.nxo * cfg.nyo * cfg.nzo * 8 / 1e9
    res = {"value": round(1e3 / wall, 4), "unit": "solves/s", "ms_per_solve": round(wall, 3),
           "copy_ms": round(float(np.median(copies)), 3), "iters": its[0],
           "ratio_to_device_path": round(wall / device_ms, 3),
           "host_bytes_gb": round(gb, 3),
           "note": "ngap_solve with pageable host arrays (rhs, guess in; sol out, ghosted), median of 5; "
                   "copies staged through pinned chunks on 4 host threads, an all-zero guess chunk cleared "
                   "on the device instead of copied"}
    log(f"[bench] host path: {res['value']:.2f} solves/s, copy {res['copy_ms']:.1f} ms, "
        f"{res['ratio_to_device_path']:.2f}x the device path")
    return res


def two_phase_leg(args, dev, solves=3):
    """256^3 two-phase spheres (density ratio 1000): setup + solve, mixed vs FP64."""
    import torch

    from nga2_amd import inputs
    from nga2_amd._lib import FP64, MIXED
    from nga2_amd.linsol import DeviceSolve

    t0 = time.time()
    P = inputs.spheres_two_phase_config3(n=args.n)
    log(f"[bench] generated {P.name} in {time.time() - t0:.1f}s")
    cfg = P.cfg
    d_opr = torch.from_numpy(np.ascontiguousarray(P.opr.ravel(order="F"))).to(dev)
    d_rhs = torch.from_numpy(np.ascontiguousarray(P.rhs.ravel(order="F"))).to(dev)
    d_sol = torch.zeros_like(d_rhs)
    res = {"workload": f"{P.name}: {cfg.nx}^3 periodic, 50 spheres r 0.05-0.15, VF by four-level refinement, "
                       f"density ratio 1000, arithmetic face densities, rcvg {args.rcvg:g}, zero guess"}
    for name, prec in (("mixed", MIXED), ("fp64", FP64)):
        ds = DeviceSolve(cfg)
        ds.setup(d_opr, maxit=500, rcvg=args.rcvg, precision=prec)
        d_sol.zero_()
        ds.solve(d_rhs, d_sol)  # warm-up
        torch.cuda.synchronize()
        setup_ms = []
        solve_ms = []
        its = []
        for _ in range(solves):
            t = time.perf_counter()
            ds.setup(d_opr, maxit=500, rcvg=args.rcvg, precision=prec)
            torch.cuda.synchronize()
            setup_ms.append((time.perf_counter() - t) * 1e3)
            d_sol.zero_()
            t = time.perf_counter()
            it, _ = ds.solve(d_rhs, d_sol)
            torch.cuda.synchronize()
            solve_ms.append((time.perf_counter() - t) * 1e3)
            its.append(it)
        st = ds.stats()
        res[name] = {"iters": its[0], "solve_ms": round(float(np.median(solve_ms)), 3),
                     "setup_ms": round(float(np.median(setup_ms)), 3), "true_rerr": st["true_rerr"],
                     "fallback": st["fallback"]}
        ds.destroy()
        log(f"[bench] two-phase {name}: it={its[0]} solve {res[name]['solve_ms']} ms setup {res[name]['setup_ms']} ms")
    res["speedup_solve_mixed_vs_fp64"] = round(res["fp64"]["solve_ms"] / res["mixed"]["solve_ms"], 3)
    res["speedup_setup_solve_mixed_vs_fp64"] = round(
        (res["fp64"]["solve_ms"] + res["fp64"]["setup_ms"]) / (res["mixed"]["solve_ms"] + res["mixed"]["setup_ms"]), 3)
    if not args.no_two_phase_unet:
        res["unet_mixed"] = two_phase_unet(args, P, d_opr, d_rhs, solves)
        u = res["unet_mixed"]
        res["north_star_speedup"] = round(
            (res["fp64"]["solve_ms"] + res["fp64"]["setup_ms"]) / (u["solve_ms"] + u["setup_ms"]), 3)
        res["unet_iteration_saving"] = res["mixed"]["iters"] - u["iters"]
    return res


def two_phase_unet(args, P, d_opr, d_rhs, solves):
    """Mixed precision with the learned stage on the two-phase system: the
    U-Net (input D^-1 r) is trained on crops of this operator's solves of
    other configuration-1 right-hand sides (seeds 7001-7004), then gives the
    initial guess and the first flexible-PCG stage of the timed solves.
    setup_ms counts the mixed setup and the input-scale update; the training
    is reported apart (it is paid once per operator, not per solve)."""
    import torch

    from nga2_amd import _lib, incomp, learned
    from nga2_amd._lib import MIXED
    from nga2_amd.linsol import DeviceSolve
    from nga2_amd.proj import DeviceProjection
    from nga2_amd.unet import UNet

    cfg = P.cfg
    ds = DeviceSolve(cfg)
    ds.setup(d_opr, maxit=500, rcvg=args.rcvg, precision=MIXED)
    proj = DeviceProjection(incomp.Metrics(cfg))
    state, info = learned.train_for_operator(ds, proj, d_opr, cfg, seeds=range(7001, 7005),
                                             log=lambda m: log("[bench] two-phase " + m))
    proj.destroy()
    net = UNet((cfg.nx, cfg.ny, cfg.nz), seed=0)
    net.load_state(state)
    dinv = torch.zeros(cfg.nx * cfg.ny * cfg.nz, dtype=torch.float64, device=d_opr.device)
    d_sol = torch.zeros_like(d_rhs)
    setup_ms, solve_ms, its, st = [], [], [], None
    for rep in range(solves + 1):
        torch.cuda.synchronize()
        t = time.perf_counter()
        ds.setup(d_opr, maxit=500, rcvg=args.rcvg, precision=MIXED)
        learned.diag_inverse(d_opr, cfg, dinv)
        net.set_input_scale(dinv)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ds.set_stage(net.stage(), flags=_lib.STAGE_GUESS | _lib.STAGE_PRECOND, precond_iters=1)
        d_sol.zero_()
        it, _ = ds.solve(d_rhs, d_sol)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        st = ds.stats()
        ds.set_stage(None)
        if rep:  # the first is a warm-up
            setup_ms.append((t1 - t) * 1e3)
            solve_ms.append((t2 - t1) * 1e3)
            its.append(it)
    # the network's cost per call, and the guess it gives on its own
    res_buf = torch.zeros_like(dinv)
    bn = float(torch.linalg.vector_norm(learned.interior_flat(d_rhs, cfg)))
    e = net.apply(learned.interior_flat(d_rhs, cfg))
    r_guess = ds.residual(d_rhs, learned.to_ghosted(e, cfg), res_buf) / bn
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(3):
        net.apply(res_buf)
    torch.cuda.synchronize()
    fwd_ms = (time.perf_counter() - t) / 3 * 1e3
    net.close()
    ds.destroy()
    res = {"iters": its[0], "solve_ms": round(float(np.median(solve_ms)), 3),
           "setup_ms": round(float(np.median(setup_ms)), 3), "true_rerr": st["true_rerr"],
           "stage_calls": st["stage_calls"], "guess_rejected": int(bool(st["fallback"] & 2)),
           "guess_rel_residual": r_guess, "stage_ms": round(fwd_ms, 3),
           "training": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in info.items()}}
    log(f"[bench] two-phase mixed + U-Net: it={res['iters']} solve {res['solve_ms']} ms setup {res['setup_ms']} ms "
        f"(guess residual {r_guess:.3f}, stage {fwd_ms:.2f} ms)")
    return res


def device_step_leg(args, dev, reps=3):
    """One pressure-projection step of the two-phase workload (configuration 3's
    spheres at n^3) with every array in HBM, in the order of
    examples/falling_drop/src/simulation.f90:298-312 and tpns_class.f90:1565-1716:
    face densities and the operator (update_laplacian), setup, div of the
    velocity, rhs = -vol div / dt, solve from zero, shift_p, the pressure
    gradient, and the velocity correction U -= dt Px / rho_U with P += sol.
    The velocity is a configuration-1 Gaussian face field. Reports each part
    and the whole, median of `reps` after a warm-up, and the divergence left
    after the correction relative to before (the projection's own check)."""
    import torch

    from nga2_amd import drift, incomp, inputs, learned
    from nga2_amd._lib import MIXED
    from nga2_amd.config import Config, uniform
    from nga2_amd.linsol import DeviceSolve
    from nga2_amd.proj import DeviceProjection

    n = args.n
    x = uniform(n)
    cfg = Config(x, x, x, no=1, xper=True, yper=True, zper=True, name="two_phase_step")
    met = incomp.Metrics(cfg)
    proj = DeviceProjection(met)
    centres, radii = inputs.sphere_set(50, 0.05, 0.15, 3003)
    vf = drift.sphere_vf_device(n, centres, radii, device=dev)
    No = cfg.nxo * cfg.nyo * cfg.nzo
    z = lambda: torch.zeros(No, dtype=torch.float64, device=dev)  # noqa: E731
    vel = [z() for _ in range(3)]
    opr = torch.zeros(7 * No, dtype=torch.float64, device=dev)
    div, rhs, sol, P = z(), z(), z(), z()
    grads = [z() for _ in range(3)]
    ds = DeviceSolve(cfg)
    dt = 1.0
    parts = {k: [] for k in ("operator", "setup", "rhs", "solve", "shift_pgrad_correct", "step")}
    its, div_ratio = [], []
    for rep in range(reps + 1):
        learned.gaussian_rhs_device(proj, cfg, 8001 + rep, rhs, faces=vel, div=div)  # vel <- Gaussian faces
        torch.cuda.synchronize()
        t = [time.perf_counter()]
        rf = drift.face_densities_device(cfg, vf, 1000.0, 1.0)
        proj.laplacian(opr, rf)
        torch.cuda.synchronize()
        t.append(time.perf_counter())
        ds.setup(opr, maxit=500, rcvg=args.rcvg, precision=MIXED)
        torch.cuda.synchronize()
        t.append(time.perf_counter())
        proj.div(*vel, div)
        proj.rhs(div, 1.0, dt, rhs)
        torch.cuda.synchronize()
        t.append(time.perf_counter())
        sol.zero_()
        it, _ = ds.solve(rhs, sol)
        torch.cuda.synchronize()
        t.append(time.perf_counter())
        proj.shift_p(sol)
        proj.sync(sol)
        proj.pgrad(sol, *grads)
        proj.correct(dt, 1.0, grads, vel, rho_faces=rf, P=P, sol=sol)
        torch.cuda.synchronize()
        t.append(time.perf_counter())
        if rep:
            for k, (a, b) in zip(("operator", "setup", "rhs", "solve", "shift_pgrad_correct"), zip(t, t[1:])):
                parts[k].append((b - a) * 1e3)
            parts["step"].append((t[-1] - t[0]) * 1e3)
            its.append(it)
            d0 = float(torch.linalg.vector_norm(learned.interior_flat(div, cfg)))
            for v in vel:
                proj.sync(v)
            proj.div(*vel, div)
            div_ratio.append(float(torch.linalg.vector_norm(learned.interior_flat(div, cfg))) / d0)
    ds.destroy()
    proj.destroy()
    res = {k: round(float(np.median(v)), 3) for k, v in parts.items()}
    res = {f"{k}_ms": v for k, v in res.items()}
    res.update({"iters": its[0], "div_after_over_before": max(div_ratio),
                "workload": f"{n}^3 periodic, 50 spheres r 0.05-0.15 (VF 4^3 sub-cells), density ratio 1000, "
                            "Gaussian face velocity; mixed precision, rcvg "
                            f"{args.rcvg:g}; U, V, W, P, opr, rhs, sol resident in HBM throughout"})
    log(f"[bench] device time step: {res['step_ms']} ms ({res})")
    return res


def config1_leg(args, dev, solves=10):
    """Configuration 1 at its size: the 64^3 triply periodic Gaussian case
    (examples/hit/src/simulation.f90:108-116), FP64 and mixed on the GPU (device
    entry points, median of `solves`) against the FP64 CPU oracle (median of 5
    after a warm-up, the box's threads) and the FFT closed form of the uniform
    7-point operator (src/solver/fourier3d_class.f90:1103-1196)."""
    import torch

    from nga2_amd import inputs
    from nga2_amd._lib import FP64, MIXED
    from nga2_amd.linsol import DeviceSolve
    from oracle.orc import Oracle, interior, num_threads

    P = inputs.periodic_gaussian(n=64)
    cfg = P.cfg
    n = cfg.nx
    b = interior(P.rhs, cfg)
    # the closed form: -vol lap x = b with eigenvalues vol * sum_d (2 - 2 cos(2 pi k_d / n)) / h^2
    k = np.arange(n)
    lam1 = (1.0 / n) * (2.0 - 2.0 * np.cos(2.0 * np.pi * k / n))
    lam = lam1[:, None, None] + lam1[None, :, None] + lam1[None, None, :]
    lam[0, 0, 0] = 1.0
    X = np.fft.fftn(b.reshape(n, n, n)) / lam
    X[0, 0, 0] = 0.0
    xc = np.fft.ifftn(X).real.ravel()

    def err(x, ref):
        x = x - x.mean()
        r = ref - ref.mean()
        return float(np.linalg.norm(x - r) / np.linalg.norm(r))

    res = {"workload": "periodic_gaussian_64: 64^3 triply periodic, uniform, Gaussian faces seed 1001, "
                       f"rcvg {args.rcvg:g}, zero guess"}
    o = Oracle.from_cfg(cfg, P.opr)
    times = []
    for rep in range(6):
        t = time.perf_counter()
        xr, st_r = o.solve(b, rcvg=args.rcvg, maxit=500)
        if rep:
            times.append(time.perf_counter() - t)
    o.close()
    res["cpu_oracle"] = {"iters": st_r["it"], "solve_ms": round(1e3 * float(np.median(times)), 3),
                         "threads": num_threads(), "err_vs_closed_form": err(xr, xc)}
    d_opr = torch.from_numpy(np.ascontiguousarray(P.opr.ravel(order="F"))).to(dev)
    d_rhs = torch.from_numpy(np.ascontiguousarray(P.rhs.ravel(order="F"))).to(dev)
    d_sol = torch.zeros_like(d_rhs)
    for name, prec in (("fp64", FP64), ("mixed", MIXED)):
        ds = DeviceSolve(cfg)
        ds.setup(d_opr, maxit=500, rcvg=args.rcvg, precision=prec)
        ms = []
        for rep in range(solves + 1):
            d_sol.zero_()
            torch.cuda.synchronize()
            t = time.perf_counter()
            it, _ = ds.solve(d_rhs, d_sol)
            torch.cuda.synchronize()
            if rep:
                ms.append((time.perf_counter() - t) * 1e3)
        st = ds.stats()
        ds.destroy()
        x = interior(d_sol.cpu().numpy().reshape(cfg.shape, order="F"), cfg)
        res[name] = {"iters": it, "solve_ms": round(float(np.median(ms)), 3), "true_rerr": st["true_rerr"],
                     "err_vs_fp64_oracle": err(x, xr), "err_vs_closed_form": err(x, xc)}
    res["speedup_fp64_gpu_vs_cpu"] = round(res["cpu_oracle"]["solve_ms"] / res["fp64"]["solve_ms"], 2)
    log(f"[bench] configuration 1 (64^3): {res}")
    return res


def fourier_leg(args, dev, solves=10):
    """fourier3d on the GPU at n^3 periodic uniform (configuration 1's recipe at
    the metric's size): setup and solve time, the true residual, and the PCG
    solve of the same system (mixed) for comparison."""
    import torch

    from nga2_amd import fourier, inputs
    from nga2_amd._lib import MIXED
    from nga2_amd.linsol import DeviceSolve

    t0 = time.time()
    P = inputs.periodic_gaussian(n=args.n)
    log(f"[bench] generated {P.name} in {time.time() - t0:.1f}s")
    cfg = P.cfg
    d_opr = torch.from_numpy(np.ascontiguousarray(P.opr.ravel(order="F"))).to(dev)
    d_rhs = torch.from_numpy(np.ascontiguousarray(P.rhs.ravel(order="F"))).to(dev)
    d_x = torch.zeros_like(d_rhs)
    ds = fourier.DeviceFourier(cfg)
    torch.cuda.synchronize()
    t = time.perf_counter()
    ds.setup(d_opr)
    setup_ms = (time.perf_counter() - t) * 1e3
    ds.solve(d_rhs, d_x)  # warm-up
    wall, dev_ms = [], []
    for _ in range(solves):
        torch.cuda.synchronize()
        t = time.perf_counter()
        ds.solve(d_rhs, d_x)
        wall.append((time.perf_counter() - t) * 1e3)
        dev_ms.append(ds.last_ms())
    ds.destroy()
    # true residual with the exact operator, and the PCG solution of the same system
    ps = DeviceSolve(cfg)
    ps.setup(d_opr, maxit=200, rcvg=args.rcvg, precision=MIXED)
    d_r = torch.zeros(cfg.nx * cfg.ny * cfg.nz, dtype=torch.float64, device=dev)
    rnorm = ps.residual(d_rhs, d_x, d_r)
    from nga2_amd import learned

    bnorm = float(torch.linalg.vector_norm(learned.interior_flat(d_rhs, cfg)))
    d_y = torch.zeros_like(d_rhs)
    torch.cuda.synchronize()
    t = time.perf_counter()
    it, _ = ps.solve(d_rhs, d_y)
    pcg_ms = (time.perf_counter() - t) * 1e3
    x = learned._mean_free(learned.interior_flat(d_x, cfg))
    y = learned._mean_free(learned.interior_flat(d_y, cfg))
    diff = float(torch.linalg.vector_norm(x - y) / torch.linalg.vector_norm(x))
    ps.destroy()
    N = cfg.nx * cfg.ny * cfg.nz
    res = {"workload": f"{P.name}: {cfg.nx}^3 triply periodic, uniform, Gaussian faces seed 1001, direct FFT solve "
                       "(rocFFT D2Z/Z2D FP64 + HIP divide, copy and ghost fill), inputs resident in HBM",
           "setup_ms": round(setup_ms, 3), "solve_ms": round(float(np.median(wall)), 3),
           "solve_ms_device": round(float(np.median(dev_ms)), 3),
           "solves_per_s": round(1e3 / float(np.median(wall)), 2),
           "true_rerr": rnorm / bnorm, "pcg_mixed_same_system": {"iters": it, "solve_ms": round(pcg_ms, 3),
                                                                 "rel_diff_vs_fft": diff},
           "speedup_vs_pcg_mixed": round(pcg_ms / float(np.median(wall)), 2),
           "cells": N}
    log(f"[bench] fourier3d: solve {res['solve_ms']} ms (device {res['solve_ms_device']} ms), setup {setup_ms:.1f} ms, "
        f"true rerr {res['true_rerr']:.2e}, PCG mixed {pcg_ms:.1f} ms ({it} it), diff {diff:.2e}")
    return res


def implicit_leg(args, dev, solves=3):
    """gmres_pfmg on the nonsymmetric implicit-solve test system at n^3 (the
    channel grid: x/z periodic, y walls, tanh stretch 2; visc 0.02, CFL 1):
    setup and solve, plain FP64 against the FP32 V-cycle."""
    import torch

    from nga2_amd import inputs
    from nga2_amd._lib import FP64, GMRES_PFMG, MIXED
    from nga2_amd.config import Config, tanh_stretched, uniform
    from nga2_amd.linsol import DeviceSolve

    t0 = time.time()
    n = args.n
    cfg = Config(uniform(n), tanh_stretched(n), uniform(n), no=1, yper=False, name="implicit_channel")
    P = inputs.implicit_problem(cfg, visc=0.02)
    log(f"[bench] generated {P.name} in {time.time() - t0:.1f}s")
    d_opr = torch.from_numpy(np.ascontiguousarray(P.opr.ravel(order="F"))).to(dev)
    d_rhs = torch.from_numpy(np.ascontiguousarray(P.rhs.ravel(order="F"))).to(dev)
    d_x = torch.zeros_like(d_rhs)
    res = {"workload": f"{P.name}: {n}^3, x/z periodic, y walls tanh stretch 2, rho 1 + dt/2 (central convection "
                       f"by Gaussian faces |u|max 1 - 0.02 lap), dt = h_min, rcvg {args.rcvg:g}, zero guess; "
                       "GMRES(5) right-preconditioned by the V-cycle of the symmetric part"}
    xs = {}
    for name, prec in (("fp64", FP64), ("mixed", MIXED)):
        ds = DeviceSolve(cfg, method=GMRES_PFMG)
        ds.setup(d_opr, maxit=200, rcvg=args.rcvg, precision=prec)
        d_x.zero_()
        ds.solve(d_rhs, d_x)  # warm-up
        setup_ms, solve_ms = [], []
        for _ in range(solves):
            torch.cuda.synchronize()
            t = time.perf_counter()
            ds.setup(d_opr, maxit=200, rcvg=args.rcvg, precision=prec)
            torch.cuda.synchronize()
            setup_ms.append((time.perf_counter() - t) * 1e3)
            d_x.zero_()
            t = time.perf_counter()
            it, _ = ds.solve(d_rhs, d_x)
            torch.cuda.synchronize()
            solve_ms.append((time.perf_counter() - t) * 1e3)
        st = ds.stats()
        xs[name] = d_x.clone()
        res[name] = {"iters": it, "restarts": st["restarts"], "solve_ms": round(float(np.median(solve_ms)), 3),
                     "setup_ms": round(float(np.median(setup_ms)), 3), "true_rerr": st["true_rerr"]}
        ds.destroy()
        log(f"[bench] implicit gmres {name}: it={it} solve {res[name]['solve_ms']} ms setup {res[name]['setup_ms']} ms")
    res["rel_diff_mixed_vs_fp64"] = float(torch.linalg.vector_norm(xs["mixed"] - xs["fp64"]) /
                                          torch.linalg.vector_norm(xs["fp64"]))
    res["speedup_solve_mixed_vs_fp64"] = round(res["fp64"]["solve_ms"] / res["mixed"]["solve_ms"], 3)
    return res


def uns_leg(args, dev, solves=3):
    """hypre_uns (pcg_amg) on immersed solids: the masked-spheres system at
    uns_n^3 (40 solid spheres, Neumann at their faces, masked cells outside
    the system): host hierarchy + upload, then the device solve from a zero
    guess; the structured PCG+PFMG solve of the same system (masked cells as
    identity rows) beside it."""
    import torch

    from nga2_amd import inputs, uns
    from nga2_amd._lib import FP64, MIXED
    from nga2_amd.linsol import DeviceSolve

    t0 = time.time()
    n = args.uns_n
    P = inputs.masked_spheres(n=n)
    cfg = P.cfg
    log(f"[bench] generated {P.name} in {time.time() - t0:.1f}s ({P.info['solid_cells']} solid cells)")
    d_rhs = torch.from_numpy(np.ascontiguousarray(P.rhs.ravel(order="F"))).to(dev)
    d_x = torch.zeros_like(d_rhs)
    res = {"workload": f"{P.name}: {n}^3 periodic unit cube, {P.info['nsph']} solid spheres r "
                       f"{P.info['rmin']}-{P.info['rmax']} (seed {P.info['seed']}) as masked cells with Neumann faces, "
                       f"N(0,1) rhs on the fluid, rcvg {args.rcvg:g}, zero guess"}
    t = time.perf_counter()
    ds = uns.DeviceUns(cfg, P.opr, method=uns.PCG_AMG, maxit=200, rcvg=args.rcvg)
    setup_ms = (time.perf_counter() - t) * 1e3
    st = ds.stats()
    ds.solve(d_rhs, d_x)  # warm-up
    ms = []
    for _ in range(solves):
        d_x.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        it, rerr, _ = ds.solve(d_rhs, d_x)
        torch.cuda.synchronize()
        ms.append((time.perf_counter() - t) * 1e3)
    st = ds.stats()
    xa = d_x.clone()
    res["pcg_amg"] = {"rows": st["rows"], "nnz": st["nnz"], "levels": st["nlevels"],
                      "op_complexity": round(st["op_complexity"], 3), "setup_ms": round(setup_ms, 1),
                      "hierarchy_host_ms": round(st["hier_ms"], 1), "iters": it, "true_rerr": rerr,
                      "solve_ms": round(float(np.median(ms)), 3), "vcycle_ms": round(st["vcycle_ms"], 3)}
    ds.destroy()
    log(f"[bench] uns pcg_amg: rows {st['rows']}, levels {st['nlevels']}, it={it}, solve {res['pcg_amg']['solve_ms']} "
        f"ms, V-cycle {st['vcycle_ms']:.3f} ms, setup {setup_ms:.0f} ms")
    # the same with the FP32 cycle under FP64 flexible PCG
    dm = uns.DeviceUns(cfg, P.opr, method=uns.PCG_AMG, maxit=200, rcvg=args.rcvg, precision=1)
    dm.solve(d_rhs, d_x)
    ms = []
    for _ in range(solves):
        d_x.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        it, rerr, _ = dm.solve(d_rhs, d_x)
        torch.cuda.synchronize()
        ms.append((time.perf_counter() - t) * 1e3)
    stm = dm.stats()
    # solutions of the singular system agree up to a constant: compare on the
    # participating cells with their means removed
    act = torch.from_numpy(np.ascontiguousarray((np.abs(P.opr[0]) > 1e-14).ravel(order="F"))).to(dev)
    a, b = xa[act], d_x[act]
    diffm = float(torch.linalg.vector_norm((a - a.mean()) - (b - b.mean())) / torch.linalg.vector_norm(a - a.mean()))
    res["pcg_amg_mixed"] = {"iters": it, "true_rerr": rerr, "solve_ms": round(float(np.median(ms)), 3),
                            "vcycle_ms": round(stm["vcycle_ms"], 3), "rel_diff_vs_fp64": diffm}
    dm.destroy()
    log(f"[bench] uns pcg_amg mixed: it={it}, solve {res['pcg_amg_mixed']['solve_ms']} ms, V-cycle "
        f"{stm['vcycle_ms']:.3f} ms, diff {diffm:.1e}")
    d_opr = torch.from_numpy(np.ascontiguousarray(P.opr.ravel(order="F"))).to(dev)
    for name, prec in (("pcg_pfmg_fp64", FP64), ("pcg_pfmg_mixed", MIXED)):
        s2 = DeviceSolve(cfg)
        s2.setup(d_opr, maxit=200, rcvg=args.rcvg, precision=prec)
        d_x.zero_()
        s2.solve(d_rhs, d_x)
        ms = []
        for _ in range(solves):
            d_x.zero_()
            torch.cuda.synchronize()
            t = time.perf_counter()
            it2, _ = s2.solve(d_rhs, d_x)
            torch.cuda.synchronize()
            ms.append((time.perf_counter() - t) * 1e3)
        st2 = s2.stats()
        a, b = xa[act], d_x[act]
        diff = float(torch.linalg.vector_norm((a - a.mean()) - (b - b.mean())) / torch.linalg.vector_norm(b - b.mean()))
        res[name] = {"iters": it2, "solve_ms": round(float(np.median(ms)), 3), "true_rerr": st2["true_rerr"],
                     "rel_diff_vs_pcg_amg": diff}
        s2.destroy()
        log(f"[bench] uns same system, {name}: it={it2} solve {res[name]['solve_ms']} ms, diff {diff:.1e}")
    return res


def main():
    args = parse()
    ws = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or ws > 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus > 1 is not supported (the solve would "
                                   "shard as a 3-D block decomposition with halo exchange, not implemented here)"}))
        return 2
    if args.kernel_only:
        return kernel_only(args)
    import torch

    from nga2_amd import _lib, inputs
    from nga2_amd._lib import FP64, MIXED
    from nga2_amd.linsol import DeviceSolve

    assert torch.cuda.is_available(), "bench needs a GPU"
    dev = torch.device("cuda:0")
    torch.cuda.set_device(0)
    t0 = time.time()
    P = inputs.WORKLOADS[args.workload](n=args.n)
    cfg = P.cfg
    log(f"[bench] generated {P.name} in {time.time() - t0:.1f}s")
    d_opr = torch.from_numpy(np.ascontiguousarray(P.opr.ravel(order="F"))).to(dev)
    d_rhs = torch.from_numpy(np.ascontiguousarray(P.rhs.ravel(order="F"))).to(dev)
    d_sol = torch.zeros_like(d_rhs)
    N = cfg.nx * cfg.ny * cfg.nz

    def run_mode(prec, steps, warmup):
        ds = DeviceSolve(cfg)
        ts = time.perf_counter()
        ds.setup(d_opr, maxit=500, rcvg=args.rcvg, precision=prec)
        torch.cuda.synchronize()
        setup_s = time.perf_counter() - ts
        for _ in range(warmup):
            d_sol.zero_()
            ds.solve(d_rhs, d_sol)
        torch.cuda.synchronize()
        t = time.perf_counter()
        its = []
        for _ in range(steps):
            d_sol.zero_()
            it, _ = ds.solve(d_rhs, d_sol)
            its.append(it)
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        st = ds.stats()
        sol = d_sol.cpu().numpy().reshape(cfg.shape, order="F")
        # a setup again on the same handle (what a time step repeats): median of 3
        again = []
        for _ in range(3):
            ts = time.perf_counter()
            ds.setup(d_opr, maxit=500, rcvg=args.rcvg, precision=prec)
            torch.cuda.synchronize()
            again.append(time.perf_counter() - ts)
        return ds, dict(elapsed=el, setup_s=setup_s, setup_again_s=float(np.median(again)), its=its, stats=st,
                        sol=sol)

    log("[bench] mixed-precision solves")
    ds_mixed, mixed = run_mode(MIXED, args.steps, args.warmup)
    log(f"[bench] mixed: {args.steps / mixed['elapsed']:.2f} solves/s, it={mixed['its'][0]}")
    fp64 = None
    if not args.skip_fp64:
        log("[bench] plain FP64 solves")
        ds_fp64, fp64 = run_mode(FP64, max(1, args.steps // 2), 1)
        log(f"[bench] fp64: {len(fp64['its']) / fp64['elapsed']:.2f} solves/s, it={fp64['its'][0]}")

    host_res = None
    if not args.no_host_path:
        host_res = host_path_leg(args, P, 1e3 * mixed["elapsed"] / args.steps)

    # Roofline of the kernel that runs on the fine level of the headline (mixed) cycle: its
    # post-smoothing pair k_sweep2<float,16,false>, the largest fine-level launch. "achieved"
    # counts the algorithmic bytes of its one pass over the arrays (7 x 4 B per cell);
    # "sweep_equiv_gbs" the bytes of the two single-sweep launches it replaces. The other
    # fine-level kernels, the single sweep k_sweep<float> (timed on the fine level; in the cycle
    # it runs on the levels with rows of more than 128 cells below it, and the whole-row
    # k_sweep_row on the others) and the FP64 sweep and SpMV of the plain method are timed
    # beside it.
    hk = ds_mixed.hooks()
    rb_ms, rb_bytes = hk.time_kernel(_lib.K_RBGS, prec=1, reps=args.roofline_reps)
    r64_ms, r64_bytes = hk.time_kernel(_lib.K_RBGS, prec=0, reps=args.roofline_reps)
    sp_ms, sp_bytes = hk.time_kernel(_lib.K_SPMV, prec=0, reps=args.roofline_reps)
    p2_ms, p2_bytes = hk.time_kernel(_lib.K_SWEEP2, prec=1, reps=args.roofline_reps)
    p0_ms, p0_bytes = hk.time_kernel(_lib.K_SWEEP2_ZERO, prec=1, reps=args.roofline_reps)
    profiled = args.n == 256 and args.workload == "channel_spectral"

    def entry(ms, nb, prof, kernel, sweep_b=None):
        gbs = nb / (ms * 1e-3) / 1e9
        e = {"kernel": kernel, "achieved": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4), "avg_ms": ms,
             "algorithmic_bytes": nb, "traffic": profiled_traffic(prof) if profiled else None,
             "traffic_source": f"profiles/{prof}" if profiled else None}
        if sweep_b is not None:
            e["sweep_equiv_gbs"] = round(sweep_b / (ms * 1e-3) / 1e9, 1)
            e["sweep_equiv_frac"] = round(sweep_b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        return e

    post = entry(p2_ms, p2_bytes, "sweep2_post_fp32_channel256.txt",
                 "k_sweep2<float,16,false>: the fine level's two post-smoothing red/black GS sweeps in one pass",
                 2 * rb_bytes)
    roofline = {"bound": "hbm", "achieved": post["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": post["frac"], "traffic": post["traffic"]}
    roofline.update({k: v for k, v in post.items() if k not in roofline})
    roofline["sweep2_zero_fp32"] = entry(p0_ms, p0_bytes, "sweep2_pre_fp32_channel256.txt",
                                         "k_sweep2<float,16,true>: the fine level's pre-smoothing pair from a zero iterate",
                                         6.0 / 7.0 * rb_bytes + rb_bytes)
    roofline["sweep_fp32"] = entry(rb_ms, rb_bytes, "sweep_fp32_channel256.txt",
                                   "k_sweep<float>: one fused red/black GS sweep (64 x 16 tiles), on the fine level")
    roofline["sweep_fp64"] = entry(r64_ms, r64_bytes, "sweep_fp64_channel256.txt",
                                   "k_sweep<double>: the plain method's sweep (RelaxType 2), fine level")
    roofline["spmv_fp64"] = entry(sp_ms, sp_bytes, "spmv_fp64_channel256.txt",
                                  "k_stencil_rw<double,2,APPLY>: the FP64 SpMV of the outer PCG")

    # CPU baseline + FP64 oracle error
    cpu = None
    err_mixed = err_fp64 = None
    if not args.no_cpu:
        from oracle.orc import Oracle, interior, num_threads, set_threads

        if args.cpu_threads > 0:
            set_threads(args.cpu_threads)
        log("[bench] CPU oracle (FP64 PCG-PFMG, OpenMP): 1 warm-up + median of 5, then one solve on 1 core")
        ts = time.perf_counter()
        o = Oracle.from_cfg(cfg, P.opr)
        osetup = time.perf_counter() - ts
        b = interior(P.rhs, cfg)
        threads = num_threads()
        times = []
        for rep in range(6):
            ts = time.perf_counter()
            xr, st_r = o.solve(b, rcvg=args.rcvg, maxit=500)
            if rep:
                times.append(time.perf_counter() - ts)
        osolve = float(np.median(times))
        one = None
        if not args.no_cpu_1core:
            set_threads(1)
            ts = time.perf_counter()
            o.solve(b, rcvg=args.rcvg, maxit=500)
            one = time.perf_counter() - ts
            set_threads(threads)
        cpu = {"value": round(1.0 / osolve, 4), "unit": "solves/s", "cores": threads, "kind": "port",
               "sample": f"{P.name}, full solve to {args.rcvg:g}: 1 warm-up then the median of 5 "
                         f"(setup {osetup:.1f}s excluded)",
               "cpu_model": cpu_model(), "iters": st_r["it"], "solve_s_median": round(osolve, 3),
               "solve_s_runs": [round(t, 3) for t in times]}
        if one is not None:
            cpu["one_core"] = {"value": round(1.0 / one, 4), "solve_s": round(one, 3), "cores": 1,
                               "sample": "one solve after the threaded runs"}
        xr -= xr.mean()

        def err(sol):
            x = interior(sol, cfg)
            x = x - x.mean()
            return float(np.linalg.norm(x - xr) / np.linalg.norm(xr))

        err_mixed = err(mixed["sol"])
        if fp64:
            err_fp64 = err(fp64["sol"])
        o.close()

    two_phase = None
    if not args.no_two_phase:
        two_phase = two_phase_leg(args, dev)
    step_res = None
    if not args.no_device_step:
        step_res = device_step_leg(args, dev)
    c1_res = None
    if not args.no_cpu and not args.no_config1:
        c1_res = config1_leg(args, dev)
    sequence = None
    if not args.no_sequence:
        from nga2_amd import learned

        log(f"[bench] learned-guess sequence ({args.seq_n}^3, 200 steps, train on 0-149)")
        sequence = learned.run(n=args.seq_n, m=args.seq_m, epochs=args.seq_epochs, log=log)
        ev = sequence["eval"]
        sequence["workload"] = (f"{args.seq_n}^3 periodic, 200 correlated RHS (k<=16, per-mode phase speeds), "
                                f"mixed precision, rcvg {args.rcvg:g}; steps 0-149 train, 150-199 evaluate")
        sequence["speedup_learned_vs_prev"] = round(ev["ms_prev"] / ev["ms_learned"], 3)
        sequence["speedup_extrap_vs_prev"] = round(ev["ms_prev"] / ev["ms_extrap"], 3)

    fft_res = None
    if not args.no_fourier:
        fft_res = fourier_leg(args, dev)
    impl_res = None
    if not args.no_implicit:
        impl_res = implicit_leg(args, dev)

    uns_res = None
    if not args.no_uns:
        uns_res = uns_leg(args, dev)

    drift_res = None
    if not args.no_drift:
        from nga2_amd import drift

        checks = {}

        def drift_check(step, seq, x64, xmx):
            # the FP64 CPU oracle on this step's operator and right-hand side
            from oracle.orc import Oracle, interior

            c = seq.cfg
            opr_h = seq.opr.cpu().numpy().reshape((7,) + c.shape, order="F")
            b = interior(seq.rhs.cpu().numpy().reshape(c.shape, order="F"), c)
            ts = time.perf_counter()
            o = Oracle.from_cfg(c, opr_h)
            xr, st_r = o.solve(b, rcvg=1e-12, maxit=500)
            o.close()
            xr = xr - xr.mean()
            nr = np.linalg.norm(xr)
            checks[step] = {"err_fp64_vs_oracle": float(np.linalg.norm(x64.cpu().numpy() - xr) / nr),
                            "err_mixed_vs_oracle": float(np.linalg.norm(xmx.cpu().numpy() - xr) / nr),
                            "oracle_it": st_r["it"], "oracle_s": round(time.perf_counter() - ts, 1)}
            log(f"[bench] drift step {step} vs FP64 oracle: {checks[step]}")

        log(f"[bench] drifting-spheres sequence ({args.drift_n}^3, 200 spheres, density ratio 1000, "
            f"{args.drift_steps} steps)")
        drift_res = drift.run(n=args.drift_n, steps=args.drift_steps, train_steps=args.drift_train,
                              unet=not args.no_drift_unet, check=drift_check,
                              check_steps=() if args.no_cpu else (args.drift_steps - 1,), log=log)
        if checks:
            drift_res["oracle_check"] = checks
            drift_res["max_err_mixed_vs_fp64_oracle"] = max(v["err_mixed_vs_oracle"] for v in checks.values())
        drift_res["workload"] = (f"{args.drift_n}^3 periodic, 200 spheres r 0.02-0.08 drifting 0.5 cell/step (rms), "
                                 "density ratio 1000, operator assembled on the device every step, a new "
                                 "configuration-1 RHS every step; zero guess, setup + solve per step, rcvg 1e-10; "
                                 f"U-Net trained on crops of steps 0-{args.drift_train - 1}, medians over the rest")

    steps = args.steps
    el = mixed["elapsed"]
    value = steps / el
    out = {
        "metric": METRIC, "value": round(value, 4), "unit": "solves/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * el / steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": f"{P.name}: {cfg.nx}x{cfg.ny}x{cfg.nz}, x/z periodic, y walls tanh stretch 2, "
                               f"spectral RHS seed 2002; FP64 flexible PCG + FP32 PFMG-style V(2,2), "
                               f"rcvg {args.rcvg:g}, zero guess, inputs resident in HBM",
                   "cells": N, "precision": "mixed"},
        "iters": mixed["its"][0], "true_rerr": mixed["stats"]["true_rerr"], "err_vs_fp64_oracle": err_mixed,
        "setup_ms": round(1e3 * mixed["setup_s"], 2), "setup_ms_again": round(1e3 * mixed["setup_again_s"], 3),
        "levels": mixed["stats"]["nlevels"],
        "roofline": roofline, "cpu_baseline": cpu,
    }
    if fp64:
        out["fp64_plain"] = {"value": round(len(fp64["its"]) / fp64["elapsed"], 4), "unit": "solves/s",
                             "iters": fp64["its"][0], "true_rerr": fp64["stats"]["true_rerr"],
                             "err_vs_fp64_oracle": err_fp64,
                             "ms_per_solve": round(1e3 * fp64["elapsed"] / len(fp64["its"]), 3)}
        out["speedup_mixed_vs_fp64"] = round(out["value"] / out["fp64_plain"]["value"], 3)
    if host_res:
        out["host_path"] = host_res
    if two_phase:
        out["two_phase_256"] = two_phase
    if step_res:
        out["device_time_step"] = step_res
    if c1_res:
        out["config1_64"] = c1_res
    if sequence:
        out["learned_sequence"] = sequence
    if drift_res:
        out["two_phase_drift"] = drift_res
    if fft_res:
        out["fourier3d"] = fft_res
    if impl_res:
        out["implicit_gmres"] = impl_res
    if uns_res:
        out["uns_amg"] = uns_res
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
