"""Two-phase sequence with drifting spheres: a periodic unit cube of n^3 cells
(512^3 by default) with 200 spheres, radii 0.02-0.08, and density ratio 1000.
Each sphere drifts with its own seeded constant velocity, so the operator
changes every step.

Per step, everything stays on the GPU:
1. VF of the spheres at their new positions (sphere_vf_device, the
   inputs.sphere_vf algorithm in torch);
2. face densities (the inputs.face_densities rule);
3. the operator by the device-side projection (ngap_proj_laplacian);
4. the step's right-hand side by the configuration-1 recipe: Gaussian face
   velocities drawn afresh for the step, their divergence (ngap_proj_div),
   rhs = -vol div / dt with dt = 1 (the two-phase form of
   examples/falling_drop/src/simulation.f90:302), mean removed;
5. setup (ngap_setup_device) and solve, as tpns sets the solver up before
   every solve (src/two_phase/tpns_class.f90:1592).

Compared per step, all from a zero guess (examples/channel/src/simulation.f90:319
sets sol = 0 before the solve) and to rcvg with the same stopping rule:
- plain FP64 PCG with the V(1,1) cycle (the method as NGA2 calls it);
- mixed precision: FP64 PCG around the FP32 V-cycle;
- mixed precision with the learned stage: the U-Net gives the initial guess
  and, in the first precond_iters iterations, an extra flexible-PCG stage.
  The network sees the Jacobi-scaled residual D^-1 r (D = diag A), which
  carries the local density, so one set of weights serves every step's
  operator. It is trained on the first train_steps steps' own solutions
  (crops of crop^3 cells, the residual each correction leaves as the loss),
  and the solver keeps a guess only if it lowers the residual.
The right-hand sides of different steps are independent draws, so no
temporal predictor applies here; learned.py covers the correlated sequence.
"""
from __future__ import annotations

import time

import numpy as np

from . import incomp, inputs, learned
from .config import Config, uniform
from .linsol import FP64, MIXED, DeviceSolve
from .proj import DeviceProjection


def sphere_vf_device(n, centres, radii, nsub=4, device="cuda:0"):
    """inputs.sphere_vf on the device (same arithmetic, torch float64):
    VF[i, j, k] of the union of spheres on the n^3 periodic unit cube."""
    import torch

    dev = torch.device(device)
    h = 1.0 / n
    hd = 0.5 * np.sqrt(3.0) * h
    xc = (torch.arange(n, dtype=torch.float64, device=dev) + 0.5) * h
    G = torch.full((n, n, n), -float("inf"), dtype=torch.float64, device=dev)
    for c, r in zip(centres, radii):
        lo = np.floor((c - r - 2 * hd) / h).astype(int)
        hi = np.ceil((c + r + 2 * hd) / h).astype(int)
        ii = torch.arange(lo[0], hi[0] + 1, device=dev) % n
        jj = torch.arange(lo[1], hi[1] + 1, device=dev) % n
        kk = torch.arange(lo[2], hi[2] + 1, device=dev) % n
        dx = (xc[ii] - c[0] + 0.5) % 1.0 - 0.5
        dy = (xc[jj] - c[1] + 0.5) % 1.0 - 0.5
        dz = (xc[kk] - c[2] + 0.5) % 1.0 - 0.5
        d = torch.sqrt(dx[:, None, None] ** 2 + dy[None, :, None] ** 2 + dz[None, None, :] ** 2)
        g = r - d
        ix = (ii[:, None, None], jj[None, :, None], kk[None, None, :])
        G[ix] = torch.maximum(G[ix], g)
    VF = torch.where(G > hd, 1.0, 0.0).to(torch.float64)
    ci, cj, ck = torch.nonzero(torch.abs(G) <= hd, as_tuple=True)
    M = ci.numel()
    if M:
        off = (torch.arange(nsub, dtype=torch.float64, device=dev) + 0.5) / nsub - 0.5
        ox, oy, oz = torch.meshgrid(off, off, off, indexing="ij")
        ox, oy, oz = (o.reshape(-1) * h for o in (ox, oy, oz))
        inside = torch.zeros((M, nsub ** 3), dtype=torch.bool, device=dev)
        ccx, ccy, ccz = xc[ci], xc[cj], xc[ck]
        for c, r in zip(centres, radii):
            dx = (ccx - c[0] + 0.5) % 1.0 - 0.5
            dy = (ccy - c[1] + 0.5) % 1.0 - 0.5
            dz = (ccz - c[2] + 0.5) % 1.0 - 0.5
            dc = torch.sqrt(dx * dx + dy * dy + dz * dz)
            cand = torch.nonzero(torch.abs(r - dc) <= hd, as_tuple=True)[0]
            if cand.numel() == 0:
                continue
            px = dx[cand, None] + ox[None, :]
            py = dy[cand, None] + oy[None, :]
            pz = dz[cand, None] + oz[None, :]
            inside[cand] |= (px * px + py * py + pz * pz) < r * r
        VF[ci, cj, ck] = inside.to(torch.float64).mean(dim=1)
    return VF


def face_densities_device(cfg, vf, rho_l, rho_g):
    """inputs.face_densities on the device; vf is [i, j, k] (n^3), the result
    three ghosted Fortran-ordered flat tensors (rho_U, rho_V, rho_W)."""
    import torch

    dev = vf.device
    no = cfg.no
    # ghosted [k, j, i] (C order of the Fortran-ordered flat array)
    rho = torch.zeros((cfg.nzo, cfg.nyo, cfg.nxo), dtype=torch.float64, device=dev)
    rho[no:no + cfg.nz, no:no + cfg.ny, no:no + cfg.nx] = (vf * rho_l + (1.0 - vf) * rho_g).permute(2, 1, 0)
    # periodic ghosts, x then y then z (cfg.sync; every direction periodic here)
    for d, n in ((2, cfg.nx), (1, cfg.ny), (0, cfg.nz)):
        sl = [slice(None)] * 3
        src = [slice(None)] * 3
        sl[d] = slice(0, no)
        src[d] = slice(n, n + no)
        rho[tuple(sl)] = rho[tuple(src)]
        sl[d] = slice(n + no, n + 2 * no)
        src[d] = slice(no, 2 * no)
        rho[tuple(sl)] = rho[tuple(src)]
    out = []
    for d in (2, 1, 0):  # x, y, z faces
        rf = torch.ones_like(rho)
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[d] = slice(0, -1)
        hi[d] = slice(1, None)
        rf[tuple(hi)] = 0.5 * (rho[tuple(lo)] + rho[tuple(hi)])
        out.append(rf.reshape(-1))
    return tuple(out)


class DriftingSpheres:
    """The sequence's geometry, per-step right-hand side and per-step operator."""

    def __init__(self, n=512, nsph=200, rmin=0.02, rmax=0.08, seed=5005, vel_seed=5006, rhs_seed=1001,
                 rho_l=1000.0, rho_g=1.0, drift_cells=0.5, device="cuda:0"):
        import torch

        self.n = n
        self.dev = torch.device(device)
        x = uniform(n)
        self.cfg = Config(x, x, x, no=1, xper=True, yper=True, zper=True, name="drifting_spheres")
        self.met = incomp.Metrics(self.cfg)
        self.proj = DeviceProjection(self.met)
        self.centres, self.radii = inputs.sphere_set(nsph, rmin, rmax, seed)
        rng = np.random.Generator(np.random.PCG64(vel_seed))
        v = rng.standard_normal((nsph, 3))
        self.vel = v / np.sqrt((v * v).sum(axis=1, keepdims=True).mean()) * (drift_cells / n)  # per step
        self.rho_l, self.rho_g = rho_l, rho_g
        self.rhs_seed = rhs_seed
        cfg = self.cfg
        No = cfg.nxo * cfg.nyo * cfg.nzo
        self.No = No
        self.N = cfg.nx * cfg.ny * cfg.nz
        self.opr = torch.zeros(7 * No, dtype=torch.float64, device=self.dev)
        self.rhs = torch.zeros(No, dtype=torch.float64, device=self.dev)
        self._faces = [torch.zeros(No, dtype=torch.float64, device=self.dev) for _ in range(3)]
        self._div = torch.zeros(No, dtype=torch.float64, device=self.dev)

    def positions(self, step):
        return (self.centres + step * self.vel) % 1.0

    def rhs_step(self, step):
        """self.rhs <- the step's right-hand side by the configuration-1 recipe
        (learned.gaussian_rhs_device), seeded by rhs_seed and the step."""
        return learned.gaussian_rhs_device(self.proj, self.cfg, int(self.rhs_seed) * 1000003 + int(step), self.rhs,
                                           faces=self._faces, div=self._div)

    def operator(self, step):
        """Assemble the step's operator into self.opr (device); returns VF.
        self.last_ms gets the times of the VF generation (input synthesis, a
        stand-in for the VOF transport) and of the face densities + operator
        assembly (the device projection)."""
        import torch

        torch.cuda.synchronize()
        t0 = time.perf_counter()
        vf = sphere_vf_device(self.n, self.positions(step), self.radii, device=self.dev)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        rf = face_densities_device(self.cfg, vf, self.rho_l, self.rho_g)
        self.proj.laplacian(self.opr, rf)
        torch.cuda.synchronize()
        self.last_ms = {"vf_ms": (t1 - t0) * 1e3, "laplacian_ms": (time.perf_counter() - t1) * 1e3}
        return vf


def run(n=512, steps=100, train_steps=12, nsph=200, drift_cells=0.5, unet=True, crop=128, crops_per_step=6,
        epochs=6, lr=5e-4, precond_iters=1, check=None, check_steps=(), log=print, device="cuda:0"):
    """Solve the sequence per step with plain FP64, mixed, and mixed + U-Net
    (from step train_steps on), timing setup + solve. The first train_steps
    steps collect the network's training crops from their FP64 solutions.
    check(step, seq, x_fp64, x_mixed): called for the steps in check_steps
    with the interior solutions (the caller's comparison against the oracle).
    Returns a dict of per-mode medians over the evaluated steps."""
    import torch

    from . import _lib
    from .unet import UNet

    seq = DriftingSpheres(n=n, nsph=nsph, drift_cells=drift_cells, device=device)
    cfg = seq.cfg
    N = seq.N
    dev = seq.dev
    ds64 = DeviceSolve(cfg)
    dsmx = DeviceSolve(cfg)
    dinv = torch.zeros(N, dtype=torch.float64, device=dev)
    x = torch.zeros(seq.No, dtype=torch.float64, device=dev)
    gen = np.random.default_rng(77)
    samples, kappas = [], []
    net = net_full = None
    train_info = {}
    recs = []
    t_vf, t_lap, t_rhs = [], [], []

    def solve(ds, prec, setup=True, stage=None):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if setup:
            ds.setup(seq.opr, maxit=500, rcvg=1e-10, precision=prec)
        t1 = time.perf_counter()
        ds.set_stage(stage, flags=_lib.STAGE_GUESS | (_lib.STAGE_PRECOND if precond_iters > 0 else 0),
                     precond_iters=precond_iters)
        try:
            x.zero_()
            it, _ = ds.solve(seq.rhs, x)
        finally:
            ds.set_stage(None)
        t2 = time.perf_counter()
        st = ds.stats()
        return learned.interior_flat(x, cfg).clone(), {
            "it": it, "setup_ms": (t1 - t0) * 1e3, "solve_ms": (t2 - t1) * 1e3, "true_rerr": st["true_rerr"],
            "stage_calls": int(st.get("stage_calls", 0)), "guess_rejected": int(bool(st.get("fallback", 0) & 2))}

    for step in range(steps):
        seq.operator(step)
        torch.cuda.synchronize()
        tr = time.perf_counter()
        seq.rhs_step(step)
        torch.cuda.synchronize()
        t_rhs.append((time.perf_counter() - tr) * 1e3)
        t_vf.append(seq.last_ms["vf_ms"])
        t_lap.append(seq.last_ms["laplacian_ms"])
        x64, s64 = solve(ds64, FP64)
        xmx, smx = solve(dsmx, MIXED)
        x64m, xmxm = learned._mean_free(x64), learned._mean_free(xmx)
        nref = float(torch.linalg.vector_norm(x64m))
        rec = {"step": step, "fp64": s64, "mixed": smx,
               "diff_mixed": float(torch.linalg.vector_norm(xmxm - x64m)) / nref}
        if unet and step < train_steps:
            learned.diag_inverse(seq.opr, cfg, dinv)
            b = learned.interior_flat(seq.rhs, cfg)
            smp, kap = learned.crop_samples(seq.opr, cfg, b, x64m, dinv, crop, crops_per_step, gen)
            samples += smp
            kappas += kap
            if step == train_steps - 1:
                t0 = time.perf_counter()
                kappa = float(np.median(kappas))
                net = UNet((crop, crop, crop), seed=0, device=device)
                hist, best = learned.train_on_crops(net, samples, kappa, epochs=epochs, lr=lr,
                                                    log=lambda m: log("[drift] " + m))
                state = net.state()
                net.close()
                samples.clear()
                net_full = UNet((n, n, n), seed=0, device=device)
                net_full.load_state(state)
                net_full.set_input_scale(dinv)
                train_info = {"train_s": time.perf_counter() - t0, "loss_history": hist, "best_epoch": best,
                              "kappa": kappa, "samples": train_steps * crops_per_step, "crop": crop}
                log(f"[drift] U-Net trained in {train_info['train_s']:.1f} s, loss {hist[0]:.3e} -> {min(hist):.3e}")
        elif unet and net_full is not None:
            ts = time.perf_counter()
            learned.diag_inverse(seq.opr, cfg, dinv)  # the stage's input scale for this step's operator
            torch.cuda.synchronize()
            dinv_ms = (time.perf_counter() - ts) * 1e3
            # the mixed setup of this step is already in dsmx; the U-Net solve
            # is timed with that setup and the input-scale update
            xu, su = solve(dsmx, MIXED, setup=False, stage=net_full.stage())
            su["setup_ms"] = smx["setup_ms"] + dinv_ms
            rec["unet_mixed"] = su
            rec["diff_unet"] = float(torch.linalg.vector_norm(learned._mean_free(xu) - x64m)) / nref
        if check is not None and step in check_steps:
            check(step, seq, x64m, xmxm)
        recs.append(rec)
        msg = (f"[drift] step {step}: it fp64/mixed {s64['it']}/{smx['it']}, setup+solve ms "
               f"{s64['setup_ms'] + s64['solve_ms']:.1f}/{smx['setup_ms'] + smx['solve_ms']:.1f}")
        if "unet_mixed" in rec:
            su = rec["unet_mixed"]
            msg += f", U-Net it {su['it']} {su['setup_ms'] + su['solve_ms']:.1f} ms"
        log(msg)
    ds64.destroy()
    dsmx.destroy()
    if net_full is not None:
        net_full.close()
    seq.proj.destroy()

    ev = [r for r in recs if r["step"] >= (train_steps if unet else 0)] or recs

    def summ(key):
        rr = [r[key] for r in ev if key in r]
        if not rr:
            return None
        out = {k: float(np.median([r[k] for r in rr])) for k in ("it", "setup_ms", "solve_ms")}
        out["step_ms"] = float(np.median([r["setup_ms"] + r["solve_ms"] for r in rr]))
        out["max_true_rerr"] = max(r["true_rerr"] for r in rr)
        out["steps"] = len(rr)
        return out

    out = {"n": n, "spheres": nsph, "steps": steps, "drift_cells": drift_cells, "evaluated_steps": len(ev),
           "vf_generation_ms": float(np.median(t_vf)), "face_densities_laplacian_ms": float(np.median(t_lap)),
           "rhs_ms": float(np.median(t_rhs))}
    out["zero_fp64"] = summ("fp64")
    out["mixed"] = summ("mixed")
    out["max_diff_mixed"] = max(r["diff_mixed"] for r in recs)
    out["speedup_mixed_vs_fp64"] = out["zero_fp64"]["step_ms"] / out["mixed"]["step_ms"]
    out["solves_per_s"] = {"fp64": 1e3 / out["zero_fp64"]["step_ms"], "mixed": 1e3 / out["mixed"]["step_ms"]}
    if unet and any("unet_mixed" in r for r in recs):
        u = summ("unet_mixed")
        ur = [r["unet_mixed"] for r in ev if "unet_mixed" in r]
        u["stage_calls"] = int(sum(r["stage_calls"] for r in ur))
        u["guess_rejected"] = int(sum(r["guess_rejected"] for r in ur))
        out["unet_mixed"] = u
        out["learned_guess_used"] = int(sum(1 for r in ur if not r["guess_rejected"]))
        out["max_diff_unet"] = max(r["diff_unet"] for r in recs if "diff_unet" in r)
        out["speedup_unet_mixed_vs_fp64"] = out["zero_fp64"]["step_ms"] / u["step_ms"]
        out["solves_per_s"]["unet_mixed"] = 1e3 / u["step_ms"]
        out["unet"] = dict(train_info, precond_iters=precond_iters, train_steps=train_steps)
    return out
