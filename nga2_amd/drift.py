"""Two-phase sequence with drifting spheres: a periodic unit cube of n^3 cells
with 200 spheres, radii 0.02-0.08, and density ratio 1000. Each sphere drifts
with its own seeded constant velocity, so the operator changes every step.

Per step, everything stays on the GPU:
1. VF of the spheres at their new positions (sphere_vf_device, the
   inputs.sphere_vf algorithm in torch);
2. face densities (the inputs.face_densities rule);
3. the operator by the device-side projection (ngap_proj_laplacian);
4. setup (ngap_setup_device) and solve.
The right-hand side is -vol div(U) of one seeded Gaussian face velocity field
(the configuration-1 recipe), held fixed over the steps. So the solution
moves only because the spheres move, and the previous solutions predict the
next one.

Compared per step, all to rcvg with the same stopping rule:
- plain FP64 PCG from a zero guess (the method as NGA2 calls it);
- plain FP64 from the previous solution;
- the mixed-precision solve from the learned temporal prediction
  (learned.fit_predictor fitted on the first steps' solutions). The prediction
  is used only if its residual is below ||b||, the residual of a zero guess;
  the check (one residual evaluation) is timed with the solve.
The time of a step is setup + solve, since tpns sets the solver up before
every solve (tpns_class.f90:1592).
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
    """The sequence's geometry, right-hand side and per-step operator."""

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
        cfg = self.cfg
        No = cfg.nxo * cfg.nyo * cfg.nzo
        self.No = No
        # fixed right-hand side: -vol div(U) of seeded Gaussian faces, mean removed
        U, V, W = inputs._gaussian_faces(cfg, inputs._rng(rhs_seed))
        d = [torch.from_numpy(np.ascontiguousarray(a.ravel(order="F"))).to(self.dev) for a in (U, V, W)]
        div = torch.zeros(No, dtype=torch.float64, device=self.dev)
        self.proj.div(*d, div)
        vol = torch.from_numpy(np.ascontiguousarray(np.asarray(cfg.vol).ravel(order="F"))).to(self.dev)
        rhs = -vol * div
        ri = learned.interior_flat(rhs, cfg)
        self.rhs = learned.to_ghosted(ri - ri.mean(), cfg)
        self.opr = torch.zeros(7 * No, dtype=torch.float64, device=self.dev)

    def positions(self, step):
        return (self.centres + step * self.vel) % 1.0

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


def run(n=256, steps=40, fit_steps=16, m=6, nsph=200, drift_cells=0.5, log=print, oracle_steps=(), device="cuda:0"):
    """Solve the sequence three ways per step and time setup + solve.
    Steps 0..fit_steps-1 are solved plain FP64 from the previous solution and
    fit the predictor; the remaining steps are compared. oracle_steps: steps
    whose FP64-oracle error is measured (CPU, slow at 256^3)."""
    import torch

    seq = DriftingSpheres(n=n, nsph=nsph, drift_cells=drift_cells, device=device)
    cfg = seq.cfg
    N = cfg.nx * cfg.ny * cfg.nz
    dev = seq.dev
    ds64 = DeviceSolve(cfg)
    dsmx = DeviceSolve(cfg)
    zero = torch.zeros(seq.No, dtype=torch.float64, device=dev)

    def timed(ds, prec, guess_int):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ds.setup(seq.opr, maxit=500, rcvg=1e-10, precision=prec)
        t1 = time.perf_counter()
        x = learned.to_ghosted(guess_int, cfg) if guess_int is not None else zero.clone()
        it, _ = ds.solve(seq.rhs, x)
        t2 = time.perf_counter()
        st = ds.stats()
        return learned.interior_flat(x, cfg).clone(), {"it": it, "setup_ms": (t1 - t0) * 1e3,
                                                        "solve_ms": (t2 - t1) * 1e3, "true_rerr": st["true_rerr"]}

    hist = []
    recs = []
    alpha = None
    t_asm = []
    t_vf, t_lap = [], []
    res_buf = torch.zeros(N, dtype=torch.float64, device=dev)
    b_norm = float(torch.linalg.vector_norm(learned.interior_flat(seq.rhs, cfg)))

    def dsmx_residual(x_int):
        # ||b - A x|| with the current operator (ds64 was set up from it this step)
        return ds64.residual(seq.rhs, learned.to_ghosted(x_int, cfg), res_buf)
    for step in range(steps):
        torch.cuda.synchronize()
        ta = time.perf_counter()
        seq.operator(step)
        torch.cuda.synchronize()
        t_asm.append((time.perf_counter() - ta) * 1e3)
        t_vf.append(seq.last_ms["vf_ms"])
        t_lap.append(seq.last_ms["laplacian_ms"])
        prev = hist[-1] if hist else torch.zeros(N, dtype=torch.float64, device=dev)
        x_prev, s_prev = timed(ds64, FP64, prev)
        x_prev = learned._mean_free(x_prev)
        if step < fit_steps:
            hist.append(x_prev)
            log(f"[drift] step {step}: fp64 from previous it={s_prev['it']}")
            continue
        if alpha is None:
            alpha = learned.fit_predictor(hist, m)
            log(f"[drift] predictor alpha={np.round(alpha, 3).tolist()}")
        x_zero, s_zero = timed(ds64, FP64, None)
        x_ext = learned.extrapolate(hist[-(m + 1):], alpha)
        # guard: when the operator moved so much that the prediction's residual
        # exceeds ||b|| (the residual of a zero guess), start from zero instead
        torch.cuda.synchronize()
        tg = time.perf_counter()
        r_ext = dsmx_residual(x_ext)
        use_ext = r_ext < b_norm
        guard_ms = (time.perf_counter() - tg) * 1e3
        x_mix, s_mix = timed(dsmx, MIXED, x_ext if use_ext else None)
        s_mix["solve_ms"] += guard_ms
        s_mix["guess_used"] = int(use_ext)
        ref = learned._mean_free(x_zero)
        nref = float(torch.linalg.vector_norm(ref))
        rec = {"step": step, "zero_fp64": s_zero, "prev_fp64": s_prev, "learned_mixed": s_mix,
               "diff_mixed": float(torch.linalg.vector_norm(learned._mean_free(x_mix) - ref)) / nref,
               "diff_prev": float(torch.linalg.vector_norm(x_prev - ref)) / nref}
        if step in oracle_steps:
            from oracle.orc import Oracle, interior

            opr_h = seq.opr.cpu().numpy().reshape((7,) + cfg.shape, order="F")
            b = interior(seq.rhs.cpu().numpy().reshape(cfg.shape, order="F"), cfg)
            o = Oracle.from_cfg(cfg, opr_h)
            xr, _ = o.solve(b, rcvg=1e-12, maxit=500)
            o.close()
            xr = xr - xr.mean()
            xm = learned._mean_free(x_mix).cpu().numpy()  # interior, x fastest, as the oracle's
            rec["err_mixed_vs_oracle"] = float(np.linalg.norm(xm - xr) / np.linalg.norm(xr))
        recs.append(rec)
        hist = hist[1:] + [x_prev]
        log(f"[drift] step {step}: it zero/prev/learned {s_zero['it']}/{s_prev['it']}/{s_mix['it']}, "
            f"setup+solve ms {s_zero['setup_ms'] + s_zero['solve_ms']:.1f}/"
            f"{s_prev['setup_ms'] + s_prev['solve_ms']:.1f}/{s_mix['setup_ms'] + s_mix['solve_ms']:.1f}")
    ds64.destroy()
    dsmx.destroy()
    seq.proj.destroy()

    def med(key, sub):
        return float(np.median([r[key][sub] for r in recs]))

    out = {"n": n, "spheres": nsph, "steps": steps, "fit_steps": fit_steps, "m": m, "drift_cells": drift_cells,
           "assembly_ms": float(np.median(t_asm)), "vf_generation_ms": float(np.median(t_vf)),
           "face_densities_laplacian_ms": float(np.median(t_lap))}
    for k in ("zero_fp64", "prev_fp64", "learned_mixed"):
        out[k] = {"it": med(k, "it"), "setup_ms": med(k, "setup_ms"), "solve_ms": med(k, "solve_ms"),
                  "max_true_rerr": max(r[k]["true_rerr"] for r in recs)}
        out[k]["step_ms"] = out[k]["setup_ms"] + out[k]["solve_ms"]
    out["speedup_learned_mixed_vs_zero_fp64"] = out["zero_fp64"]["step_ms"] / out["learned_mixed"]["step_ms"]
    out["speedup_learned_mixed_vs_prev_fp64"] = out["prev_fp64"]["step_ms"] / out["learned_mixed"]["step_ms"]
    out["max_diff_mixed"] = max(r["diff_mixed"] for r in recs)
    out["learned_guess_used"] = int(sum(r["learned_mixed"]["guess_used"] for r in recs))
    errs = [r["err_mixed_vs_oracle"] for r in recs if "err_mixed_vs_oracle" in r]
    if errs:
        out["max_err_mixed_vs_fp64_oracle"] = max(errs)
    return out
