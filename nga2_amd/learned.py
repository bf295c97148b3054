"""Learned initial guess over a sequence of pressure solves.

The workload is the correlated right-hand-side sequence of inputs.RhsSequence:
a 128^3 periodic box with 200 steps.

1. ``collect``: solve steps 0..T-1, each from the previous step's solution
   (the method without the network). For each step it records
   r_n = b_n - A p_{n-1} and e_n = p_n - p_{n-1}, both over the interior
   cells with x fastest.
2. ``train``: fit the U-Net (unet.UNet, trained by the project's own HIP
   forward and backward kernels) so that kappa s_n N(r_n / s_n) leaves the
   smallest residual r_n - A kappa s_n N(r_n / s_n). Here s_n = rms(r_n), and
   kappa is the median of rms(e_n)/rms(r_n) over the training pairs. The
   alternative loss fits N(r_n / s_n) to e_n / (kappa s_n). The periodic
   problem fixes p only up to a constant, so the mean of e_n is removed.
3. ``evaluate``: solve each of steps T..199 twice from the same p_{n-1}, to
   the same tolerance with the same stopping rule:
   - from the guess p_{n-1} (previous-step guess);
   - with the solver's guess stage attached (ngap_set_stage), which starts
     from p_{n-1} + kappa s N(r / s) and keeps that guess only if it lowers
     the residual.
   It reports iterations, solve wall time (the stage's own cost included) and
   the true residuals. With ``precond_iters`` > 0 the network also acts as a
   flexible-PCG stage in the first iterations.
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib, inputs
from .linsol import MIXED, DeviceSolve
from .unet import UNet


def interior_flat(t, cfg):
    """Interior of a ghosted Fortran-ordered flat device array, x fastest."""
    no = cfg.no
    return t.reshape(cfg.nzo, cfg.nyo, cfg.nxo)[no:no + cfg.nz, no:no + cfg.ny, no:no + cfg.nx].reshape(-1)


def add_interior(t, e, cfg):
    """t[interior] += e in place (t ghosted Fortran-ordered flat, e interior x fastest)."""
    no = cfg.no
    v = t.reshape(cfg.nzo, cfg.nyo, cfg.nxo)[no:no + cfg.nz, no:no + cfg.ny, no:no + cfg.nx]
    v += e.reshape(cfg.nz, cfg.ny, cfg.nx)


class SequenceSolver:
    """A DeviceSolve set up once for the sequence's fixed operator."""

    def __init__(self, n=128, seed=4004, kmax=16, precision=MIXED, rcvg=1e-10, maxit=200, device="cuda:0"):
        import torch

        self.seq = inputs.RhsSequence(n=n, seed=seed, kmax=kmax)
        self.cfg = self.seq.cfg
        self.dev = torch.device(device)
        self.N = self.cfg.nx * self.cfg.ny * self.cfg.nz
        d_opr = torch.from_numpy(np.ascontiguousarray(self.seq.opr.ravel(order="F"))).to(self.dev)
        self.ds = DeviceSolve(self.cfg)
        self.ds.setup(d_opr, maxit=maxit, rcvg=rcvg, precision=precision)
        self.rcvg = rcvg
        self._res = torch.zeros(self.N, dtype=torch.float64, device=self.dev)

    def rhs(self, step):
        return self.seq.rhs_device(step, self.dev)

    def residual(self, b, x):
        """(r = b - A x over the interior, ||r||_2)"""
        nrm = self.ds.residual(b, x, self._res)
        return self._res.clone(), nrm

    def solve(self, b, x0, stage=None, flags=_lib.STAGE_GUESS, precond_iters=0):
        """Solve from the guess x0 (not modified); returns (x, stats)."""
        import torch

        x = x0.clone()
        self.ds.set_stage(stage, flags=flags, precond_iters=precond_iters)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        self.ds.solve(b, x)
        wall = (time.perf_counter() - t0) * 1e3
        st = self.ds.stats()
        st["wall_ms"] = wall
        self.ds.set_stage(None)
        return x, st

    def close(self):
        self.ds.destroy()


def collect(S, steps, x0=None):
    """Solve `steps` from the previous solution; return the pairs and the
    last solution. Pairs are (r_n, e_n) FP64 device vectors."""
    import torch

    x = x0 if x0 is not None else torch.zeros(S.cfg.nxo * S.cfg.nyo * S.cfg.nzo, dtype=torch.float64, device=S.dev)
    pairs, stats = [], []
    for n in steps:
        b = S.rhs(n)
        r, _ = S.residual(b, x)
        xn, st = S.solve(b, x)
        e = interior_flat(xn, S.cfg) - interior_flat(x, S.cfg)
        pairs.append((r, e - e.mean()))
        stats.append(st)
        x = xn
    return pairs, stats, x


def fit_kappa(pairs):
    import torch

    ratios = [float(torch.sqrt((e * e).mean()) / torch.sqrt((r * r).mean())) for r, e in pairs]
    return float(np.median(ratios))


def train(net, pairs, kappa, S=None, epochs=8, lr=2e-3, lr_end=1e-4, seed=0, loss="residual", log=None):
    """Adam on the normalised pairs, one field per step, cosine-decayed rate.
    Returns the per-epoch mean losses.

    With x = r / s (s = rms(r)) and y = net(x):
    - loss "residual" (needs the solver S) is ||x - kappa A y||^2 / V. This is
      the relative residual left by the guess, which decides the iterations
      that follow; its gradient is -2 kappa A (x - kappa A y) / V, with A
      symmetric.
    - loss "error" is ||y - e / (kappa s)||^2 / V."""
    import torch

    if loss not in ("residual", "error"):
        raise ValueError(loss)
    if loss == "residual" and S is None:
        raise ValueError("the residual loss needs the solver's operator")
    xs, ts = [], []
    for r, e in pairs:
        s = torch.sqrt((r * r).mean())
        xs.append((r / s).to(torch.float32))
        ts.append((e / (s * kappa)).to(torch.float32) if loss == "error" else None)
    net.set_kappa(kappa)
    V = xs[0].numel()
    q = torch.empty(V, dtype=torch.float64, device=xs[0].device)
    g = torch.empty_like(q)
    rng = np.random.default_rng(seed)
    total = epochs * len(pairs)
    k = 0
    hist = []
    for ep in range(epochs):
        order = rng.permutation(len(pairs))
        losses = []
        for i in order:
            a = lr_end + 0.5 * (lr - lr_end) * (1.0 + np.cos(np.pi * k / max(total - 1, 1)))
            if loss == "error":
                losses.append(net.train_step(xs[i], ts[i], lr=float(a)))
            else:
                y = net.forward(xs[i])
                S.ds.apply((kappa * y).to(torch.float64), q)
                res = xs[i].to(torch.float64) - q
                losses.append(float((res * res).mean()))
                S.ds.apply(res, g)
                net.backward(((-2.0 * kappa / V) * g).to(torch.float32))
                net.adam_step(float(a))
            k += 1
        hist.append(float(np.mean(losses)))
        if log:
            log(f"epoch {ep}: loss {hist[-1]:.4e}")
    return hist


def evaluate(S, net, steps, x0, precond_iters=0):
    """Solve each step from p_{n-1} with and without the guess stage.
    Returns per-step records."""
    import torch

    out = []
    x = x0
    stage = net.stage()
    for n in steps:
        b = S.rhs(n)
        r, rn = S.residual(b, x)
        bn = float(torch.linalg.vector_norm(interior_flat(b, S.cfg)))
        xa, sa = S.solve(b, x)
        flags = _lib.STAGE_GUESS | (_lib.STAGE_PRECOND if precond_iters > 0 else 0)
        xb, sb = S.solve(b, x, stage=stage, flags=flags, precond_iters=precond_iters)
        # residual of the learned guess alone (what the stage starts from)
        g = x.clone()
        add_interior(g, net.apply(r), S.cfg)
        _, rg = S.residual(b, g)
        dx = interior_flat(xb, S.cfg) - interior_flat(xa, S.cfg)
        dx = dx - dx.mean()
        xa_i = interior_flat(xa, S.cfg)
        xa_i = xa_i - xa_i.mean()
        out.append({
            "step": n,
            "r0_prev": rn / bn, "r0_learned": rg / bn,
            "it_prev": sa["it"], "it_learned": sb["it"],
            "ms_prev": sa["solve_ms"], "ms_learned": sb["solve_ms"],
            "true_prev": sa["true_rerr"], "true_learned": sb["true_rerr"],
            "stage_calls": sb["stage_calls"], "fallback": sb["fallback"],
            "diff": float(torch.linalg.vector_norm(dx) / torch.linalg.vector_norm(xa_i)),
        })
        x = xa
    return out, x


def summary(recs):
    f = lambda k: float(np.mean([r[k] for r in recs]))  # noqa: E731
    return {
        "steps": len(recs),
        "it_prev": f("it_prev"), "it_learned": f("it_learned"),
        "ms_prev": f("ms_prev"), "ms_learned": f("ms_learned"),
        "r0_prev": f("r0_prev"), "r0_learned": f("r0_learned"),
        "max_true_prev": max(r["true_prev"] for r in recs),
        "max_true_learned": max(r["true_learned"] for r in recs),
        "max_solution_diff": max(r["diff"] for r in recs),
        "guess_rejected": int(sum(1 for r in recs if r["fallback"] & 2)),
    }


def run(n=128, steps=200, train_steps=150, epochs=8, lr=2e-3, seed=0, precond_iters=0, log=print):
    """The whole workflow; returns a dict of timings and summaries."""
    t0 = time.perf_counter()
    S = SequenceSolver(n=n)
    pairs, st_train, x = collect(S, range(train_steps))
    t_collect = time.perf_counter() - t0
    kappa = fit_kappa(pairs)
    net = UNet((n, n, n), seed=seed)
    t1 = time.perf_counter()
    hist = train(net, pairs, kappa, S=S, epochs=epochs, lr=lr, log=log)
    t_train = time.perf_counter() - t1
    del pairs
    recs, _ = evaluate(S, net, range(train_steps, steps), x, precond_iters=precond_iters)
    res = {"kappa": kappa, "loss_history": hist, "collect_s": t_collect, "train_s": t_train,
           "train_solves_it": float(np.mean([s["it"] for s in st_train])), "eval": summary(recs)}
    net.close()
    S.close()
    return res
