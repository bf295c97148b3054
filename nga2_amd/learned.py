"""Learned initial guess over a sequence of pressure solves.

The workload is the correlated right-hand-side sequence of inputs.RhsSequence:
a 128^3 periodic box with 200 steps. Training uses only what the solver
itself produces on the training steps. The guess for step n has two learned
parts.

- Temporal part: a linear predictor of the solution increment from the m
  previous increments, x_ext = p_{n-1} + sum_j alpha_j (p_{n-j} - p_{n-j-1}).
  The m coefficients are the least-squares fit over the training solutions.
- Spatial part: the U-Net (unet.UNet, trained by the project's own HIP
  forward and backward kernels). It maps the residual r' = b_n - A x_ext of
  the extrapolated guess to a correction e' ~ A^{-1} r'. It is applied
  scale-free (kappa s N(r' / s), s = rms(r')) by the solver's guess stage
  (ngap_set_stage), which keeps the corrected guess only if it lowers the
  residual.

Steps:

1. ``collect``: solve steps 0..T-1, each from the previous solution (the
   method without the network), and keep the solutions.
2. ``fit_predictor``: fit the m coefficients alpha from the training
   solutions.
3. ``stage_pairs`` + ``train``: build the pairs (r', e' = p_n - x_ext) on the
   training steps and fit the network to them. The default loss is the
   residual the correction leaves, ||r'/s - kappa A N(r'/s)||^2. Here
   kappa is the median of rms(e')/rms(r'). The periodic problem fixes p only
   up to a constant, so increments and corrections are mean-free.
4. ``evaluate``: solve each of steps T..199 three times from the same
   history, to the same tolerance with the same stopping rule:
   - from the previous solution (previous-step guess);
   - from x_ext (temporal part alone);
   - from x_ext with the network stage (the learned guess).
   It reports iterations, solve wall time (the stage's own cost included),
   initial and true residuals.
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


def to_ghosted(e, cfg):
    """Ghosted Fortran-ordered flat array holding the interior vector e (ghosts 0)."""
    import torch

    g = torch.zeros(cfg.nxo * cfg.nyo * cfg.nzo, dtype=e.dtype, device=e.device)
    no = cfg.no
    g.reshape(cfg.nzo, cfg.nyo, cfg.nxo)[no:no + cfg.nz, no:no + cfg.ny, no:no + cfg.nx] = \
        e.reshape(cfg.nz, cfg.ny, cfg.nx)
    return g


def _mean_free(v):
    return v - v.mean()


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

    def residual(self, b, x_int):
        """(r = b - A x over the interior, ||r||_2) for an interior guess x_int."""
        nrm = self.ds.residual(b, to_ghosted(x_int, self.cfg), self._res)
        return self._res.clone(), nrm

    def solve(self, b, x_int, stage=None, flags=_lib.STAGE_GUESS, precond_iters=0):
        """Solve from the interior guess x_int; returns (interior solution, stats)."""
        import torch

        x = to_ghosted(x_int, self.cfg)
        self.ds.set_stage(stage, flags=flags, precond_iters=precond_iters)
        try:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            self.ds.solve(b, x)
            wall = (time.perf_counter() - t0) * 1e3
            st = self.ds.stats()
        finally:
            self.ds.set_stage(None)
        st["wall_ms"] = wall
        return interior_flat(x, self.cfg).clone(), st

    def close(self):
        self.ds.destroy()


def collect(S, steps, x0=None):
    """Solve `steps` in order, each from the previous solution. Returns the
    mean-free interior solutions and the solve statistics."""
    import torch

    x = x0 if x0 is not None else torch.zeros(S.N, dtype=torch.float64, device=S.dev)
    sols, stats = [], []
    for n in steps:
        xn, st = S.solve(S.rhs(n), x)
        x = _mean_free(xn)
        sols.append(x)
        stats.append(st)
    return sols, stats


def fit_predictor(sols, m):
    """alpha minimising sum_n ||d_n - sum_j alpha_j d_{n-j}||^2 over the
    increments d_n = p_n - p_{n-1} of the training solutions (n >= m + 1).
    Solved from the m x m normal equations in float64, with a 1e-14 relative
    ridge."""
    import torch

    d = [sols[i] - sols[i - 1] for i in range(1, len(sols))]
    if len(d) <= m:
        raise ValueError(f"need more than {m + 1} training solutions")
    G = np.zeros((m, m))
    h = np.zeros(m)
    for t in range(m, len(d)):
        past = [d[t - j] for j in range(1, m + 1)]
        for i in range(m):
            h[i] += float(torch.dot(past[i], d[t]))
            for j in range(i, m):
                G[i, j] += float(torch.dot(past[i], past[j]))
    G = G + np.triu(G, 1).T
    G += 1e-14 * np.trace(G) / m * np.eye(m)
    return np.linalg.solve(G, h)


def extrapolate(hist, alpha):
    """x_ext from the last m + 1 solutions (oldest first)."""
    m = len(alpha)
    x = hist[-1].clone()
    for j in range(1, m + 1):
        x += float(alpha[j - 1]) * (hist[-j] - hist[-j - 1])
    return x


def stage_pairs(S, sols, alpha, steps):
    """(r', e') pairs of the network stage for the given training steps
    (indices into sols, each >= m + 1)."""
    m = len(alpha)
    pairs = []
    for n in steps:
        x_ext = extrapolate(sols[n - m - 1:n], alpha)
        r, _ = S.residual(S.rhs(n), x_ext)
        pairs.append((r, _mean_free(sols[n] - x_ext)))
    return pairs


def fit_kappa(pairs):
    import torch

    ratios = [float(torch.sqrt((e * e).mean()) / torch.sqrt((r * r).mean())) for r, e in pairs]
    return float(np.median(ratios))


def train(net, pairs, kappa, S=None, epochs=8, lr=2e-3, lr_end=1e-4, seed=0, loss="residual", log=None):
    """Adam on the normalised pairs, one field per step, cosine-decayed rate.
    Returns the per-epoch mean losses; the net is left holding the weights
    of the epoch with the lowest mean loss (not those of the last epoch).

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
        if hist[-1] <= min(hist):
            best = net.state()
    net.load_state(best)
    return hist


def evaluate(S, net, alpha, sols, steps, precond_iters=0):
    """Solve each step from the same history three ways (previous solution,
    temporal extrapolation, extrapolation + network stage). `sols` holds at
    least the last m + 1 solutions before the first step; the history is
    advanced with the previous-step solve's solution. Returns per-step
    records."""
    import torch

    m = len(alpha)
    hist = list(sols[-(m + 1):])
    stage = net.stage() if net is not None else None
    flags = _lib.STAGE_GUESS | (_lib.STAGE_PRECOND if precond_iters > 0 else 0)
    out = []
    for n in steps:
        b = S.rhs(n)
        bn = float(torch.linalg.vector_norm(interior_flat(b, S.cfg)))
        x_prev = hist[-1]
        x_ext = extrapolate(hist, alpha)
        _, r_prev = S.residual(b, x_prev)
        r_ext_v, r_ext = S.residual(b, x_ext)
        xa, sa = S.solve(b, x_prev)
        xb, sb = S.solve(b, x_ext)
        rec = {"step": n, "r0_prev": r_prev / bn, "r0_extrap": r_ext / bn,
               "it_prev": sa["it"], "it_extrap": sb["it"],
               "ms_prev": sa["solve_ms"], "ms_extrap": sb["solve_ms"],
               "true_prev": sa["true_rerr"], "true_extrap": sb["true_rerr"]}
        ref = _mean_free(xa)
        nref = float(torch.linalg.vector_norm(ref))
        diffs = [float(torch.linalg.vector_norm(_mean_free(xb) - ref)) / nref]
        if net is not None:
            xc, sc = S.solve(b, x_ext, stage=stage, flags=flags, precond_iters=precond_iters)
            _, r_learn = S.residual(b, x_ext + net.apply(r_ext_v))
            rec.update({"r0_learned": r_learn / bn, "it_learned": sc["it"], "ms_learned": sc["solve_ms"],
                        "true_learned": sc["true_rerr"], "stage_calls": sc["stage_calls"],
                        "rejected": int(bool(sc["fallback"] & 2))})
            diffs.append(float(torch.linalg.vector_norm(_mean_free(xc) - ref)) / nref)
        rec["diff"] = max(diffs)
        out.append(rec)
        hist = hist[1:] + [_mean_free(xa)]
    return out


def summary(recs):
    keys = [k for k in recs[0] if k.startswith(("it_", "ms_", "r0_"))]
    s = {"steps": len(recs)}
    for k in keys:
        s[k] = float(np.mean([r[k] for r in recs]))
    for k in [k for k in recs[0] if k.startswith("true_")]:
        s["max_" + k] = max(r[k] for r in recs)
    s["max_solution_diff"] = max(r["diff"] for r in recs)
    if "rejected" in recs[0]:
        s["guess_rejected"] = int(sum(r["rejected"] for r in recs))
    return s


def run(n=128, steps=200, train_steps=150, m=6, epochs=6, lr=5e-4, seed=0, precond_iters=0, log=print):
    """The whole workflow; returns a dict of timings and summaries."""
    t0 = time.perf_counter()
    S = SequenceSolver(n=n)
    sols, st_train = collect(S, range(train_steps))
    t_collect = time.perf_counter() - t0
    t1 = time.perf_counter()
    alpha = fit_predictor(sols, m)
    pairs = stage_pairs(S, sols, alpha, range(m + 1, train_steps))
    kappa = fit_kappa(pairs)
    net = UNet((n, n, n), seed=seed)
    hist = train(net, pairs, kappa, S=S, epochs=epochs, lr=lr, log=log)
    t_train = time.perf_counter() - t1
    del pairs
    recs = evaluate(S, net, alpha, sols, range(train_steps, steps), precond_iters=precond_iters)
    res = {"alpha": [float(a) for a in alpha], "kappa": kappa, "loss_history": hist, "collect_s": t_collect,
           "train_s": t_train, "train_solves_it": float(np.mean([s["it"] for s in st_train])),
           "eval": summary(recs)}
    net.close()
    S.close()
    return res


# ------------------------------------------------------------------ variable-density operators
# For a variable-coefficient operator the network is given the Jacobi-scaled
# residual D^-1 r (D = diag A, the input scale of UNet.set_input_scale). Inside
# one phase of a two-phase operator D^-1 A is close to (h^2/6) times the
# constant-coefficient Laplacian whatever the density, so one set of weights
# serves every density field and the network need not see it separately. The
# network is trained on crops (a few m^3 windows of full-size solves) with the
# Jacobi-scaled residual the correction leaves as the loss.

def gaussian_rhs_device(proj, cfg, seed, out, faces=None, div=None):
    """out (ghosted, Fortran order) <- a right-hand side by the configuration-1
    recipe: U, V, W ~ N(0, 1) i.i.d. on the interior faces (a torch generator
    on the device, seeded by `seed`), per-component mean removed, ghosts
    synced, div by the device projection (ngap_proj_div), rhs = -vol div / dt
    with dt = 1 (ngap_proj_rhs), mean removed over the interior.
    inputs._gaussian_faces is the same recipe drawn with numpy."""
    import torch

    dev = out.device
    No = cfg.nxo * cfg.nyo * cfg.nzo
    N = cfg.nx * cfg.ny * cfg.nz
    faces = faces if faces is not None else [torch.zeros(No, dtype=torch.float64, device=dev) for _ in range(3)]
    div = div if div is not None else torch.zeros(No, dtype=torch.float64, device=dev)
    g = torch.Generator(device=dev)
    g.manual_seed(int(seed))
    no = cfg.no
    for F in faces:
        v = torch.randn(N, generator=g, dtype=torch.float64, device=dev)
        v -= v.mean()
        F.zero_()
        F.view(cfg.nzo, cfg.nyo, cfg.nxo)[no:no + cfg.nz, no:no + cfg.ny, no:no + cfg.nx] = \
            v.view(cfg.nz, cfg.ny, cfg.nx)
        proj.sync(F)
    proj.div(*faces, div)
    proj.rhs(div, 1.0, 1.0, out)
    torch.cuda.synchronize()
    ri = out.view(cfg.nzo, cfg.nyo, cfg.nxo)[no:no + cfg.nz, no:no + cfg.ny, no:no + cfg.nx]
    ri -= ri.mean()
    return out


def _interior_opr(opr, cfg):
    no = cfg.no
    return opr.view(cfg.nzo, cfg.nyo, cfg.nxo, 7)[no:no + cfg.nz, no:no + cfg.ny, no:no + cfg.nx, :]


def diag_inverse(opr, cfg, out):
    """out (interior, x fastest) <- 1 / opr(1) of a ghosted Fortran-ordered
    operator (0 on identity rows, opr(1) == 0)."""
    import torch

    d = _interior_opr(opr, cfg)[..., 0].reshape(-1)
    torch.where(d != 0, 1.0 / torch.where(d != 0, d, torch.ones_like(d)), torch.zeros_like(d), out=out)
    return out


def _crop_index(cfg, off, m, dev):
    import torch

    return [(torch.arange(m, device=dev) + int(q)) % n for q, n in zip(off, (cfg.nx, cfg.ny, cfg.nz))]


def crop_field(a, cfg, off, m):
    """m^3 periodic window of an interior field (x fastest) at offset (i, j, k),
    as an (m, m, m) tensor indexed [k, j, i]."""
    ix = _crop_index(cfg, off, m, a.device)
    a3 = a.view(cfg.nz, cfg.ny, cfg.nx)
    return a3.index_select(0, ix[2]).index_select(1, ix[1]).index_select(2, ix[0])


def crop_operator(opr, cfg, off, m):
    """The 7 stencil entries of a ghosted Fortran-ordered operator over an m^3
    window, (7, m, m, m) indexed [s, k, j, i] in the stencil order
    (0, +x, -x, +y, -y, +z, -z)."""
    ix = _crop_index(cfg, off, m, opr.device)
    o = _interior_opr(opr, cfg)
    o = o.index_select(0, ix[2]).index_select(1, ix[1]).index_select(2, ix[0])
    return o.permute(3, 0, 1, 2).contiguous()


def crop_apply(o, y):
    """A_c y for the operator restricted to a window (its principal submatrix:
    couplings to cells outside the window dropped); o (7, m, m, m) from
    crop_operator, y (m, m, m) indexed [k, j, i]."""
    out = o[0] * y
    out[:, :, :-1] += o[1][:, :, :-1] * y[:, :, 1:]
    out[:, :, 1:] += o[2][:, :, 1:] * y[:, :, :-1]
    out[:, :-1, :] += o[3][:, :-1, :] * y[:, 1:, :]
    out[:, 1:, :] += o[4][:, 1:, :] * y[:, :-1, :]
    out[:-1] += o[5][:-1] * y[1:]
    out[1:] += o[6][1:] * y[:-1]
    return out


def crop_loss_grad(w, o, c, y, kappa, mask, M=None):
    """(L, dL/dy) of the residual a correction leaves on a window, in units of
    the window's rms residual t: with the network input u = D^-1 r / s and
    output y, the correction is e = s kappa y and
        (r - A e) / t = w - c kappa A_c y,   w = r / t,  c = s / t,
    L = sum(mask (w - c kappa A_c y)^2) / M (M = sum(mask) by default) and
    dL/dy = -2 c kappa / M  A_c (mask res), since A_c is symmetric. This is
    the 2-norm residual by which the solver accepts a guess."""
    if M is None:
        M = float(mask.sum())
    res = mask * (w - (c * kappa) * crop_apply(o, y))
    return float((res * res).sum()) / M, (-2.0 * c * kappa / M) * crop_apply(o, res)


def crop_samples(opr, cfg, b, x, dinv, m, count, gen):
    """`count` training windows of one solve: b the right-hand side and x the
    (mean-free) solution, both interior and x fastest, dinv = 1 / diag A.
    Returns ([(u, o, w, c)], [kappa estimates]): on each window u = D^-1 b / s
    (s its rms, rounded through FP32 as the network sees it), o the window's
    operator, w = b / t (t = rms of b on the window), c = s / t, and the
    ratio rms(x) / s."""
    import torch

    samples, kappas = [], []
    for _ in range(count):
        off = gen.integers(0, (cfg.nx, cfg.ny, cfg.nz), 3)
        o = crop_operator(opr, cfg, off, m)
        db = crop_field(dinv * b, cfg, off, m)
        bw = crop_field(b, cfg, off, m)
        s = float(torch.sqrt((db * db).mean()))
        t = float(torch.sqrt((bw * bw).mean()))
        p = crop_field(x, cfg, off, m)
        kappas.append(float(torch.sqrt((p * p).mean())) / s)
        samples.append(((db / s).to(torch.float32).to(torch.float64), o, bw / t, s / t))
    return samples, kappas


def train_on_crops(net, samples, kappa, border=16, epochs=6, lr=5e-4, lr_end=5e-5, seed=0, log=None):
    """Adam over window samples (u, o, w, c) of crop_samples. The loss is the
    residual the correction leaves (crop_loss_grad), mean over the window
    without a border of `border` cells, where the window's edge and the
    network's periodic padding differ from the full grid. Returns (per-epoch
    mean losses, best epoch); the net is left holding the weights of the
    epoch with the lowest mean loss."""
    import torch

    net.set_kappa(kappa)
    m = samples[0][0].shape[0]
    mask = torch.zeros((m, m, m), dtype=torch.float64, device=samples[0][0].device)
    mask[border:m - border, border:m - border, border:m - border] = 1.0
    M = float(mask.sum())
    rng = np.random.default_rng(seed)
    total = epochs * len(samples)
    k = 0
    hist = []
    best = (np.inf, None, -1)
    for ep in range(epochs):
        losses = []
        for i in rng.permutation(len(samples)):
            u, o, w, c = samples[i]
            a = lr_end + 0.5 * (lr - lr_end) * (1.0 + np.cos(np.pi * k / max(total - 1, 1)))
            y = net.forward(u.reshape(-1)).to(torch.float64).view(m, m, m)
            loss, g = crop_loss_grad(w, o, c, y, kappa, mask, M)
            losses.append(loss)
            net.backward(g.reshape(-1).to(torch.float32))
            net.adam_step(float(a))
            k += 1
        hist.append(float(np.mean(losses)))
        if log:
            log(f"U-Net epoch {ep}: loss {hist[-1]:.4e}")
        if hist[-1] < best[0]:
            best = (hist[-1], net.state(), ep)
    net.load_state(best[1])
    return hist, best[2]


def train_for_operator(ds, proj, opr, cfg, seeds, crop=128, crops_per_solve=8, epochs=6, lr=5e-4, log=None,
                       device="cuda:0"):
    """Train the network for one operator from its own solves: for every seed
    a configuration-1 right-hand side (gaussian_rhs_device) is solved by ds
    (set up on opr) from a zero guess, and crops of (rhs, solution) become
    training windows (crop_samples, train_on_crops). Returns
    (state, info): the best epoch's weights with kappa, and the training
    record. The net is trained on crop^3 windows; load the state into a
    full-size UNet to apply it."""
    import torch

    from .unet import UNet

    No = cfg.nxo * cfg.nyo * cfg.nzo
    N = cfg.nx * cfg.ny * cfg.nz
    dev = opr.device
    rhs = torch.zeros(No, dtype=torch.float64, device=dev)
    x = torch.zeros(No, dtype=torch.float64, device=dev)
    dinv = diag_inverse(opr, cfg, torch.zeros(N, dtype=torch.float64, device=dev))
    gen = np.random.default_rng(91)
    samples, kappas, its = [], [], []
    t0 = time.perf_counter()
    for sd in seeds:
        gaussian_rhs_device(proj, cfg, sd, rhs)
        x.zero_()
        it, _ = ds.solve(rhs, x)
        its.append(it)
        smp, kap = crop_samples(opr, cfg, interior_flat(rhs, cfg), _mean_free(interior_flat(x, cfg)), dinv, crop,
                                crops_per_solve, gen)
        samples += smp
        kappas += kap
    t1 = time.perf_counter()
    kappa = float(np.median(kappas))
    net = UNet((crop, crop, crop), seed=0, device=device)
    try:
        hist, best = train_on_crops(net, samples, kappa, epochs=epochs, lr=lr, log=log)
        state = net.state()
    finally:
        net.close()
    info = {"solves": len(seeds), "solve_iters": its, "windows": len(samples), "crop": crop, "kappa": kappa,
            "loss_history": hist, "best_epoch": best, "collect_s": t1 - t0, "train_s": time.perf_counter() - t1}
    return state, info
