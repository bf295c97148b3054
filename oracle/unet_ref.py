"""CPU oracle of the U-Net (torch, float64) — test infrastructure only.

It restates nga2_amd/unet.py's architecture with the same weight layouts:

- 3x3x3 weights are [27][cout][cin], tap = (dz+1)*9 + (dy+1)*3 + (dx+1);
- the 1x1 output is [cin];
- activations are channels-last [voxel][channel] with x fastest;
- padding is periodic.

With quantize=True the weights of the MFMA layers and every post-activation are
rounded to bfloat16 where the GPU path stores them, so the forward pass differs
from the GPU only by FP32 accumulation order.

Pinned in tests against direct loops on tiny inputs (tap convention, periodic
padding, pooling, upsampling).
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F

LAYERS = [
    ("c0", 0, 1, 32, "small"), ("c1", 1, 32, 32, "mfma"), ("c2a", 2, 32, 64, "mfma"), ("c2", 2, 64, 64, "mfma"),
    ("b1", 3, 64, 64, "mfma"), ("b", 3, 64, 64, "mfma"), ("d2", 2, 128, 64, "mfma"), ("d1", 1, 96, 32, "mfma"),
    ("d0", 0, 64, 32, "mfma"), ("out", 0, 32, 1, "1x1"),
]


def to_torch(a_vc, shape):
    """[V][C] (x fastest) -> [1, C, nz, ny, nx]"""
    nx, ny, nz = shape
    a = a_vc if torch.is_tensor(a_vc) else torch.as_tensor(np.asarray(a_vc), dtype=torch.float64)
    c = a.numel() // (nx * ny * nz)
    return a.reshape(nz, ny, nx, c).permute(3, 0, 1, 2).unsqueeze(0)


def from_torch(t):
    """[1, C, nz, ny, nx] -> [V][C]"""
    return t[0].permute(1, 2, 3, 0).reshape(-1, t.shape[1])


def weight3(w_flat, cin, cout):
    """[27][cout][cin] -> torch [cout][cin][kz][ky][kx]"""
    w = w_flat if torch.is_tensor(w_flat) else torch.as_tensor(np.asarray(w_flat), dtype=torch.float64)
    w = w.reshape(3, 3, 3, cout, cin)
    return w.permute(3, 4, 0, 1, 2).contiguous()


def conv3(x, w):
    return F.conv3d(F.pad(x, (1, 1, 1, 1, 1, 1), mode="circular"), w)


def _bf(t, on):
    """Round to bfloat16 in the forward pass; the gradient passes through
    unrounded (straight-through), so the oracle's gradients stay float64."""
    if not on:
        return t
    return t + (t.detach().to(torch.bfloat16).to(torch.float64) - t.detach())


class _GradToBF16(torch.autograd.Function):
    """Identity forward; the incoming gradient is rounded to bfloat16."""

    @staticmethod
    def forward(ctx, t):
        return t.view_as(t)

    @staticmethod
    def backward(ctx, g):
        return g.to(torch.bfloat16).to(torch.float64)


def forward(weights, r, shape, quantize=True):
    """weights: dict name -> flat array (torch tensors with requires_grad for
    autograd, or numpy). r: [V] input. Returns y [V] (float64 tensor).

    quantize=True rounds to bfloat16 where the GPU stores BF16:
    - in the forward pass, the MFMA weights and every activation;
    - in the backward pass, the gradient at the output of each MFMA
      convolution, which is the operand of its weight and data gradients."""
    W = {k: (v if torch.is_tensor(v) else torch.as_tensor(np.asarray(v), dtype=torch.float64))
         for k, v in weights.items()}
    q = quantize
    lw = {name: (weight3(_bf(W[name], q and kind == "mfma"), cin, cout) if kind != "1x1" else W[name])
          for name, lvl, cin, cout, kind in LAYERS}

    def cv(t, name):
        z = conv3(t, lw[name])
        return _GradToBF16.apply(z) if q and name != "c0" else z

    x = to_torch(r, shape)
    c0 = _bf(F.relu(cv(x, "c0")), q)
    c1 = _bf(F.relu(cv(_bf(F.avg_pool3d(c0, 2), q), "c1")), q)
    p2 = _bf(F.avg_pool3d(c1, 2), q)
    c2a = _bf(F.relu(cv(p2, "c2a")), q)
    c2 = _bf(F.relu(cv(c2a, "c2")), q)
    p3 = _bf(F.avg_pool3d(c2, 2), q)
    b1 = _bf(F.relu(cv(p3, "b1")), q)
    b = _bf(F.relu(cv(b1, "b")), q)
    up = lambda t: F.interpolate(t, scale_factor=2, mode="nearest")  # noqa: E731
    d2 = _bf(F.relu(cv(torch.cat([up(b), c2], dim=1), "d2")), q)
    d1 = _bf(F.relu(cv(torch.cat([up(d2), c1], dim=1), "d1")), q)
    d0 = _bf(F.relu(cv(torch.cat([up(d1), c0], dim=1), "d0")), q)
    y = torch.einsum("bczyx,c->bzyx", d0, lw["out"])
    return y.reshape(-1)


def gradients(weights, r, dy, shape, quantize=True):
    """d(sum(dy * y))/dW for every layer (float64 numpy, flat)."""
    W = {k: torch.tensor(np.asarray(v, dtype=np.float64), requires_grad=True) for k, v in weights.items()}
    y = forward(W, r, shape, quantize=quantize)
    (y * torch.as_tensor(np.asarray(dy), dtype=torch.float64)).sum().backward()
    return {k: v.grad.numpy().copy() for k, v in W.items()}, y.detach().numpy()


def conv3_direct(x_vc, w_flat, shape, cin, cout):
    """Direct loops: out[v][co] = sum_tap sum_ci w[tap][co][ci] x[v+off][ci] (periodic)."""
    nx, ny, nz = shape
    x = np.asarray(x_vc, dtype=np.float64).reshape(nz, ny, nx, cin)
    w = np.asarray(w_flat, dtype=np.float64).reshape(27, cout, cin)
    out = np.zeros((nz, ny, nx, cout))
    for tap in range(27):
        dx, dy, dz = tap % 3 - 1, (tap // 3) % 3 - 1, tap // 9 - 1
        xs = np.roll(x, shift=(-dz, -dy, -dx), axis=(0, 1, 2))
        out += xs @ w[tap].T
    return out.reshape(-1, cout)
