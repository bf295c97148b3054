"""GPU parity of the U-Net kernels (libngapnet.so) against the float64 CPU
oracle (oracle/unet_ref.py) and direct numpy formulas.

Inputs of the BF16 kernels are rounded to bfloat16 first, and the oracle
computes in float64 on the rounded values. The only difference left is the
FP32 accumulation order of the MFMA path. Tolerances:

- single conv forward / data gradient: 2e-6 of max|ref|;
- weight gradient (sums over up to 2^18 voxels): 2e-5 of max|ref|;
- whole net: the oracle rounds to BF16 at the same points as the GPU. This
  covers the weights and activations in the forward pass, and the gradient at
  each MFMA convolution's output in the backward pass. A BF16 value still
  differs by one ulp wherever the two accumulation orders straddle a rounding
  boundary, and a ReLU mask flips where a pre-activation is near zero.
  - The forward output is held to relative L2 <= 1e-2.
  - Each layer's gradient is held to the oracle's own sensitivity. The test
    perturbs the input by 1e-7 relative (FP32 noise) and measures how far the
    oracle's own BF16-rounded gradient moves: 3-12% per layer for a randomly
    initialised net. The GPU gradient must lie within 2x that distance plus
    0.5%, and at cosine >= 0.97. A wiring mistake (a dropped skip path, a
    wrong channel slice) moves a gradient by O(1).
"""
import numpy as np
import pytest
import torch

from oracle import unet_ref as U

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _lib():
    from nga2_amd import unet

    return unet, unet.lib()


def _bf16(a):
    """numpy -> (bf16 device tensor, float64 numpy of the rounded values)"""
    t = torch.as_tensor(np.asarray(a, dtype=np.float32)).to(torch.bfloat16)
    return t.to(DEV), t.to(torch.float64).numpy()


def _f32(a):
    t = torch.as_tensor(np.asarray(a, dtype=np.float32))
    return t.to(DEV), t.to(torch.float64).numpy()


def _np(t):
    return t.detach().to("cpu", torch.float64).numpy()


def _run(fn):
    unet, L = _lib()
    unet._enter()
    rc = fn(L, unet._p)
    unet._leave()
    assert rc == 0, rc


def _wgrad_direct(x, gy, shape, cin, cout):
    nx, ny, nz = shape
    X = x.reshape(nz, ny, nx, cin)
    dw = np.zeros((27, cout, cin))
    for tap in range(27):
        dx, dy, dz = tap % 3 - 1, (tap // 3) % 3 - 1, tap // 9 - 1
        xs = np.roll(X, shift=(-dz, -dy, -dx), axis=(0, 1, 2)).reshape(-1, cin)
        dw[tap] = gy.T @ xs
    return dw.reshape(-1)


CONV_CASES = [((16, 4, 4), 32, 32), ((32, 8, 4), 64, 32), ((16, 8, 8), 96, 64), ((32, 4, 8), 128, 64)]


@pytest.mark.parametrize("shape,cin,cout", CONV_CASES)
def test_conv3_forward_mfma(shape, cin, cout):
    print(f"\n[conv3 fwd {shape} {cin}->{cout}]", flush=True)
    rng = np.random.default_rng(10 + cin)
    V = shape[0] * shape[1] * shape[2]
    x, xr = _bf16(rng.standard_normal((V, cin)))
    w, wr = _bf16(rng.standard_normal(27 * cout * cin) / np.sqrt(27 * cin))
    out = torch.zeros(V * cout, dtype=torch.float32, device=DEV)
    outb = torch.zeros(V * cout, dtype=torch.bfloat16, device=DEV)
    _run(lambda L, p: L.ngapnet_conv3_fwd(p(x), p(w), p(out), p(outb), *shape, cin, cout, 1, None))
    ref = U.conv3_direct(xr, wr, shape, cin, cout).reshape(-1)
    got = _np(out)
    assert np.abs(got - ref).max() <= 2e-6 * np.abs(ref).max()
    # the BF16 ReLU output is the round-to-nearest-even of the FP32 value
    assert torch.equal(outb, torch.relu(out).to(torch.bfloat16))


def test_conv3_rejects_untiled_shapes():
    unet, L = _lib()
    assert L.ngapnet_conv3_fwd(None, None, None, None, 16, 4, 4, 16, 32, 0, None) == -1
    assert L.ngapnet_conv3_fwd(None, None, None, None, 8, 4, 4, 32, 32, 0, None) == -1
    assert L.ngapnet_conv3_wgrad(None, None, None, None, 16, 6, 4, 32, 32, None) == -1


@pytest.mark.parametrize("shape,cin,cout", CONV_CASES[:3])
def test_conv3_data_gradient(shape, cin, cout):
    """dx = conv3(gy, flip-transposed W) equals autograd of the oracle conv."""
    print(f"\n[conv3 dgrad {shape} {cin}->{cout}]", flush=True)
    rng = np.random.default_rng(20 + cin)
    V = shape[0] * shape[1] * shape[2]
    w, wr = _bf16(rng.standard_normal(27 * cout * cin) / np.sqrt(27 * cin))
    gy, gyr = _bf16(rng.standard_normal((V, cout)))
    wd = torch.zeros_like(w)
    dx = torch.zeros(V * cin, dtype=torch.float32, device=DEV)
    _run(lambda L, p: L.ngapnet_flip_transpose(p(w), p(wd), cin, cout, None))
    wdn = _np(wd).reshape(27, cin, cout)
    assert np.array_equal(wdn, wr.reshape(27, cout, cin)[::-1].transpose(0, 2, 1))
    _run(lambda L, p: L.ngapnet_conv3_fwd(p(gy), p(wd), p(dx), None, *shape, cout, cin, 0, None))
    xt = torch.zeros((V, cin), dtype=torch.float64, requires_grad=True)
    y = U.conv3(U.to_torch(xt, shape), U.weight3(wr, cin, cout))
    (U.from_torch(y) * torch.as_tensor(gyr)).sum().backward()
    ref = xt.grad.numpy().reshape(-1)
    got = _np(dx)
    assert np.abs(got - ref).max() <= 2e-6 * np.abs(ref).max()


@pytest.mark.parametrize("shape,cin,cout", [((16, 4, 4), 32, 32), ((64, 32, 32), 32, 64), ((64, 32, 64), 64, 32),
                                            ((32, 8, 8), 128, 64)])
def test_conv3_weight_gradient(shape, cin, cout):
    print(f"\n[conv3 wgrad {shape} {cin}->{cout}]", flush=True)
    unet, L = _lib()
    rng = np.random.default_rng(30 + cin)
    V = shape[0] * shape[1] * shape[2]
    x, xr = _bf16(rng.standard_normal((V, cin)))
    gy, gyr = _bf16(rng.standard_normal((V, cout)))
    G = L.ngapnet_wgrad_parts(*shape, cin, cout)
    part = torch.zeros(G * 27 * cin * cout, dtype=torch.float32, device=DEV)
    dw = torch.zeros(27 * cin * cout, dtype=torch.float32, device=DEV)
    _run(lambda L, p: L.ngapnet_conv3_wgrad(p(x), p(gy), p(dw), p(part), *shape, cin, cout, None))
    ref = _wgrad_direct(xr, gyr, shape, cin, cout)
    got = _np(dw)
    assert np.abs(got - ref).max() <= 2e-5 * np.abs(ref).max()


@pytest.mark.parametrize("shape", [(16, 8, 8), (18, 6, 5)])
def test_small_conv_forward_and_weight_gradient(shape):
    print(f"\n[small conv {shape}]", flush=True)
    rng = np.random.default_rng(40)
    V = shape[0] * shape[1] * shape[2]
    cin, cout = 1, 32
    x, xr = _f32(rng.standard_normal((V, cin)))
    w, wr = _f32(rng.standard_normal(27 * cout * cin) / np.sqrt(27))
    out = torch.zeros(V * cout, dtype=torch.float32, device=DEV)
    outb = torch.zeros(V * cout, dtype=torch.bfloat16, device=DEV)
    _run(lambda L, p: L.ngapnet_conv3_small_fwd(p(x), p(w), p(out), p(outb), *shape, cin, cout, 1, None))
    ref = U.conv3_direct(xr, wr, shape, cin, cout).reshape(-1)
    assert np.abs(_np(out) - ref).max() <= 1e-5 * np.abs(ref).max()
    assert torch.equal(outb, torch.relu(out).to(torch.bfloat16))
    gy, gyr = _f32(rng.standard_normal((V, cout)))
    ref = _wgrad_direct(xr, gyr, shape, cin, cout)
    unet, L = _lib()
    part = torch.zeros(int(L.ngapnet_small_wgrad_part_size(cout)), dtype=torch.float32, device=DEV)
    for scratch in (None, part):  # one workgroup per weight / one pass per tap row
        dw = torch.zeros(27 * cout * cin, dtype=torch.float32, device=DEV)
        _run(lambda L, p: L.ngapnet_conv3_small_wgrad(p(x), p(gy), p(dw), p(scratch), *shape, cin, cout, None))
        assert np.abs(_np(dw) - ref).max() <= 1e-5 * np.abs(ref).max()


def test_pool_upsample_concat_relu_kernels():
    print("\n[pool/up/concat/relu]", flush=True)
    rng = np.random.default_rng(50)
    nx, ny, nz, c = 8, 6, 4, 32
    V, Vc = nx * ny * nz, nx * ny * nz // 8
    a, ar = _bf16(rng.standard_normal((V, c)))
    pooled = torch.zeros(Vc * c, dtype=torch.bfloat16, device=DEV)
    _run(lambda L, p: L.ngapnet_pool2(p(a), p(pooled), nx, ny, nz, c, None))
    ref = ar.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, c).mean(axis=(1, 3, 5)).reshape(-1)
    assert np.array_equal(_np(pooled), torch.as_tensor(ref).to(torch.bfloat16).double().numpy())
    # pool backward (with accumulation)
    g, gr = _f32(rng.standard_normal((Vc, c)))
    din = torch.ones(V * c, dtype=torch.float32, device=DEV)
    _run(lambda L, p: L.ngapnet_pool2_bwd(p(g), p(din), nx, ny, nz, c, 1, None))
    t = gr.reshape(nz // 2, ny // 2, nx // 2, c).repeat(2, 0).repeat(2, 1).repeat(2, 2).reshape(-1) / 8 + 1
    assert np.allclose(_np(din), t, rtol=1e-6, atol=1e-7)
    # upsample into channels [16, 48) of a 64-channel buffer, and its adjoint
    cat = torch.zeros(V * 64, dtype=torch.bfloat16, device=DEV)
    cp, cpr = _bf16(rng.standard_normal((Vc, c)))
    _run(lambda L, p: L.ngapnet_up2(p(cp), p(cat), nx, ny, nz, c, 64, 16, None))
    catn = _np(cat).reshape(V, 64)
    up = cpr.reshape(nz // 2, ny // 2, nx // 2, c).repeat(2, 0).repeat(2, 1).repeat(2, 2).reshape(V, c)
    assert np.array_equal(catn[:, 16:48], up) and not catn[:, :16].any() and not catn[:, 48:].any()
    gc, gcr = _f32(rng.standard_normal((V, 64)))
    dc = torch.zeros(Vc * c, dtype=torch.float32, device=DEV)
    _run(lambda L, p: L.ngapnet_up2_bwd(p(gc), p(dc), nx, ny, nz, c, 64, 16, None))
    ref = gcr[:, 16:48].reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, c).sum(axis=(1, 3, 5)).reshape(-1)
    assert np.allclose(_np(dc), ref, rtol=1e-6, atol=1e-6)
    # channel slice copies
    dst = torch.zeros(V * 48, dtype=torch.float32, device=DEV)
    _run(lambda L, p: L.ngapnet_chan_copy_f32(p(gc), p(dst), V, 20, 64, 40, 48, 8, None))
    dn = _np(dst).reshape(V, 48)
    assert np.array_equal(dn[:, 8:28], gcr[:, 40:60]) and not dn[:, :8].any() and not dn[:, 28:].any()
    dstb = torch.zeros(V * 64, dtype=torch.bfloat16, device=DEV)
    _run(lambda L, p: L.ngapnet_chan_copy_bf16(p(a), p(dstb), V, 32, 32, 0, 64, 32, None))
    assert np.array_equal(_np(dstb).reshape(V, 64)[:, 32:], ar)
    # ReLU backward through the BF16 activation, and FP32 -> BF16
    y, yr = _bf16(rng.standard_normal(V * c))
    dy, dyr = _f32(rng.standard_normal(V * c))
    dx = torch.zeros(V * c, dtype=torch.float32, device=DEV)
    dxb = torch.zeros(V * c, dtype=torch.bfloat16, device=DEV)
    _run(lambda L, p: L.ngapnet_relu_bwd(p(dy), p(y), p(dx), p(dxb), V * c, None))
    assert np.array_equal(_np(dx), np.where(yr > 0, dyr, 0.0))
    assert torch.equal(dxb, dx.to(torch.bfloat16))
    b = torch.zeros(V * c, dtype=torch.bfloat16, device=DEV)
    _run(lambda L, p: L.ngapnet_to_bf16(p(dy), p(b), V * c, None))
    assert torch.equal(b, dy.to(torch.bfloat16))


def test_output_layer_forward_backward():
    print("\n[1x1 output layer]", flush=True)
    rng = np.random.default_rng(60)
    V, c = 40000, 32
    x, xr = _bf16(rng.standard_normal((V, c)))
    w, wr = _f32(rng.standard_normal(c))
    dy, dyr = _f32(rng.standard_normal(V))
    y = torch.zeros(V, dtype=torch.float32, device=DEV)
    dx = torch.zeros(V * c, dtype=torch.float32, device=DEV)
    dw = torch.zeros(c, dtype=torch.float32, device=DEV)
    part = torch.zeros(1024 * 64, dtype=torch.float32, device=DEV)
    _run(lambda L, p: L.ngapnet_conv1_fwd(p(x), p(w), p(y), V, c, None))
    assert np.allclose(_np(y), xr @ wr, rtol=1e-5, atol=1e-5)
    _run(lambda L, p: L.ngapnet_conv1_bwd(p(x), p(w), p(dy), p(dx), p(dw), p(part), V, c, None))
    assert np.allclose(_np(dx), np.outer(dyr, wr).reshape(-1), rtol=1e-6, atol=1e-7)
    ref = xr.T @ dyr
    assert np.abs(_np(dw) - ref).max() <= 1e-5 * np.abs(ref).max()


def test_adam_matches_torch_optimizer():
    print("\n[adam]", flush=True)
    rng = np.random.default_rng(70)
    n = 5000
    w0 = rng.standard_normal(n).astype(np.float32)
    w = torch.tensor(w0, device=DEV)
    m = torch.zeros(n, device=DEV)
    v = torch.zeros(n, device=DEV)
    wb = torch.zeros(n, dtype=torch.bfloat16, device=DEV)
    wt = torch.tensor(w0, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([wt], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    for step in range(1, 4):
        gn = rng.standard_normal(n).astype(np.float32)
        g = torch.tensor(gn, device=DEV)
        _run(lambda L, p: L.ngapnet_adam(p(w), p(g), p(m), p(v), p(wb), n, 1e-2, 0.9, 0.999, 1e-8, step, 1.0, None))
        wt.grad = torch.tensor(gn, dtype=torch.float64)
        opt.step()
    assert np.allclose(_np(w), wt.detach().numpy(), rtol=1e-5, atol=1e-6)
    assert torch.equal(wb, w.to(torch.bfloat16))


NET_SHAPE = (128, 32, 32)


def test_unet_forward_and_gradients_match_oracle():
    print(f"\n[unet {NET_SHAPE} forward/backward]", flush=True)
    from nga2_amd.unet import UNet

    net = UNet(NET_SHAPE, seed=3)
    rng = np.random.default_rng(80)
    V = NET_SHAPE[0] * NET_SHAPE[1] * NET_SHAPE[2]
    r = rng.standard_normal(V)
    dy = rng.standard_normal(V) / np.sqrt(V)
    rt = torch.tensor(r, dtype=torch.float32, device=DEV)
    y = _np(net.forward(rt))
    net.backward(torch.tensor(dy, dtype=torch.float32, device=DEV))
    Gg = net.grads()
    W = {k: v.astype(np.float64) for k, v in net.weights().items()}
    net.close()
    r32 = r.astype(np.float32).astype(np.float64)
    G, yref = U.gradients(W, r32, dy, NET_SHAPE, quantize=True)
    Gp, _ = U.gradients(W, r32 * (1.0 + 1e-7 * rng.standard_normal(V)), dy, NET_SHAPE, quantize=True)
    ey = np.linalg.norm(y - yref) / np.linalg.norm(yref)
    print(f"  forward rel L2 {ey:.2e}", flush=True)
    assert ey <= 1e-2
    for name in G:
        nrm = np.linalg.norm(G[name])
        e = np.linalg.norm(Gg[name] - G[name]) / nrm
        sens = np.linalg.norm(Gp[name] - G[name]) / nrm
        cos = float(Gg[name] @ G[name]) / (np.linalg.norm(Gg[name]) * nrm)
        print(f"  grad {name:4s} rel L2 {e:.2e} (oracle sensitivity {sens:.2e}) cos {cos:.4f}", flush=True)
        assert e <= 2.0 * sens + 5e-3, name
        assert cos >= 0.97, name


def test_unet_stage_is_scale_free_and_weights_round_trip():
    print(f"\n[unet {NET_SHAPE} stage]", flush=True)
    from nga2_amd.unet import UNet, init_weights

    net = UNet(NET_SHAPE, seed=5)
    W0 = init_weights(5)
    W1 = net.weights()
    assert all(np.array_equal(W0[k], W1[k]) for k in W0)
    rng = np.random.default_rng(81)
    V = NET_SHAPE[0] * NET_SHAPE[1] * NET_SHAPE[2]
    r = torch.tensor(rng.standard_normal(V) * 3e-4, dtype=torch.float64, device=DEV)
    net.set_kappa(2.5)
    e = _np(net.apply(r))
    s = float(torch.sqrt((r * r).mean()))
    y = _np(net.forward((r / s).to(torch.float32)))
    assert np.allclose(e, s * 2.5 * y, rtol=1e-6, atol=1e-12 * s)
    # a scaled residual gives the scaled correction (up to FP32 rounding of r / s)
    e2 = _np(net.apply(r * 1e6))
    assert np.allclose(e2, 1e6 * e, rtol=1e-5, atol=1e-5 * np.abs(1e6 * e).max())
    assert not _np(net.apply(torch.zeros(V, dtype=torch.float64, device=DEV))).any()
    with pytest.raises(ValueError):
        UNet((64, 32, 32))
    net.close()


def _periodic_poisson(r, shape):
    """e with -lap(e) = r - mean(r) on the periodic unit-spacing grid (FFT)."""
    nx, ny, nz = shape
    R = np.fft.fftn(r.reshape(nz, ny, nx))
    kz, ky, kx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    lam = (2 - 2 * np.cos(2 * np.pi * kx / nx)) + (2 - 2 * np.cos(2 * np.pi * ky / ny)) + \
        (2 - 2 * np.cos(2 * np.pi * kz / nz))
    lam[0, 0, 0] = 1.0
    E = R / lam
    E[0, 0, 0] = 0.0
    return np.real(np.fft.ifftn(E)).reshape(-1)


def test_unet_training_reduces_loss():
    """A few Adam steps on (r, A^-1 r) pairs of the periodic Laplacian reduce
    the MSE on the training pair and on a held-out pair."""
    print(f"\n[unet {NET_SHAPE} training]", flush=True)
    from nga2_amd.unet import UNet

    net = UNet(NET_SHAPE, seed=4)
    rng = np.random.default_rng(90)
    V = NET_SHAPE[0] * NET_SHAPE[1] * NET_SHAPE[2]

    def pair():
        r = _periodic_poisson(rng.standard_normal(V), NET_SHAPE)  # smooth input
        r = r / np.sqrt(np.mean(r * r))
        e = _periodic_poisson(r, NET_SHAPE)
        e = e / np.sqrt(np.mean(e * e))
        return (torch.tensor(r, dtype=torch.float32, device=DEV), torch.tensor(e, dtype=torch.float32, device=DEV))

    train = [pair() for _ in range(4)]
    held = pair()

    def loss(p):
        d = net.forward(p[0]) - p[1]
        return float((d * d).mean())

    l0, h0 = np.mean([loss(p) for p in train]), loss(held)
    pre = loss(train[0])
    first = net.train_step(*train[0], lr=2e-3)
    assert abs(first - pre) <= 1e-4 * pre  # the step reports the MSE before its update
    for it in range(1, 60):
        net.train_step(*train[it % 4], lr=2e-3)
    l1, h1 = np.mean([loss(p) for p in train]), loss(held)
    print(f"  train {l0:.3e} -> {l1:.3e}, held-out {h0:.3e} -> {h1:.3e}", flush=True)
    assert np.isfinite(l1) and l1 < 0.5 * l0
    assert h1 < 0.7 * h0
