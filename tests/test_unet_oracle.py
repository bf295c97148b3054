"""CPU pins of the U-Net oracle (oracle/unet_ref.py).

- The torch formulation (F.conv3d with circular padding, permuted layouts) is
  checked against direct numpy loops in the channels-last [voxel][channel]
  layout on non-cubic grids, so a transposed axis, a flipped tap or a wrong
  roll sign fails.
- A whole-net numpy restatement fixes the pooling, upsampling and
  concatenation order.
- Autograd gradients are checked against central finite differences.
"""
import numpy as np
import pytest
import torch

from oracle import unet_ref as U


def _rand_weights(rng, scale=0.3):
    W = {}
    for name, lvl, cin, cout, kind in U.LAYERS:
        n = cin * cout if kind == "1x1" else 27 * cout * cin
        W[name] = rng.standard_normal(n) * scale / np.sqrt(max(cin, 1))
    return W


def test_conv3_direct_single_point_by_loops():
    """One output value by explicit periodic index arithmetic."""
    rng = np.random.default_rng(0)
    shape, cin, cout = (5, 4, 3), 2, 3
    nx, ny, nz = shape
    x = rng.standard_normal((nx * ny * nz, cin))
    w = rng.standard_normal(27 * cout * cin)
    out = U.conv3_direct(x, w, shape, cin, cout)
    W = w.reshape(27, cout, cin)
    for (px, py, pz, co) in [(0, 0, 0, 0), (4, 3, 2, 2), (2, 1, 0, 1)]:
        s = 0.0
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    tap = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)
                    u = ((px + dx) % nx) + nx * (((py + dy) % ny) + ny * ((pz + dz) % nz))
                    s += W[tap, co] @ x[u]
        v = px + nx * (py + ny * pz)
        assert abs(out[v, co] - s) < 1e-12


@pytest.mark.parametrize("shape,cin,cout", [((5, 4, 3), 2, 3), ((4, 6, 5), 3, 2)])
def test_torch_conv3_matches_direct(shape, cin, cout):
    rng = np.random.default_rng(1)
    V = shape[0] * shape[1] * shape[2]
    x = rng.standard_normal((V, cin))
    w = rng.standard_normal(27 * cout * cin)
    ref = U.conv3_direct(x, w, shape, cin, cout)
    got = U.from_torch(U.conv3(U.to_torch(x, shape), U.weight3(w, cin, cout))).numpy()
    assert np.abs(got - ref).max() < 1e-12


def test_layout_round_trip():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((4 * 3 * 2, 5))
    t = U.to_torch(x, (4, 3, 2))
    assert tuple(t.shape) == (1, 5, 2, 3, 4)
    # voxel v = x + nx (y + ny z) sits at [z, y, x]
    assert t[0, 3, 1, 2, 3] == x[3 + 4 * (2 + 3 * 1), 3]
    assert np.array_equal(U.from_torch(t).numpy(), x)


def _forward_numpy(W, r, shape):
    """The U-Net in the [V][C] layout with numpy only (no quantization)."""
    def lv(k):
        return tuple(s >> k for s in shape)

    def pool(a, sh):
        nx, ny, nz = sh
        c = a.shape[1]
        t = a.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, c)
        return t.mean(axis=(1, 3, 5)).reshape(-1, c)

    def up(a, sh_fine):
        nx, ny, nz = sh_fine
        c = a.shape[1]
        t = a.reshape(nz // 2, ny // 2, nx // 2, c)
        t = t.repeat(2, axis=0).repeat(2, axis=1).repeat(2, axis=2)
        return t.reshape(-1, c)

    relu = lambda a: np.maximum(a, 0.0)  # noqa: E731
    cv = lambda a, name, k, ci, co: relu(U.conv3_direct(a, W[name], lv(k), ci, co))  # noqa: E731
    c0 = cv(r.reshape(-1, 1), "c0", 0, 1, 32)
    c1 = cv(pool(c0, lv(0)), "c1", 1, 32, 32)
    c2a = cv(pool(c1, lv(1)), "c2a", 2, 32, 64)
    c2 = cv(c2a, "c2", 2, 64, 64)
    b1 = cv(pool(c2, lv(2)), "b1", 3, 64, 64)
    b = cv(b1, "b", 3, 64, 64)
    d2 = cv(np.concatenate([up(b, lv(2)), c2], axis=1), "d2", 2, 128, 64)
    d1 = cv(np.concatenate([up(d2, lv(1)), c1], axis=1), "d1", 1, 96, 32)
    d0 = cv(np.concatenate([up(d1, lv(0)), c0], axis=1), "d0", 0, 64, 32)
    return d0 @ W["out"]


def test_forward_matches_numpy_restatement():
    rng = np.random.default_rng(3)
    shape = (16, 8, 8)
    W = _rand_weights(rng)
    r = rng.standard_normal(16 * 8 * 8)
    ref = _forward_numpy(W, r, shape)
    got = U.forward(W, r, shape, quantize=False).numpy()
    assert np.abs(got - ref).max() <= 1e-11 * max(1.0, np.abs(ref).max())


def test_quantized_forward_is_close_and_rounds():
    rng = np.random.default_rng(4)
    shape = (16, 8, 8)
    W = _rand_weights(rng)
    r = rng.standard_normal(16 * 8 * 8)
    y64 = U.forward(W, r, shape, quantize=False).numpy()
    yq = U.forward(W, r, shape, quantize=True).numpy()
    rel = np.linalg.norm(yq - y64) / np.linalg.norm(y64)
    assert 1e-5 < rel < 5e-2
    t = torch.tensor([1.0 + 2.0**-10, 3.0], dtype=torch.float64, requires_grad=True)
    q = U._bf(t, True)
    assert q[0].item() == 1.0 and q[1].item() == 3.0
    q.sum().backward()
    assert torch.equal(t.grad, torch.ones(2, dtype=torch.float64))  # straight-through


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(5)
    shape = (16, 8, 8)
    W = _rand_weights(rng)
    r = rng.standard_normal(16 * 8 * 8)
    dy = rng.standard_normal(16 * 8 * 8)
    G, _ = U.gradients(W, r, dy, shape, quantize=False)
    h = 1e-6
    for name, idx in [("c0", 5), ("c1", 1000), ("b", 77), ("d2", 4321), ("d0", 99), ("out", 7)]:
        Wp = {k: v.copy() for k, v in W.items()}
        Wm = {k: v.copy() for k, v in W.items()}
        Wp[name][idx] += h
        Wm[name][idx] -= h
        fp = float(U.forward(Wp, r, shape, quantize=False).numpy() @ dy)
        fm = float(U.forward(Wm, r, shape, quantize=False).numpy() @ dy)
        fd = (fp - fm) / (2 * h)
        assert abs(G[name][idx] - fd) <= 1e-6 * max(1.0, abs(fd)), (name, G[name][idx], fd)


def test_weight_file_checksum(tmp_path):
    """U-Net weight files carry the SHA-256 of their contents: a file read back
    unchanged loads; one changed weight, a missing digest or a missing layer
    is refused."""
    from nga2_amd import unet

    st = unet.init_weights(seed=4)
    st["kappa"] = np.array([0.25])
    p = tmp_path / "w.npz"
    unet.write_state(p, st)
    back = unet.read_state(p)
    assert all(np.array_equal(back[k], np.asarray(st[k])) for k in st)
    bad = dict(st)
    bad["d1"] = np.asarray(st["d1"]).copy()
    bad["d1"][17] = np.nextafter(bad["d1"][17], np.float32(1))  # one ulp
    with np.load(p) as z:
        files = {k: z[k] for k in z.files}
    files["d1"] = bad["d1"]
    q = tmp_path / "tampered.npz"
    np.savez(q, **files)
    with pytest.raises(ValueError, match="checksum"):
        unet.read_state(q)
    del files["sha256"]
    np.savez(q, **files)
    with pytest.raises(ValueError, match="sha256"):
        unet.read_state(q)
    files = {k: v for k, v in st.items() if k != "b1"}
    np.savez(q, sha256=np.frombuffer(b"0" * 64, dtype=np.uint8), **files)
    with pytest.raises(ValueError, match="missing"):
        unet.read_state(q)
