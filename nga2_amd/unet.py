"""Small 3-D U-Net on the project's own HIP kernels (libngapnet.so).

It maps a residual r of the pressure equation to a correction e ~ A^-1 r:

- as the initial guess x0 = p_prev + N(b - A p_prev) for the next solve in a
  time sequence;
- as an extra stage of the flexible-PCG preconditioner.

N is applied scale-free: N(r) = s * kappa * net(r / s), with s = rms(r) and
kappa a fixed output scale fitted on the training pairs.

Architecture: 4 levels, 32 base channels, periodic 3x3x3 convolutions, BF16
weights and activations, FP32 accumulation. Average pooling goes down,
nearest-neighbour upsampling with channel concatenation goes up, and a 1x1
output layer closes the net:

    L0 (n^3)    c0 = relu(conv(r, 1->32))                   d0 = relu(conv([up d1, c0], 64->32))   y = conv1(d0)
    L1 (n/2)^3  c1 = relu(conv(pool c0, 32->32))            d1 = relu(conv([up d2, c1], 96->32))
    L2 (n/4)^3  c2 = relu(conv(relu(conv(pool c1,32->64)),64->64))   d2 = relu(conv([up b, c2], 128->64))
    L3 (n/8)^3  b  = relu(conv(relu(conv(pool c2,64->64)),64->64))

The convolutions with >= 32 input channels run as BF16 MFMA implicit GEMMs,
forward, data gradient (flipped weights) and weight gradient. The first layer
(1 input channel) and the 1x1 output are small FP32 kernels. The network is one
native object (csrc/ngap_netmodel.hip) that sequences the kernels; this module
holds its handle.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
NET_LIB = os.path.join(_HERE, "lib", "libngapnet.so")
_L = None

_vp = C.c_void_p
_i = C.c_int
_ll = C.c_longlong
_f = C.c_float


def lib():
    global _L
    if _L is None:
        if not os.path.exists(NET_LIB):
            raise RuntimeError(f"libngapnet.so not found at {NET_LIB}; run __graft_entry__.build()")
        L = C.CDLL(NET_LIB)
        L.ngapnet_conv3_fwd.argtypes = [_vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _vp]
        L.ngapnet_wgrad_parts.argtypes = [_i, _i, _i, _i, _i]
        L.ngapnet_conv3_wgrad.argtypes = [_vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp]
        L.ngapnet_flip_transpose.argtypes = [_vp, _vp, _i, _i, _vp]
        L.ngapnet_conv3_small_fwd.argtypes = [_vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _vp]
        L.ngapnet_conv3_small_wgrad.argtypes = [_vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp]
        L.ngapnet_small_wgrad_part_size.argtypes = [_i]
        L.ngapnet_small_wgrad_part_size.restype = _ll
        L.ngapnet_relu_bwd.argtypes = [_vp, _vp, _vp, _vp, _ll, _vp]
        L.ngapnet_to_bf16.argtypes = [_vp, _vp, _ll, _vp]
        L.ngapnet_pool2.argtypes = [_vp, _vp, _i, _i, _i, _i, _vp]
        L.ngapnet_pool2_bwd.argtypes = [_vp, _vp, _i, _i, _i, _i, _i, _vp]
        L.ngapnet_up2.argtypes = [_vp, _vp, _i, _i, _i, _i, _i, _i, _vp]
        L.ngapnet_up2_bwd.argtypes = [_vp, _vp, _i, _i, _i, _i, _i, _i, _vp]
        L.ngapnet_chan_copy_bf16.argtypes = [_vp, _vp, _ll, _i, _i, _i, _i, _i, _vp]
        L.ngapnet_chan_copy_f32.argtypes = [_vp, _vp, _ll, _i, _i, _i, _i, _i, _vp]
        L.ngapnet_conv1_fwd.argtypes = [_vp, _vp, _vp, _ll, _i, _vp]
        L.ngapnet_conv1_bwd.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _ll, _i, _vp]
        L.ngapnet_adam.argtypes = [_vp, _vp, _vp, _vp, _vp, _ll, _f, _f, _f, _f, _i, _f, _vp]
        L.ngapnet_sync.argtypes = []
        L.ngapnet_layer_info.argtypes = [_i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), C.POINTER(_i),
                                         C.POINTER(_ll)]
        L.ngapnet_num_weights.argtypes = []
        L.ngapnet_num_weights.restype = _ll
        L.ngapnet_create.argtypes = [C.POINTER(_vp), _i, _i, _i]
        L.ngapnet_destroy.argtypes = [_vp]
        L.ngapnet_set_weights.argtypes = [_vp, C.POINTER(_f)]
        L.ngapnet_get_weights.argtypes = [_vp, C.POINTER(_f)]
        L.ngapnet_get_grads.argtypes = [_vp, C.POINTER(_f)]
        L.ngapnet_forward.argtypes = [_vp, _vp, _vp, _vp]
        L.ngapnet_backward.argtypes = [_vp, _vp, _vp]
        L.ngapnet_model_adam.argtypes = [_vp, _f, _vp]
        L.ngapnet_train_step.argtypes = [_vp, _vp, _vp, _f, C.POINTER(C.c_double), _vp]
        L.ngapnet_set_kappa.argtypes = [_vp, C.c_double]
        L.ngapnet_stage.argtypes = [_vp, _vp, _vp, _ll, _vp]
        L.ngapnet_set_input_scale.argtypes = [_vp, _vp, _ll]
        _L = L
    return _L


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _ck(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc})")


def _stream():
    # The library runs on its own HIP runtime (the one in /opt/rocm); torch's
    # streams belong to the runtime bundled with torch and are not valid
    # handles there. Kernels go to the library's default stream, bracketed by
    # _enter() and _leave().
    return None


def _enter():
    """Finish torch's queued work before the library reads its buffers."""
    import torch

    torch.cuda.synchronize()


def _leave():
    """Finish the library's work before torch reads or frees the buffers."""
    _ck(lib().ngapnet_sync(), "ngapnet_sync")


# layer table: name, level, cin, cout, kind ("small" FP32 3^3, "mfma" 3^3, "1x1")
LAYERS = [
    ("c0", 0, 1, 32, "small"),
    ("c1", 1, 32, 32, "mfma"),
    ("c2a", 2, 32, 64, "mfma"),
    ("c2", 2, 64, 64, "mfma"),
    ("b1", 3, 64, 64, "mfma"),
    ("b", 3, 64, 64, "mfma"),
    ("d2", 2, 128, 64, "mfma"),
    ("d1", 1, 96, 32, "mfma"),
    ("d0", 0, 64, 32, "mfma"),
    ("out", 0, 32, 1, "1x1"),
]


def layer_table():
    """(name, level, cin, cout, kind, offset, size) per layer, read from the library."""
    L = lib()
    out = []
    for i, (name, *_rest) in enumerate(LAYERS):
        lvl, cin, cout, kind, off = _i(), _i(), _i(), _i(), _ll()
        _ck(L.ngapnet_layer_info(i, C.byref(lvl), C.byref(cin), C.byref(cout), C.byref(kind), C.byref(off)),
            "layer_info")
        size = cin.value * cout.value * (1 if kind.value == 2 else 27)
        out.append((name, lvl.value, cin.value, cout.value, ("small", "mfma", "1x1")[kind.value], off.value, size))
    return out


def init_weights(seed=0):
    """He-normal 3x3x3 layers, 1/sqrt(cin) output layer; dict name -> float32."""
    rng = np.random.default_rng(seed)
    W = {}
    for name, lvl, cin, cout, kind in LAYERS:
        if kind == "1x1":
            W[name] = (rng.standard_normal(cout * cin) * np.sqrt(1.0 / cin)).astype(np.float32)
        else:
            W[name] = (rng.standard_normal(27 * cout * cin) * np.sqrt(2.0 / (27 * cin))).astype(np.float32)
    return W


class UNet:
    """Handle to the native network object (ngapnet_create) for one grid size.

    Tensors passed in and out are torch device tensors; each call finishes
    torch's queued work first and the library's work before it returns."""

    def __init__(self, dims, seed=0, device="cuda:0"):
        import torch

        if isinstance(dims, int):
            dims = (dims, dims, dims)
        self.dims = tuple(int(v) for v in dims)
        self.dev = torch.device(device)
        self.V = self.dims[0] * self.dims[1] * self.dims[2]
        L = lib()
        self.table = layer_table()
        assert [t[:5] for t in self.table] == LAYERS, "layer table differs from the library"
        self.nw = int(L.ngapnet_num_weights())
        h = _vp()
        rc = L.ngapnet_create(C.byref(h), *self.dims)
        if rc == -1:
            raise ValueError(f"grid {self.dims}: every level needs nx % 16 == 0 and ny, nz % 4 == 0 "
                             "(4 levels: nx % 128, ny % 32, nz % 32)")
        _ck(rc, "ngapnet_create")
        self.h = h
        self.kappa = 1.0
        self.set_weights(init_weights(seed))

    # ------------------------------------------------------------------ weights
    def _flat(self, W):
        flat = np.zeros(self.nw, dtype=np.float32)
        for name, lvl, cin, cout, kind, off, size in self.table:
            flat[off:off + size] = np.asarray(W[name], dtype=np.float32).reshape(-1)
        return flat

    def _split(self, flat):
        return {name: flat[off:off + size].copy() for name, lvl, cin, cout, kind, off, size in self.table}

    def set_weights(self, W):
        flat = np.ascontiguousarray(self._flat(W))
        _enter()
        _ck(lib().ngapnet_set_weights(self.h, flat.ctypes.data_as(C.POINTER(_f))), "set_weights")

    def weights(self):
        flat = np.zeros(self.nw, dtype=np.float32)
        _ck(lib().ngapnet_get_weights(self.h, flat.ctypes.data_as(C.POINTER(_f))), "get_weights")
        return self._split(flat)

    def grads(self):
        flat = np.zeros(self.nw, dtype=np.float32)
        _ck(lib().ngapnet_get_grads(self.h, flat.ctypes.data_as(C.POINTER(_f))), "get_grads")
        return self._split(flat)

    def set_kappa(self, kappa):
        _ck(lib().ngapnet_set_kappa(self.h, float(kappa)), "set_kappa")
        self.kappa = float(kappa)

    def set_input_scale(self, d):
        """Per-cell input scale of the correction stage (FP64 device tensor of
        nx*ny*nz cells, e.g. 1/diag(A)), or None. The stage then computes
        e = s kappa net(d r / s) with s = rms(d r). The tensor is kept alive here."""
        import torch

        if d is None:
            self._dscale = None
            _ck(lib().ngapnet_set_input_scale(self.h, None, 0), "set_input_scale")
            return
        d = self._vec(d, torch.float64)
        self._dscale = d
        _enter()
        _ck(lib().ngapnet_set_input_scale(self.h, _p(d), self.V), "set_input_scale")

    # ------------------------------------------------------------------ compute
    def _vec(self, t, dtype):
        import torch

        t = t.reshape(-1)
        if t.numel() != self.V:
            raise ValueError(f"expected {self.V} cells, got {t.numel()}")
        return t.to(device=self.dev, dtype=dtype).contiguous()

    def forward(self, r):
        """y = net(r) for r a flat device tensor of nx*ny*nz cells (x fastest); FP32 result."""
        import torch

        x = self._vec(r, torch.float32)
        y = torch.empty(self.V, dtype=torch.float32, device=self.dev)
        _enter()
        _ck(lib().ngapnet_forward(self.h, _p(x), _p(y), None), "forward")
        _leave()
        return y

    def backward(self, dy):
        """Gradients of sum(dy * y) for the last forward (read them with grads())."""
        import torch

        d = self._vec(dy, torch.float32)
        _enter()
        _ck(lib().ngapnet_backward(self.h, _p(d), None), "backward")
        _leave()

    def adam_step(self, lr=1e-3):
        _enter()
        _ck(lib().ngapnet_model_adam(self.h, float(lr), None), "adam")
        _leave()

    def train_step(self, x, target, lr=1e-3):
        """Forward, MSE against target, backward and one Adam step; returns the loss."""
        import torch

        xx, tt = self._vec(x, torch.float32), self._vec(target, torch.float32)
        loss = C.c_double()
        _enter()
        _ck(lib().ngapnet_train_step(self.h, _p(xx), _p(tt), float(lr), C.byref(loss), None), "train_step")
        _leave()
        return loss.value

    def apply(self, r):
        """Scale-free correction e = s kappa net(r/s), s = rms(r), FP64 in and out
        (the correction stage the solver calls)."""
        import torch

        rr = self._vec(r, torch.float64)
        e = torch.empty(self.V, dtype=torch.float64, device=self.dev)
        _enter()
        _ck(lib().ngapnet_stage(self.h, _p(rr), _p(e), self.V, None), "stage")
        _leave()
        return e

    def stage(self):
        """(function pointer, context) of the correction stage for ngap_set_stage."""
        return C.cast(lib().ngapnet_stage, _vp), self.h

    def close(self):
        if getattr(self, "h", None):
            lib().ngapnet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ checkpoint
    def state(self):
        st = self.weights()
        st["kappa"] = np.array([self.kappa])
        return st

    def load_state(self, st):
        self.set_weights({k: v for k, v in st.items() if k != "kappa"})
        if "kappa" in st:
            self.set_kappa(float(np.asarray(st["kappa"]).reshape(-1)[0]))

    def save(self, path):
        """Write the weights and kappa to an .npz file together with the
        SHA-256 of their bytes (see state_digest); load() refuses a file whose
        digest does not match."""
        write_state(path, self.state())

    def load(self, path):
        self.load_state(read_state(path))


def state_digest(st):
    """SHA-256 (hex) over the layer names, shapes and float32 bytes of a
    state dict, in the library's layer order, then kappa as float64."""
    import hashlib

    h = hashlib.sha256()
    for name, *_rest in LAYERS:
        a = np.ascontiguousarray(np.asarray(st[name], dtype=np.float32).reshape(-1))
        h.update(name.encode())
        h.update(np.int64(a.size).tobytes())
        h.update(a.tobytes())
    h.update(np.asarray(st["kappa"], dtype=np.float64).reshape(-1).tobytes())
    return h.hexdigest()


def write_state(path, st):
    """Write a state dict (layer name -> float32 weights, "kappa") to an .npz
    file with its digest."""
    np.savez(path, sha256=np.frombuffer(state_digest(st).encode(), dtype=np.uint8), **st)


def read_state(path):
    """State dict of a file written by UNet.save, after checking its digest
    (ValueError on a mismatch or a missing digest). Loaded without pickle."""
    with np.load(path, allow_pickle=False) as z:
        if "sha256" not in z.files:
            raise ValueError(f"{path}: no sha256 entry")
        want = bytes(z["sha256"]).decode()
        st = {k: z[k].copy() for k in z.files if k != "sha256"}
    missing = [name for name, *_r in LAYERS if name not in st] + ([] if "kappa" in st else ["kappa"])
    if missing:
        raise ValueError(f"{path}: missing {missing}")
    got = state_digest(st)
    if got != want:
        raise ValueError(f"{path}: checksum mismatch (file {want[:16]}..., contents {got[:16]}...)")
    return st


def time_net(n=128, reps=5, device="cuda:0"):
    """Milliseconds per forward pass and per training step (forward, backward,
    Adam) of the net on an n^3 grid, random input and target, warm."""
    import time

    import torch

    torch.zeros(1, device=device)
    net = UNet((n, n, n), seed=0, device=device)
    x = torch.randn(n ** 3, device=device)
    t = torch.randn(n ** 3, device=device)
    try:
        for _ in range(3):
            net.forward(x)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            net.forward(x)
        torch.cuda.synchronize()
        fwd = (time.perf_counter() - t0) / reps * 1e3
        for _ in range(2):
            net.train_step(x, t, lr=1e-9)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            net.train_step(x, t, lr=1e-9)
        torch.cuda.synchronize()
        step = (time.perf_counter() - t0) / reps * 1e3
    finally:
        net.close()
    return fwd, step


if __name__ == "__main__":
    import sys

    _n = int(sys.argv[1]) if len(sys.argv) > 1 else 128
    _fwd, _step = time_net(_n)
    print(f"unet {_n}^3: forward {_fwd:.3f} ms, training step {_step:.3f} ms", flush=True)
