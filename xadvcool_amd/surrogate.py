"""Learned warm start: a 3-D convolutional encoder-decoder that predicts (u, p, T) of a cold-plate design.

Network (after the 3-D U-Net of the reference's benchmark suite,
benchmark/models/unet3d.py:49-125, reduced and made additive): four levels
with 16, 32, 64, 128 channels; two 3x3x3 conv + SiLU per level; 2x2x2
average pooling down, nearest x2 upsampling + conv up, additive skips; a
linear 3x3x3 head.  ~1.33 M parameters.  Every kernel is this project's HIP
code (libxadv_nn.so, include/xadv_nn.h): BF16 activations and weights into
v_mfma_f32_16x16x32_bf16, FP32 accumulation, FP32 master weights and Adam
moments.  Torch supplies device memory and the stream only.

Inputs (16 channels, channels-last, lattice padded to multiples of 8):
fluid / solid / inlet / outlet indicators, the heat map, inlet velocity on
fluid cells, tau, heat power, normalised coordinates.  Targets (5 channels):
u / u_in, (rho - 1) / u_in, (T - T_in) / P, each standardised by training-set
statistics kept with the weights.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
NN_LIB = os.path.join(HERE, "libxadv_nn.so")
CIN = 16
NOUT = 5
T_IN = 25.0
BC_SOLID, BC_INLET, BC_OUTLET = 1, 2, 3

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(NN_LIB):
            raise ImportError(f"{NN_LIB} is missing: build it with __graft_entry__.build(); there is no fallback")
        L = ctypes.CDLL(NN_LIB)
        vp, i, i64, f, fp = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_float, ctypes.c_void_p
        L.xnn_conv_fwd.argtypes = [vp, vp, vp, vp, i, i, i, i, i, i, i, vp]
        L.xnn_wflip.argtypes = [vp, vp, i, i, i, i, vp]
        L.xnn_conv_wgrad.argtypes = [vp, vp, vp, vp, i, i, i, i, i, i, i, i, vp]
        L.xnn_wgrad_part_floats.argtypes = [i, i, i, i, i, i, i, i]
        L.xnn_wgrad_part_floats.restype = i64
        L.xnn_bias_grad.argtypes = [vp, vp, vp, i64, i, vp]
        L.xnn_bias_part_floats.argtypes = [i64, i]
        L.xnn_bias_part_floats.restype = i64
        L.xnn_silu_fwd.argtypes = [vp, vp, i64, vp]
        L.xnn_silu_bwd.argtypes = [vp, vp, vp, i64, vp]
        L.xnn_to_bf16.argtypes = [vp, vp, i64, vp]
        L.xnn_add_f32.argtypes = [vp, vp, vp, i64, vp]
        L.xnn_add_bf16.argtypes = [vp, vp, vp, i64, vp]
        L.xnn_pool_fwd.argtypes = [vp, vp, i, i, i, i, i, vp]
        L.xnn_pool_bwd.argtypes = [vp, vp, i, i, i, i, i, vp]
        L.xnn_up_fwd.argtypes = [vp, vp, i, i, i, i, i, vp]
        L.xnn_up_bwd.argtypes = [vp, vp, i, i, i, i, i, vp]
        L.xnn_mse.argtypes = [vp, vp, vp, vp, vp, i64, i, i, f, vp]
        L.xnn_mse_blocks.argtypes = [i64, i]
        L.xnn_adam.argtypes = [vp, vp, vp, vp, vp, i64, f, f, f, f, i, f, vp]
        _lib = L
    return _lib


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _st():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed (code {rc})")


def round8(n):
    return (n + 7) // 8 * 8


# ------------------------------------------------------------ primitives ---

def conv_fwd(x, W, b, shape, cin, cout, Kp):
    """x: bf16 (N,X,Y,Z,cin); returns fp32 (N,X,Y,Z,cout)."""
    N, X, Y, Z = shape
    out = torch.empty((N, X, Y, Z, cout), dtype=torch.float32, device=x.device)
    _ok(lib().xnn_conv_fwd(_p(x), _p(W), _p(b), _p(out), N, X, Y, Z, cin, cout, Kp, _st()), "xnn_conv_fwd")
    return out


def to_bf16(x):
    y = torch.empty(x.shape, dtype=torch.bfloat16, device=x.device)
    _ok(lib().xnn_to_bf16(_p(x), _p(y), x.numel(), _st()), "xnn_to_bf16")
    return y


def silu_fwd(x):
    y = torch.empty(x.shape, dtype=torch.bfloat16, device=x.device)
    _ok(lib().xnn_silu_fwd(_p(x), _p(y), x.numel(), _st()), "xnn_silu_fwd")
    return y


def silu_bwd(x, gy):
    g = torch.empty(x.shape, dtype=torch.bfloat16, device=x.device)
    _ok(lib().xnn_silu_bwd(_p(x), _p(gy), _p(g), x.numel(), _st()), "xnn_silu_bwd")
    return g


def add_f32(a, b):
    y = torch.empty_like(a)
    _ok(lib().xnn_add_f32(_p(a), _p(b), _p(y), a.numel(), _st()), "xnn_add_f32")
    return y


def add_bf16(a, b):
    y = torch.empty_like(a)
    _ok(lib().xnn_add_bf16(_p(a), _p(b), _p(y), a.numel(), _st()), "xnn_add_bf16")
    return y


def pool_fwd(x, shape_c):
    N, X, Y, Z = shape_c
    C = x.shape[-1]
    y = torch.empty((N, X, Y, Z, C), dtype=torch.bfloat16, device=x.device)
    _ok(lib().xnn_pool_fwd(_p(x), _p(y), N, X, Y, Z, C, _st()), "xnn_pool_fwd")
    return y


def pool_bwd(gc, shape_f):
    N, X, Y, Z = shape_f
    C = gc.shape[-1]
    g = torch.empty((N, X, Y, Z, C), dtype=torch.float32, device=gc.device)
    _ok(lib().xnn_pool_bwd(_p(gc), _p(g), N, X, Y, Z, C, _st()), "xnn_pool_bwd")
    return g


def up_fwd(x, shape_f):
    N, X, Y, Z = shape_f
    C = x.shape[-1]
    y = torch.empty((N, X, Y, Z, C), dtype=torch.bfloat16, device=x.device)
    _ok(lib().xnn_up_fwd(_p(x), _p(y), N, X, Y, Z, C, _st()), "xnn_up_fwd")
    return y


def up_bwd(gf, shape_c):
    N, X, Y, Z = shape_c
    C = gf.shape[-1]
    g = torch.empty((N, X, Y, Z, C), dtype=torch.float32, device=gf.device)
    _ok(lib().xnn_up_bwd(_p(gf), _p(g), N, X, Y, Z, C, _st()), "xnn_up_bwd")
    return g


class Conv:
    """3x3x3 convolution: FP32 master weights W[cout][Kp] (k = tap*cin + ci), BF16 copy, Adam moments."""

    def __init__(self, cin, cout, gen, gain=1.0, device="cuda"):
        self.cin, self.cout = cin, cout
        self.Kp = (27 * cin + 31) // 32 * 32
        self.Kpt = (27 * cout + 31) // 32 * 32
        w = torch.zeros(cout, self.Kp, dtype=torch.float32)
        w[:, :27 * cin] = torch.randn(cout, 27 * cin, generator=gen) * (gain * math.sqrt(2.0 / (27 * cin)))
        self.w = w.to(device)
        self.b = torch.zeros(cout, dtype=torch.float32, device=device)
        self.wb = self.w.to(torch.bfloat16)
        self.m_w, self.v_w = torch.zeros_like(self.w), torch.zeros_like(self.w)
        self.m_b, self.v_b = torch.zeros_like(self.b), torch.zeros_like(self.b)
        self.dw, self.db = torch.zeros_like(self.w), torch.zeros_like(self.b)
        self.wt = torch.zeros(cin, self.Kpt, dtype=torch.bfloat16, device=device)
        self._parts = {}

    def fwd(self, x, shape):
        return conv_fwd(x, self.wb, self.b, shape, self.cin, self.cout, self.Kp)

    def dgrad(self, g_bf16, shape):
        """Gradient w.r.t. the input (fp32) of an upstream gradient g (bf16, w.r.t. this conv's output)."""
        L = lib()
        _ok(L.xnn_wflip(_p(self.wb), _p(self.wt), self.cin, self.cout, self.Kp, self.Kpt, _st()), "xnn_wflip")
        return conv_fwd(g_bf16, self.wt, None, shape, self.cout, self.cin, self.Kpt)

    def wgrad(self, x, g_bf16, shape, vchunk=4096):
        """Accumulate dW, db from the input x (bf16) and upstream gradient g (bf16)."""
        L = lib()
        N, X, Y, Z = shape
        nw = L.xnn_wgrad_part_floats(N, X, Y, Z, self.cin, self.cout, self.Kp, vchunk)
        nbp = L.xnn_bias_part_floats(N * X * Y * Z, self.cout)
        key = (nw, nbp)
        parts = self._parts.get(key)
        if parts is None:  # zeroed once: the chunked path relies on zero padded K entries
            parts = (torch.zeros(nw, dtype=torch.float32, device=x.device),
                     torch.zeros(nbp, dtype=torch.float32, device=x.device))
            self._parts[key] = parts
        part, bpart = parts
        dw = torch.empty_like(self.w)
        _ok(L.xnn_conv_wgrad(_p(x), _p(g_bf16), _p(part), _p(dw), N, X, Y, Z, self.cin, self.cout, self.Kp, vchunk,
                             _st()), "xnn_conv_wgrad")
        db = torch.empty_like(self.b)
        _ok(L.xnn_bias_grad(_p(g_bf16), _p(db), _p(bpart), N * X * Y * Z, self.cout, _st()), "xnn_bias_grad")
        self.dw.add_(dw)
        self.db.add_(db)

    def zero_grad(self):
        self.dw.zero_()
        self.db.zero_()

    def adam(self, step, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
        L = lib()
        _ok(L.xnn_adam(_p(self.w), _p(self.dw), _p(self.m_w), _p(self.v_w), _p(self.wb), self.w.numel(), lr, b1, b2,
                       eps, step, wd, _st()), "xnn_adam")
        _ok(L.xnn_adam(_p(self.b), _p(self.db), _p(self.m_b), _p(self.v_b), None, self.b.numel(), lr, b1, b2, eps,
                       step, 0.0, _st()), "xnn_adam")

    def state(self):
        return {"w": self.w.cpu().numpy(), "b": self.b.cpu().numpy()}

    def load(self, st):
        self.w.copy_(torch.as_tensor(st["w"]))
        self.b.copy_(torch.as_tensor(st["b"]))
        self.wb.copy_(self.w.to(torch.bfloat16))


class UNet3D:
    """Encoder-decoder with levels of CH channels; forward keeps the tape for backward."""

    CH = (16, 32, 64, 128)
    LAYERS = ("c1a", "c1b", "c2a", "c2b", "c3a", "c3b", "c4a", "c4b", "d3a", "d3b", "d2a", "d2b", "d1a", "d1b",
              "head")

    def __init__(self, seed=0, device="cuda"):
        gen = torch.Generator().manual_seed(seed)
        c = self.CH
        spec = {"c1a": (CIN, c[0]), "c1b": (c[0], c[0]), "c2a": (c[0], c[1]), "c2b": (c[1], c[1]),
                "c3a": (c[1], c[2]), "c3b": (c[2], c[2]), "c4a": (c[2], c[3]), "c4b": (c[3], c[3]),
                "d3a": (c[3], c[2]), "d3b": (c[2], c[2]), "d2a": (c[2], c[1]), "d2b": (c[1], c[1]),
                "d1a": (c[1], c[0]), "d1b": (c[0], c[0]), "head": (c[0], 16)}
        self.conv = {k: Conv(*spec[k], gen, gain=(0.5 if k == "head" else 1.0), device=device) for k in self.LAYERS}
        self.step_count = 0
        self.norm = None  # target standardisation (mean, std) per channel

    def n_params(self):
        return sum(c.cout * 27 * c.cin + c.cout for c in self.conv.values())

    @staticmethod
    def shapes(N, X, Y, Z):
        return [(N, X >> k, Y >> k, Z >> k) for k in range(4)]

    def forward(self, x, keep=False):
        """x: bf16 (N,X,Y,Z,16), X,Y,Z multiples of 8.  Returns fp32 (N,X,Y,Z,16) (first 5 channels used)."""
        N, X, Y, Z, _ = x.shape
        s = self.shapes(N, X, Y, Z)
        cv = self.conv
        t = {"x": x}
        # encoder
        t["p1a"] = cv["c1a"].fwd(x, s[0]); t["a1a"] = silu_fwd(t["p1a"])
        t["p1b"] = cv["c1b"].fwd(t["a1a"], s[0]); t["s1"] = silu_fwd(t["p1b"])
        t["q1"] = pool_fwd(t["s1"], s[1])
        t["p2a"] = cv["c2a"].fwd(t["q1"], s[1]); t["a2a"] = silu_fwd(t["p2a"])
        t["p2b"] = cv["c2b"].fwd(t["a2a"], s[1]); t["s2"] = silu_fwd(t["p2b"])
        t["q2"] = pool_fwd(t["s2"], s[2])
        t["p3a"] = cv["c3a"].fwd(t["q2"], s[2]); t["a3a"] = silu_fwd(t["p3a"])
        t["p3b"] = cv["c3b"].fwd(t["a3a"], s[2]); t["s3"] = silu_fwd(t["p3b"])
        t["q3"] = pool_fwd(t["s3"], s[3])
        t["p4a"] = cv["c4a"].fwd(t["q3"], s[3]); t["a4a"] = silu_fwd(t["p4a"])
        t["p4b"] = cv["c4b"].fwd(t["a4a"], s[3]); t["a4b"] = silu_fwd(t["p4b"])
        # decoder
        t["u3"] = up_fwd(t["a4b"], s[2])
        t["pd3a"] = cv["d3a"].fwd(t["u3"], s[2]); t["ad3a"] = silu_fwd(t["pd3a"])
        t["k3"] = add_bf16(t["ad3a"], t["s3"])
        t["pd3b"] = cv["d3b"].fwd(t["k3"], s[2]); t["ad3b"] = silu_fwd(t["pd3b"])
        t["u2"] = up_fwd(t["ad3b"], s[1])
        t["pd2a"] = cv["d2a"].fwd(t["u2"], s[1]); t["ad2a"] = silu_fwd(t["pd2a"])
        t["k2"] = add_bf16(t["ad2a"], t["s2"])
        t["pd2b"] = cv["d2b"].fwd(t["k2"], s[1]); t["ad2b"] = silu_fwd(t["pd2b"])
        t["u1"] = up_fwd(t["ad2b"], s[0])
        t["pd1a"] = cv["d1a"].fwd(t["u1"], s[0]); t["ad1a"] = silu_fwd(t["pd1a"])
        t["k1"] = add_bf16(t["ad1a"], t["s1"])
        t["pd1b"] = cv["d1b"].fwd(t["k1"], s[0]); t["ad1b"] = silu_fwd(t["pd1b"])
        out = cv["head"].fwd(t["ad1b"], s[0])
        if keep:
            self._tape = (t, s)
        return out

    def backward(self, gout):
        """gout: fp32 gradient of the loss w.r.t. the forward output.  Accumulates parameter gradients."""
        t, s = self._tape
        cv = self.conv

        def lin(name, x_in, g_out_bf16, shape):  # conv backward: wgrad + dgrad
            cv[name].wgrad(x_in, g_out_bf16, shape)
            return cv[name].dgrad(g_out_bf16, shape)

        g = to_bf16(gout)
        g_ad1b = lin("head", t["ad1b"], g, s[0])
        g_k1 = lin("d1b", t["k1"], silu_bwd(t["pd1b"], g_ad1b), s[0])
        g_s1 = g_k1
        g_u1 = lin("d1a", t["u1"], silu_bwd(t["pd1a"], g_k1), s[0])
        g_ad2b = up_bwd(g_u1, s[1])
        g_k2 = lin("d2b", t["k2"], silu_bwd(t["pd2b"], g_ad2b), s[1])
        g_s2 = g_k2
        g_u2 = lin("d2a", t["u2"], silu_bwd(t["pd2a"], g_k2), s[1])
        g_ad3b = up_bwd(g_u2, s[2])
        g_k3 = lin("d3b", t["k3"], silu_bwd(t["pd3b"], g_ad3b), s[2])
        g_s3 = g_k3
        g_u3 = lin("d3a", t["u3"], silu_bwd(t["pd3a"], g_k3), s[2])
        g_a4b = up_bwd(g_u3, s[3])
        g_a4a = lin("c4b", t["a4a"], silu_bwd(t["p4b"], g_a4b), s[3])
        g_q3 = lin("c4a", t["q3"], silu_bwd(t["p4a"], g_a4a), s[3])
        g_s3 = add_f32(g_s3, pool_bwd(g_q3, s[2]))
        g_a3a = lin("c3b", t["a3a"], silu_bwd(t["p3b"], g_s3), s[2])
        g_q2 = lin("c3a", t["q2"], silu_bwd(t["p3a"], g_a3a), s[2])
        g_s2 = add_f32(g_s2, pool_bwd(g_q2, s[1]))
        g_a2a = lin("c2b", t["a2a"], silu_bwd(t["p2b"], g_s2), s[1])
        g_q1 = lin("c2a", t["q1"], silu_bwd(t["p2a"], g_a2a), s[1])
        g_s1 = add_f32(g_s1, pool_bwd(g_q1, s[0]))
        g_a1a = lin("c1b", t["a1a"], silu_bwd(t["p1b"], g_s1), s[0])
        cv["c1a"].wgrad(t["x"], silu_bwd(t["p1a"], g_a1a), s[0])
        self._tape = None

    def zero_grad(self):
        for c in self.conv.values():
            c.zero_grad()

    def adam_step(self, lr, wd=0.0):
        self.step_count += 1
        for c in self.conv.values():
            c.adam(self.step_count, lr, wd=wd)

    def state_dict(self):
        d = {f"{k}.{n}": v for k, c in self.conv.items() for n, v in c.state().items()}
        if self.norm is not None:
            d["norm.mean"], d["norm.std"] = self.norm
        return d

    def load_state_dict(self, d):
        for k, c in self.conv.items():
            c.load({"w": d[f"{k}.w"], "b": d[f"{k}.b"]})
        if "norm.mean" in d:
            self.norm = (np.asarray(d["norm.mean"], np.float32), np.asarray(d["norm.std"], np.float32))

    def save(self, path):
        np.savez_compressed(path, **{k: (v.astype(np.float16) if k.endswith(".w") else v)
                                     for k, v in self.state_dict().items()})

    @classmethod
    def load_file(cls, path, device="cuda"):
        net = cls(device=device)
        with np.load(path) as z:
            net.load_state_dict({k: z[k].astype(np.float32) for k in z.files})
        return net


def mse_loss(pred, target, wmask):
    """Weighted MSE on the first NOUT channels; returns (loss (device scalar), grad (fp32, pred's shape))."""
    V = pred.numel() // pred.shape[-1]
    C = pred.shape[-1]
    # the kernel reads target and wmask as (V, NOUT): refuse anything else before it runs
    if C < NOUT or target.numel() != V * NOUT or wmask.numel() != V * NOUT:
        raise ValueError(f"mse_loss: pred {tuple(pred.shape)}, target {tuple(target.shape)} and wmask "
                         f"{tuple(wmask.shape)} must share the voxel shape, with {NOUT} target / weight channels")
    if not (pred.is_contiguous() and target.is_contiguous() and wmask.is_contiguous()):
        raise ValueError("mse_loss: tensors must be contiguous")
    L = lib()
    nb = L.xnn_mse_blocks(V, C)
    part = torch.empty(nb, dtype=torch.float32, device=pred.device)
    grad = torch.empty_like(pred)
    inv = 1.0 / max(float(wmask.sum().item()), 1.0)  # mean over weighted entries
    _ok(L.xnn_mse(_p(pred), _p(target), _p(wmask), _p(grad), _p(part), V, C, NOUT, inv, _st()), "xnn_mse")
    return part.sum(), grad


# --------------------------------------------------------- data mapping ---

def features(mask, heat, u_in, tau, P, pad_to=None):
    """Network input (X8, Y8, Z8, 16) float32 for one design, zero-padded to multiples of 8."""
    nx, ny, nz = mask.shape
    X, Y, Z = pad_to or (round8(nx), round8(ny), round8(nz))
    f = np.zeros((X, Y, Z, CIN), np.float32)
    fluid = (mask != BC_SOLID).astype(np.float32)
    f[:nx, :ny, :nz, 0] = fluid
    f[:nx, :ny, :nz, 1] = 1.0 - fluid
    f[:nx, :ny, :nz, 2] = (mask == BC_INLET)
    f[:nx, :ny, :nz, 3] = (mask == BC_OUTLET)
    hm = np.asarray(heat, np.float32)
    if hm.max() > 0:
        hm = hm / hm.max()
    f[:nx, :ny, :nz, 4] = hm
    f[:nx, :ny, :nz, 5] = fluid * (u_in * 30.0)
    f[:nx, :ny, :nz, 6] = (tau - 0.54) * 50.0
    f[:nx, :ny, :nz, 7] = P * 8.0
    f[:nx, :, :, 8] = (np.arange(nx, dtype=np.float32) / nx)[:, None, None]
    f[:, :ny, :, 9] = (np.arange(ny, dtype=np.float32) / ny)[None, :, None]
    f[:, :, :nz, 10] = (np.arange(nz, dtype=np.float32) / nz)[None, None, :]
    return f


def features_device(mask, heat, u_in, tau, P, pad_to=None):
    """``features`` built on the GPU from device tensors (int8 mask, FP32 heat): the same FP32 values.

    The heat map is divided by a device tensor (an IEEE division per cell, as numpy's) and the coordinate
    channels are formed on the host with numpy, so every value equals the host builder's."""
    import torch
    nx, ny, nz = (int(v) for v in mask.shape)
    X, Y, Z = pad_to or (round8(nx), round8(ny), round8(nz))
    dev = mask.device
    f = torch.zeros((X, Y, Z, CIN), dtype=torch.float32, device=dev)
    fluid = (mask != BC_SOLID).to(torch.float32)
    f[:nx, :ny, :nz, 0] = fluid
    f[:nx, :ny, :nz, 1] = 1.0 - fluid
    f[:nx, :ny, :nz, 2] = (mask == BC_INLET).to(torch.float32)
    f[:nx, :ny, :nz, 3] = (mask == BC_OUTLET).to(torch.float32)
    hm = heat.to(torch.float32)
    mx = hm.max()
    f[:nx, :ny, :nz, 4] = torch.where(mx > 0, hm / mx, hm)
    f[:nx, :ny, :nz, 5] = fluid * float(np.float32(u_in * 30.0))
    f[:nx, :ny, :nz, 6] = float(np.float32((tau - 0.54) * 50.0))
    f[:nx, :ny, :nz, 7] = float(np.float32(P * 8.0))
    f[:nx, :, :, 8] = torch.from_numpy(np.arange(nx, dtype=np.float32) / nx).to(dev)[:, None, None]
    f[:, :ny, :, 9] = torch.from_numpy(np.arange(ny, dtype=np.float32) / ny).to(dev)[None, :, None]
    f[:, :, :nz, 10] = torch.from_numpy(np.arange(nz, dtype=np.float32) / nz).to(dev)[None, None, :]
    return f


def targets(mask, rho, u, T, u_in, P, pad_to=None):
    """(X8, Y8, Z8, 5) raw targets and (.., 5) loss weights for one solved design."""
    nx, ny, nz = mask.shape
    X, Y, Z = pad_to or (round8(nx), round8(ny), round8(nz))
    t = np.zeros((X, Y, Z, NOUT), np.float32)
    w = np.zeros((X, Y, Z, NOUT), np.float32)
    fluid = (mask != BC_SOLID).astype(np.float32)
    t[:nx, :ny, :nz, 0:3] = u / u_in
    t[:nx, :ny, :nz, 3] = (rho - 1.0) / u_in
    t[:nx, :ny, :nz, 4] = (T - T_IN) / P
    w[:nx, :ny, :nz, 0:4] = fluid[..., None]
    w[:nx, :ny, :nz, 4] = 1.0
    return t, w
