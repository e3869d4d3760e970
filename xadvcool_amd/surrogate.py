"""Learned warm start: a 3-D convolutional encoder-decoder that predicts (u, p, T) of a cold-plate design.

Network (after the 3-D U-Net of the reference's benchmark suite,
benchmark/models/unet3d.py:49-125, made additive): four levels with 16, 32,
64, 128 channels; two 3x3x3 conv + SiLU per level; 2x2x2 average pooling
down, nearest x2 upsampling + conv up, additive skips; a residual block of
two more 128-channel convolutions at the bottleneck; a linear 3x3x3 head.
The design's scalars (tau, heat power, inlet velocity) condition the
network through FiLM (Perez et al. 2018; the reference template conditions
its bottleneck the same way, unet3d.py:34-45, 83-84): a two-layer MLP
(3 -> 64 -> 480) gives a per-channel scale and shift of the pre-activations
of the last convolution of the bottleneck and of each decoder level.  The
residual block's second convolution and the MLP's last layer start at zero,
so a network without them (the first shipped weights) loads as the same
function.  ~2.24 M parameters.  Every kernel is this project's HIP code
(libxadv_nn.so, include/xadv_nn.h): BF16 activations and weights into
v_mfma_f32_16x16x32_bf16, FP32 accumulation, FP32 master weights and Adam
moments; the FiLM and MLP layers in FP32.  Torch supplies device memory and
the stream only.

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
        L.xnn_conv_fwd_silu.argtypes = [vp, vp, vp, vp, vp, i, i, i, i, i, i, i, vp]
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
        L.xnn_pool_bwd_add.argtypes = [vp, vp, vp, i, i, i, i, i, vp]
        L.xnn_up_fwd.argtypes = [vp, vp, i, i, i, i, i, vp]
        L.xnn_up_bwd.argtypes = [vp, vp, i, i, i, i, i, vp]
        L.xnn_mse.argtypes = [vp, vp, vp, vp, vp, i64, i, i, f, vp]
        L.xnn_mse_blocks.argtypes = [i64, i]
        L.xnn_adam.argtypes = [vp, vp, vp, vp, vp, i64, f, f, f, f, i, f, vp]
        L.xnn_silu_bwd_f32.argtypes = [vp, vp, vp, i64, vp]
        L.xnn_film_fwd.argtypes = [vp, vp, vp, i, i64, i, i, i, vp]
        L.xnn_film_bwd.argtypes = [vp, vp, vp, vp, vp, vp, i, i64, i, i, i, vp]
        L.xnn_film_silu_fwd.argtypes = [vp, vp, vp, i, i64, i, i, i, vp]
        L.xnn_film_silu_bwd.argtypes = [vp, vp, vp, vp, vp, vp, i, i64, i, i, i, vp]
        L.xnn_film_part_floats.argtypes = [i, i64, i]
        L.xnn_film_part_floats.restype = i64
        L.xnn_dense_fwd.argtypes = [vp, vp, vp, vp, vp, i, i, i, i, vp]
        L.xnn_dense_bwd.argtypes = [vp, vp, vp, vp, vp, vp, vp, i, i, i, i, vp]
        L.xnn_act_bwd.argtypes = [vp, vp, vp, vp, vp, vp, vp, i, i64, i, i, i, vp]
        L.xnn_act_part_floats.argtypes = [i, i64, i, i]
        L.xnn_act_part_floats.restype = i64
        L.xnn_mse_bf16.argtypes = [vp, vp, vp, vp, vp, vp, i64, i, i, vp]
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


def conv_fwd_silu(x, W, b, shape, cin, cout, Kp):
    """conv_fwd and bf16(SiLU) of its output in one call: returns (fp32 output, bf16 activation)."""
    N, X, Y, Z = shape
    out = torch.empty((N, X, Y, Z, cout), dtype=torch.float32, device=x.device)
    act = torch.empty((N, X, Y, Z, cout), dtype=torch.bfloat16, device=x.device)
    _ok(lib().xnn_conv_fwd_silu(_p(x), _p(W), _p(b), _p(out), _p(act), N, X, Y, Z, cin, cout, Kp, _st()),
        "xnn_conv_fwd_silu")
    return out, act


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


def silu_bwd_f32(x, gy):
    g = torch.empty(x.shape, dtype=torch.float32, device=x.device)
    _ok(lib().xnn_silu_bwd_f32(_p(x), _p(gy), _p(g), x.numel(), _st()), "xnn_silu_bwd_f32")
    return g


def film_fwd(p, film, off):
    """p (N, X, Y, Z, C) FP32 modulated by rows of ``film`` (N, L): p * (1 + film[:, off:off+C]) + film[:, off+C:]."""
    N, C = p.shape[0], p.shape[-1]
    y = torch.empty_like(p)
    _ok(lib().xnn_film_fwd(_p(p), _p(film), _p(y), N, p.numel() // (N * C), C, off, film.shape[1], _st()),
        "xnn_film_fwd")
    return y


def film_bwd(p, gy, film, dfilm, off):
    """Gradient w.r.t. p of a film_fwd output gradient gy (FP32); writes dfilm[:, off:off+2C]."""
    N, C = p.shape[0], p.shape[-1]
    Vn = p.numel() // (N * C)
    L = lib()
    part = torch.empty(L.xnn_film_part_floats(N, Vn, C), dtype=torch.float32, device=p.device)
    gx = torch.empty_like(p)
    _ok(L.xnn_film_bwd(_p(p), _p(gy), _p(film), _p(gx), _p(part), _p(dfilm), N, Vn, C, off, film.shape[1], _st()),
        "xnn_film_bwd")
    return gx


def film_silu_fwd(p, film, off):
    """silu(film_fwd(p, film, off)) in BF16, one pass (the same bits)."""
    N, C = p.shape[0], p.shape[-1]
    y = torch.empty(p.shape, dtype=torch.bfloat16, device=p.device)
    _ok(lib().xnn_film_silu_fwd(_p(p), _p(film), _p(y), N, p.numel() // (N * C), C, off, film.shape[1], _st()),
        "xnn_film_silu_fwd")
    return y


def film_silu_bwd(p, gy, film, dfilm, off):
    """Backward of film_silu_fwd from gy (FP32, w.r.t. its output): BF16 gradient w.r.t. p; writes
    dfilm[:, off:off+2C] (the bits of silu_bwd_f32 + film_bwd + to_bf16)."""
    N, C = p.shape[0], p.shape[-1]
    Vn = p.numel() // (N * C)
    L = lib()
    part = torch.empty(L.xnn_film_part_floats(N, Vn, C), dtype=torch.float32, device=p.device)
    gx = torch.empty(p.shape, dtype=torch.bfloat16, device=p.device)
    _ok(L.xnn_film_silu_bwd(_p(p), _p(gy), _p(film), _p(gx), _p(part), _p(dfilm), N, Vn, C, off, film.shape[1],
                            _st()), "xnn_film_silu_bwd")
    return gx


def act_bwd(p, gy, film=None, dfilm=None, off=0, want_db=True):
    """Backward of an activation site (SiLU, or FiLM then SiLU with ``film``) from gy (FP32, w.r.t. its output)
    in one pass: returns (BF16 gradient w.r.t. p, bias gradient of the convolution that produced p, or None);
    writes dfilm[:, off:off+2C] when ``film`` is given.  Elementwise the bits of silu_bwd / film_silu_bwd."""
    N, C = p.shape[0], p.shape[-1]
    Vn = p.numel() // (N * C)
    L = lib()
    part = torch.empty(L.xnn_act_part_floats(N, Vn, C, int(film is not None)), dtype=torch.float32, device=p.device)
    gx = torch.empty(p.shape, dtype=torch.bfloat16, device=p.device)
    db = torch.empty(C, dtype=torch.float32, device=p.device) if want_db else None
    Lf = film.shape[1] if film is not None else 0
    _ok(L.xnn_act_bwd(_p(p), _p(gy), _p(film), _p(gx), _p(part), _p(dfilm), _p(db), N, Vn, C, off, Lf, _st()),
        "xnn_act_bwd")
    return gx, db


class _Grads:
    """Gradient accumulators dw, db of a layer with parameters w, b.  zero_grad() only marks them cleared: the
    next accumulation takes the new gradients as they are (no fill kernel, no add kernel), and reading dw or db
    while cleared zeroes them first."""

    def _init_grads(self):
        self._dw, self._db = torch.zeros_like(self.w), torch.zeros_like(self.b)
        self._cleared = True

    @property
    def dw(self):
        self._materialize()
        return self._dw

    @property
    def db(self):
        self._materialize()
        return self._db

    def _materialize(self):
        if self._cleared:
            self._dw.zero_()
            self._db.zero_()
            self._cleared = False

    def _accumulate(self, dw, db):
        if self._cleared:  # fresh tensors of the caller's: they become the accumulators
            self._dw, self._db, self._cleared = dw, db, False
        else:
            self._dw.add_(dw)
            self._db.add_(db)

    def zero_grad(self):
        self._cleared = True


class Dense(_Grads):
    """Fully connected FP32 layer (W: O x I) with Adam moments; ``zero`` starts it at zero."""

    def __init__(self, fin, fout, gen, act, zero=False, device="cuda"):
        self.fin, self.fout, self.act = fin, fout, int(bool(act))
        w = torch.zeros(fout, fin) if zero else torch.randn(fout, fin, generator=gen) * math.sqrt(2.0 / fin)
        self.w = w.to(device)
        self.b = torch.zeros(fout, dtype=torch.float32, device=device)
        self.m_w, self.v_w = torch.zeros_like(self.w), torch.zeros_like(self.w)
        self.m_b, self.v_b = torch.zeros_like(self.b), torch.zeros_like(self.b)
        self._init_grads()

    def fwd(self, x):
        N = x.shape[0]
        pre = torch.empty(N, self.fout, dtype=torch.float32, device=x.device)
        post = torch.empty_like(pre)
        _ok(lib().xnn_dense_fwd(_p(x), _p(self.w), _p(self.b), _p(pre), _p(post), N, self.fin, self.fout, self.act,
                                _st()), "xnn_dense_fwd")
        return pre, post

    def bwd(self, x, pre, gy, want_gx=True):
        N = x.shape[0]
        dw, db = torch.empty_like(self.w), torch.empty_like(self.b)
        gx = torch.empty_like(x) if want_gx else None
        _ok(lib().xnn_dense_bwd(_p(x), _p(self.w), _p(pre), _p(gy), _p(dw), _p(db), _p(gx), N, self.fin, self.fout,
                                self.act, _st()), "xnn_dense_bwd")
        self._accumulate(dw, db)
        return gx

    def adam(self, step, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
        L = lib()
        _ok(L.xnn_adam(_p(self.w), _p(self.dw), _p(self.m_w), _p(self.v_w), None, self.w.numel(), lr, b1, b2, eps,
                       step, wd, _st()), "xnn_adam")
        _ok(L.xnn_adam(_p(self.b), _p(self.db), _p(self.m_b), _p(self.v_b), None, self.b.numel(), lr, b1, b2, eps,
                       step, 0.0, _st()), "xnn_adam")

    def state(self):
        return {"w": self.w.cpu().numpy(), "b": self.b.cpu().numpy()}

    def load(self, st):
        self.w.copy_(torch.as_tensor(st["w"]))
        self.b.copy_(torch.as_tensor(st["b"]))


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


def pool_bwd_add(gc, acc, shape_f):
    """acc + pool_bwd(gc, shape_f) in one pass (the same bits as add_f32 of the two)."""
    N, X, Y, Z = shape_f
    C = gc.shape[-1]
    if tuple(acc.shape) != (N, X, Y, Z, C) or acc.dtype != torch.float32:
        raise ValueError("pool_bwd_add: acc must be the fp32 fine-level gradient")
    y = torch.empty_like(acc)
    _ok(lib().xnn_pool_bwd_add(_p(gc), _p(acc), _p(y), N, X, Y, Z, C, _st()), "xnn_pool_bwd_add")
    return y


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


class Conv(_Grads):
    """3x3x3 convolution: FP32 master weights W[cout][Kp] (k = tap*cin + ci), BF16 copy, Adam moments."""

    def __init__(self, cin, cout, gen, gain=1.0, device="cuda"):
        self.cin, self.cout = cin, cout
        self.Kp = (27 * cin + 31) // 32 * 32
        self.Kpt = (27 * cout + 31) // 32 * 32
        w = torch.zeros(cout, self.Kp, dtype=torch.float32)
        r = torch.randn(cout, 27 * cin, generator=gen)  # drawn even at gain 0, so later layers' draws stay put
        w[:, :27 * cin] = r * (gain * math.sqrt(2.0 / (27 * cin)))
        self.w = w.to(device)
        self.b = torch.zeros(cout, dtype=torch.float32, device=device)
        self.wb = self.w.to(torch.bfloat16)
        self.m_w, self.v_w = torch.zeros_like(self.w), torch.zeros_like(self.w)
        self.m_b, self.v_b = torch.zeros_like(self.b), torch.zeros_like(self.b)
        self._init_grads()
        self.wt = torch.zeros(cin, self.Kpt, dtype=torch.bfloat16, device=device)
        self._parts = {}

    def fwd(self, x, shape):
        return conv_fwd(x, self.wb, self.b, shape, self.cin, self.cout, self.Kp)

    def fwd_silu(self, x, shape):
        """(conv output, bf16 SiLU of it): the same values as fwd then silu_fwd."""
        return conv_fwd_silu(x, self.wb, self.b, shape, self.cin, self.cout, self.Kp)

    def dgrad(self, g_bf16, shape):
        """Gradient w.r.t. the input (fp32) of an upstream gradient g (bf16, w.r.t. this conv's output)."""
        L = lib()
        _ok(L.xnn_wflip(_p(self.wb), _p(self.wt), self.cin, self.cout, self.Kp, self.Kpt, _st()), "xnn_wflip")
        return conv_fwd(g_bf16, self.wt, None, shape, self.cout, self.cin, self.Kpt)

    def wgrad(self, x, g_bf16, shape, vchunk=4096, db=None):
        """Accumulate dW, db from the input x (bf16) and upstream gradient g (bf16); ``db`` (the bias gradient,
        when the activation backward already summed it) skips the bias reduction."""
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
        if db is None:
            db = torch.empty_like(self.b)
            _ok(L.xnn_bias_grad(_p(g_bf16), _p(db), _p(bpart), N * X * Y * Z, self.cout, _st()), "xnn_bias_grad")
        self._accumulate(dw, db)

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


def scalars(tau, P, u_in):
    """The design's conditioning scalars, scaled as the constant input channels are (``features``)."""
    return np.array([(tau - 0.54) * 50.0, P * 8.0, u_in * 30.0], np.float32)


class UNet3D:
    """Encoder-decoder with levels of CH channels, a residual bottleneck block and FiLM conditioning on the
    design's scalars; forward keeps the tape for backward."""

    CH = (16, 32, 64, 128)
    LAYERS = ("c1a", "c1b", "c2a", "c2b", "c3a", "c3b", "c4a", "c4b", "c4c", "c4d", "d3a", "d3b", "d2a", "d2b",
              "d1a", "d1b", "head")
    # FiLM sites: the pre-activation each modulates and its offset in the FiLM row ([scale | shift] per site).
    FILM = (("p4b", 0, 128), ("pd3b", 256, 64), ("pd2b", 384, 32), ("pd1b", 448, 16))
    FILM_LEN = 480
    FILM_HIDDEN = 64

    def __init__(self, seed=0, device="cuda"):
        gen = torch.Generator().manual_seed(seed)
        c = self.CH
        spec = {"c1a": (CIN, c[0]), "c1b": (c[0], c[0]), "c2a": (c[0], c[1]), "c2b": (c[1], c[1]),
                "c3a": (c[1], c[2]), "c3b": (c[2], c[2]), "c4a": (c[2], c[3]), "c4b": (c[3], c[3]),
                "d3a": (c[3], c[2]), "d3b": (c[2], c[2]), "d2a": (c[2], c[1]), "d2b": (c[1], c[1]),
                "d1a": (c[1], c[0]), "d1b": (c[0], c[0]), "head": (c[0], 16)}
        # the first fifteen layers draw their initial weights in the order of the network without the
        # residual block, so a seed gives them the same values either way
        self.conv = {k: Conv(*spec[k], gen, gain=(0.5 if k == "head" else 1.0), device=device) for k in spec}
        self.conv["c4c"] = Conv(c[3], c[3], gen, device=device)
        self.conv["c4d"] = Conv(c[3], c[3], gen, gain=0.0, device=device)  # starts at zero: block = identity
        self.conv = {k: self.conv[k] for k in self.LAYERS}
        self.film1 = Dense(3, self.FILM_HIDDEN, gen, act=True, device=device)
        self.film2 = Dense(self.FILM_HIDDEN, self.FILM_LEN, gen, act=False, zero=True, device=device)
        self.step_count = 0
        self.norm = None  # target standardisation (mean, std) per channel

    def n_params(self):
        return (sum(c.cout * 27 * c.cin + c.cout for c in self.conv.values()) +
                sum(d.w.numel() + d.b.numel() for d in (self.film1, self.film2)))

    @staticmethod
    def shapes(N, X, Y, Z):
        return [(N, X >> k, Y >> k, Z >> k) for k in range(4)]

    def forward(self, x, s, keep=False):
        """x: bf16 (N,X,Y,Z,16), X,Y,Z multiples of 8; s: FP32 (N,3) design scalars (``scalars``).
        Returns fp32 (N,X,Y,Z,16) (first 5 channels used)."""
        N, X, Y, Z, _ = x.shape
        if tuple(s.shape) != (N, 3) or s.dtype != torch.float32 or not s.is_contiguous():
            raise ValueError(f"scalars must be a contiguous float32 ({N}, 3) tensor, got {tuple(s.shape)}")
        sh = self.shapes(N, X, Y, Z)
        cv = self.conv
        t = {"x": x, "s": s}
        t["h1pre"], t["h1"] = self.film1.fwd(s)
        film = self.film2.fwd(t["h1"])[0]
        t["film"] = film
        off = {name: o for name, o, _ in self.FILM}

        def act(name, conv, x_in, shape):  # conv, then FiLM (at its sites) and SiLU; the tape keeps the conv output
            if name in off:
                t[name] = cv[conv].fwd(x_in, shape)
                return film_silu_fwd(t[name], film, off[name])
            t[name], a = cv[conv].fwd_silu(x_in, shape)
            return a

        # encoder
        t["a1a"] = act("p1a", "c1a", x, sh[0])
        t["s1"] = act("p1b", "c1b", t["a1a"], sh[0])
        t["q1"] = pool_fwd(t["s1"], sh[1])
        t["a2a"] = act("p2a", "c2a", t["q1"], sh[1])
        t["s2"] = act("p2b", "c2b", t["a2a"], sh[1])
        t["q2"] = pool_fwd(t["s2"], sh[2])
        t["a3a"] = act("p3a", "c3a", t["q2"], sh[2])
        t["s3"] = act("p3b", "c3b", t["a3a"], sh[2])
        t["q3"] = pool_fwd(t["s3"], sh[3])
        t["a4a"] = act("p4a", "c4a", t["q3"], sh[3])
        t["a4b"] = act("p4b", "c4b", t["a4a"], sh[3])
        # residual block at the bottleneck
        t["a4c"] = act("p4c", "c4c", t["a4b"], sh[3])
        t["a4d"] = act("p4d", "c4d", t["a4c"], sh[3])
        t["r4"] = add_bf16(t["a4b"], t["a4d"])
        # decoder
        t["u3"] = up_fwd(t["r4"], sh[2])
        t["ad3a"] = act("pd3a", "d3a", t["u3"], sh[2])
        t["k3"] = add_bf16(t["ad3a"], t["s3"])
        t["ad3b"] = act("pd3b", "d3b", t["k3"], sh[2])
        t["u2"] = up_fwd(t["ad3b"], sh[1])
        t["ad2a"] = act("pd2a", "d2a", t["u2"], sh[1])
        t["k2"] = add_bf16(t["ad2a"], t["s2"])
        t["ad2b"] = act("pd2b", "d2b", t["k2"], sh[1])
        t["u1"] = up_fwd(t["ad2b"], sh[0])
        t["ad1a"] = act("pd1a", "d1a", t["u1"], sh[0])
        t["k1"] = add_bf16(t["ad1a"], t["s1"])
        t["ad1b"] = act("pd1b", "d1b", t["k1"], sh[0])
        out = cv["head"].fwd(t["ad1b"], sh[0])
        if keep:
            self._tape = (t, sh)
        return out

    def backward(self, gout):
        """gout: fp32 gradient of the loss w.r.t. the forward output.  Accumulates parameter gradients."""
        t, sh = self._tape
        cv = self.conv
        off = {name: o for name, o, _ in self.FILM}
        dfilm = torch.empty_like(t["film"])

        def lin(name, x_in, g_out, shape):  # conv backward: wgrad (+ bias) + dgrad
            g_bf16, db = g_out if isinstance(g_out, tuple) else (g_out, None)
            cv[name].wgrad(x_in, g_bf16, shape, db=db)
            return cv[name].dgrad(g_bf16, shape)

        def dact(name, g):  # from the gradient w.r.t. SiLU's output: (bf16 one w.r.t. the conv output p, bias grad)
            if name not in off:
                return act_bwd(t[name], g)
            return act_bwd(t[name], g, t["film"], dfilm, off[name])

        g = gout if gout.dtype == torch.bfloat16 else to_bf16(gout)
        g_ad1b = lin("head", t["ad1b"], g, sh[0])
        g_k1 = lin("d1b", t["k1"], dact("pd1b", g_ad1b), sh[0])
        g_s1 = g_k1
        g_u1 = lin("d1a", t["u1"], dact("pd1a", g_k1), sh[0])
        g_ad2b = up_bwd(g_u1, sh[1])
        g_k2 = lin("d2b", t["k2"], dact("pd2b", g_ad2b), sh[1])
        g_s2 = g_k2
        g_u2 = lin("d2a", t["u2"], dact("pd2a", g_k2), sh[1])
        g_ad3b = up_bwd(g_u2, sh[2])
        g_k3 = lin("d3b", t["k3"], dact("pd3b", g_ad3b), sh[2])
        g_s3 = g_k3
        g_u3 = lin("d3a", t["u3"], dact("pd3a", g_k3), sh[2])
        g_r4 = up_bwd(g_u3, sh[3])
        g_a4c = lin("c4d", t["a4c"], dact("p4d", g_r4), sh[3])
        g_a4b = add_f32(g_r4, lin("c4c", t["a4b"], dact("p4c", g_a4c), sh[3]))
        g_a4a = lin("c4b", t["a4a"], dact("p4b", g_a4b), sh[3])
        g_q3 = lin("c4a", t["q3"], dact("p4a", g_a4a), sh[3])
        g_s3 = pool_bwd_add(g_q3, g_s3, sh[2])
        g_a3a = lin("c3b", t["a3a"], dact("p3b", g_s3), sh[2])
        g_q2 = lin("c3a", t["q2"], dact("p3a", g_a3a), sh[2])
        g_s2 = pool_bwd_add(g_q2, g_s2, sh[1])
        g_a2a = lin("c2b", t["a2a"], dact("p2b", g_s2), sh[1])
        g_q1 = lin("c2a", t["q1"], dact("p2a", g_a2a), sh[1])
        g_s1 = pool_bwd_add(g_q1, g_s1, sh[0])
        g_a1a = lin("c1b", t["a1a"], dact("p1b", g_s1), sh[0])
        g1a, db1a = dact("p1a", g_a1a)
        cv["c1a"].wgrad(t["x"], g1a, sh[0], db=db1a)
        # every FiLM row entry was written by its site's film_bwd; on through the MLP
        g_h1 = self.film2.bwd(t["h1"], t["film"], dfilm)
        self.film1.bwd(t["s"], t["h1pre"], g_h1, want_gx=False)
        self._tape = None

    def zero_grad(self):
        for c in self.conv.values():
            c.zero_grad()
        self.film1.zero_grad()
        self.film2.zero_grad()

    def adam_step(self, lr, wd=0.0):
        self.step_count += 1
        for c in self.conv.values():
            c.adam(self.step_count, lr, wd=wd)
        self.film1.adam(self.step_count, lr)
        self.film2.adam(self.step_count, lr)

    def state_dict(self):
        d = {f"{k}.{n}": v for k, c in self.conv.items() for n, v in c.state().items()}
        for k, m in (("film1", self.film1), ("film2", self.film2)):
            for n, v in m.state().items():
                d[f"{k}.{n}"] = v
        if self.norm is not None:
            d["norm.mean"], d["norm.std"] = self.norm
        return d

    def load_state_dict(self, d):
        """Loads every layer present in ``d``.  Weights saved before the residual block and the FiLM layers
        existed leave those at their initial values (an identity block, zero modulation): the same function."""
        for k, c in self.conv.items():
            if f"{k}.w" in d:
                c.load({"w": d[f"{k}.w"], "b": d[f"{k}.b"]})
            elif k not in ("c4c", "c4d"):
                raise KeyError(f"weights file lacks layer {k}")
        for k, m in (("film1", self.film1), ("film2", self.film2)):
            if f"{k}.w" in d:
                m.load({"w": d[f"{k}.w"], "b": d[f"{k}.b"]})
        if "norm.mean" in d:
            self.norm = (np.asarray(d["norm.mean"], np.float32), np.asarray(d["norm.std"], np.float32))

    def save(self, path):
        np.savez_compressed(path, **{k: (v.astype(np.float16) if k.endswith(".w") and not k.startswith("film")
                                         else v) for k, v in self.state_dict().items()})

    @classmethod
    def load_file(cls, path, device="cuda"):
        net = cls(device=device)
        with np.load(path) as z:
            net.load_state_dict({k: z[k].astype(np.float32) for k in z.files})
        return net


def mse_loss(pred, target, wmask):
    """Weighted MSE on the first NOUT channels; returns (loss (device scalar), grad (bf16, pred's shape)), the
    normaliser (the sum of wmask) kept on the device."""
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
    grad = torch.empty(pred.shape, dtype=torch.bfloat16, device=pred.device)
    wsum = wmask.sum(dtype=torch.float32).reshape(1)  # mean over weighted entries
    _ok(L.xnn_mse_bf16(_p(pred), _p(target), _p(wmask), _p(wsum), _p(grad), _p(part), V, C, NOUT, _st()),
        "xnn_mse_bf16")
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


def targets(mask, rho, u, T, u_in, P, pad_to=None, t_weight=1.0):
    """(X8, Y8, Z8, 5) raw targets and (.., 5) loss weights for one solved design: u and rho weighted 1 on
    non-solid cells, T weighted ``t_weight`` on every cell (the thermal field sets how long a warm run takes
    to converge)."""
    nx, ny, nz = mask.shape
    X, Y, Z = pad_to or (round8(nx), round8(ny), round8(nz))
    t = np.zeros((X, Y, Z, NOUT), np.float32)
    w = np.zeros((X, Y, Z, NOUT), np.float32)
    fluid = (mask != BC_SOLID).astype(np.float32)
    t[:nx, :ny, :nz, 0:3] = u / u_in
    t[:nx, :ny, :nz, 3] = (rho - 1.0) / u_in
    t[:nx, :ny, :nz, 4] = (T - T_IN) / P
    w[:nx, :ny, :nz, 0:4] = fluid[..., None]
    w[:nx, :ny, :nz, 4] = t_weight
    return t, w
