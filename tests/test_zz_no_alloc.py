"""The allocation-free bias gradient (xnn_bias_grad_noalloc) against the workspace entry point (xnn_bias_grad).

Both run the same two kernels; the allocation-free one keeps its per-block
partials in the library's static device scratch.  Two calls in a row, the
second smaller, must give the workspace entry point's bits; shapes the
scratch cannot hold, or that the kernels do not take, are refused.
"""
import ctypes

import pytest
import torch

pytestmark = pytest.mark.gpu


def _lib():
    from xadvcool_amd import surrogate as S
    L = S.lib()
    L.xnn_bias_grad_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                                        ctypes.c_void_p]
    L.xnn_bias_grad_noalloc.restype = ctypes.c_int
    return L


def _grad(L, g, C, noalloc):
    V = g.numel() // C
    db = torch.full((C,), float("nan"), device="cuda")
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if noalloc:
        rc = L.xnn_bias_grad_noalloc(ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(db.data_ptr()), V, C, st)
    else:
        part = torch.full((L.xnn_bias_part_floats(V, C),), float("nan"), device="cuda")
        rc = L.xnn_bias_grad(ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(db.data_ptr()),
                             ctypes.c_void_p(part.data_ptr()), V, C, st)
    torch.cuda.synchronize()
    assert rc == 0
    return db.cpu()


def test_noalloc_bias_grad_matches_workspace_entry_point(gpu):
    L = _lib()
    gen = torch.Generator().manual_seed(3)
    for shape in ((4, 40, 32, 32, 64), (1, 19, 16, 4, 16)):  # the second call smaller
        C = shape[-1]
        g = torch.randn(*shape, generator=gen).to(torch.bfloat16).cuda()
        want = _grad(L, g, C, noalloc=False)
        got = _grad(L, g, C, noalloc=True)
        print(f"[noalloc] {shape}: max |db| {float(want.abs().max()):.3f}", flush=True)
        assert torch.equal(got, want), shape


def test_noalloc_bias_grad_refuses_unsupported_channels(gpu):
    L = _lib()
    g = torch.zeros(1000 * 12, dtype=torch.bfloat16, device="cuda")
    db = torch.zeros(12, device="cuda")
    rc = L.xnn_bias_grad_noalloc(ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(db.data_ptr()), 1000, 12, None)
    assert rc == 1
