"""The surrogate's HIP kernels (libxadv_nn.so) against PyTorch CPU FP32 as the oracle.

Every operand the kernels see in BF16 is rounded to BF16 on the CPU side too;
the CPU then computes in FP32, so differences are FP32 summation order only
(conv forward / data grad / weight grad: rtol 1e-3 on the tensor norm).  The
whole network's forward and parameter gradients are checked against
torch.autograd on a CPU copy with the same BF16 rounding points, and a few
Adam steps must lower the loss.
"""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

pytestmark = pytest.mark.gpu


def bf(x):
    return x.to(torch.bfloat16).to(torch.float32)


def rel(a, b):
    a, b = a.double(), b.double()
    return float((a - b).norm() / max(b.norm(), 1e-30))


def torch_weight(W, cin, cout):
    """Our W[co][tap*cin + ci] (tap = kx*9 + ky*3 + kz) -> torch conv3d weight [co, ci, kx, ky, kz]."""
    return W[:, :27 * cin].reshape(cout, 27, cin).permute(0, 2, 1).reshape(cout, cin, 3, 3, 3)


def ncdhw(x):
    return x.permute(0, 4, 1, 2, 3)


def ndhwc(x):
    return x.permute(0, 2, 3, 4, 1)


@pytest.mark.parametrize("cin,cout", [(16, 16), (16, 32), (32, 64), (64, 128), (128, 16), (32, 32)])
def test_conv_fwd_dgrad_wgrad(gpu, cin, cout):
    from xadvcool_amd import surrogate as S
    print(f"[nn] conv {cin}->{cout}", flush=True)
    gen = torch.Generator().manual_seed(cin * 7 + cout)
    shape = (2, 8, 6, 10)
    c = S.Conv(cin, cout, gen)
    c.b.copy_(torch.randn(cout, generator=gen))
    x = torch.randn(*shape, cin, generator=gen)
    xb = x.to(torch.bfloat16).cuda()
    out = c.fwd(xb, shape)
    W = c.w.cpu()
    ref = ndhwc(F.conv3d(ncdhw(bf(x)), torch_weight(bf(W), cin, cout), bias=c.b.cpu(), padding=1))
    assert rel(out.cpu(), ref) < 1e-3
    # backward
    g = torch.randn(*shape, cout, generator=gen)
    gb = g.to(torch.bfloat16).cuda()
    gin = c.dgrad(gb, shape)
    xr = ncdhw(bf(x)).clone().requires_grad_(True)
    wr = torch_weight(bf(W), cin, cout).clone().requires_grad_(True)
    br = c.b.cpu().clone().requires_grad_(True)
    o = F.conv3d(xr, wr, bias=br, padding=1)
    o.backward(ncdhw(bf(g)))
    assert rel(gin.cpu(), ndhwc(xr.grad)) < 1e-3
    c.zero_grad()
    c.wgrad(xb, gb, shape, vchunk=96)
    dw = torch_weight(c.dw.cpu(), cin, cout)
    assert rel(dw, wr.grad) < 1e-3
    assert rel(c.db.cpu(), br.grad) < 1e-3
    assert torch.all(c.dw.cpu()[:, 27 * cin:] == 0)


@pytest.mark.parametrize("cin,cout,shape", [(16, 16, (2, 3, 32, 32)), (32, 16, (2, 3, 24, 32)),
                                             (16, 32, (2, 3, 32, 32)), (16, 32, (2, 3, 16, 16)),
                                             (32, 32, (2, 3, 16, 16)), (32, 16, (2, 3, 16, 16)),
                                             (32, 64, (2, 3, 16, 16)), (64, 32, (2, 3, 16, 16)),
                                             (32, 64, (2, 3, 16, 8)), (64, 64, (2, 3, 16, 8)),
                                             (128, 64, (2, 3, 16, 8))])
def test_conv_lds_paths_match_reference_and_gather_kernel(gpu, cin, cout, shape):
    """Convolutions on lattices with Z = 32, 16 or 8 (Y a multiple of 16, or of 8 for 32 -> 16 at
    Z = 32) take the LDS-halo kernels (k_conv_fwd16_lds, k_conv_fwd_lds), forward and data gradient
    (the 128 -> 64 case covers the 64 -> 128 data gradient).

    Against PyTorch (rtol 1e-3), and bit for bit against the gather kernels: zero-padding z by
    two planes sends the same convolution through k_conv_fwd16 / k_conv_fwd (Z is then none of 8,
    16, 32), and the padded result's inner z planes must equal the LDS kernel's output exactly
    (same MFMA sequence per output element).  (The 128 -> 64 forward takes the split-K kernel with
    and without the padding; its padded run must give the same bits too.)"""
    from xadvcool_amd import surrogate as S
    print(f"[nn] conv {cin}->{cout} LDS-halo path {shape}", flush=True)
    gen = torch.Generator().manual_seed(1600 + 7 * cin + cout + shape[3])
    N, X, Y, Z = shape
    c = S.Conv(cin, cout, gen)
    c.b.copy_(torch.randn(cout, generator=gen))
    x = torch.randn(*shape, cin, generator=gen)
    xb = x.to(torch.bfloat16).cuda()
    out = c.fwd(xb, shape).cpu()
    W = c.w.cpu()
    ref = ndhwc(F.conv3d(ncdhw(bf(x)), torch_weight(bf(W), cin, cout), bias=c.b.cpu(), padding=1))
    assert rel(out, ref) < 1e-3
    xp = torch.zeros(N, X, Y, Z + 2, cin, dtype=torch.bfloat16, device="cuda")
    xp[:, :, :, 1:Z + 1] = xb
    outp = c.fwd(xp, (N, X, Y, Z + 2)).cpu()
    assert torch.equal(outp[:, :, :, 1:Z + 1], out)
    g = torch.randn(*shape, cout, generator=gen)
    gb = g.to(torch.bfloat16).cuda()
    gin = c.dgrad(gb, shape).cpu()
    xr = ncdhw(bf(x)).clone().requires_grad_(True)
    o = F.conv3d(xr, torch_weight(bf(W), cin, cout), bias=c.b.cpu(), padding=1)
    o.backward(ncdhw(bf(g)))
    assert rel(gin, ndhwc(xr.grad)) < 1e-3
    gp = torch.zeros(N, X, Y, Z + 2, cout, dtype=torch.bfloat16, device="cuda")
    gp[:, :, :, 1:Z + 1] = gb
    ginp = c.dgrad(gp, (N, X, Y, Z + 2)).cpu()
    assert torch.equal(ginp[:, :, :, 1:Z + 1], gin)


@pytest.mark.parametrize("shape,cin,cout", [((2, 5, 7, 32), 16, 16), ((2, 5, 7, 32), 32, 16),
                                             ((2, 3, 6, 16), 16, 32), ((2, 3, 6, 16), 32, 32),
                                             ((1, 4, 9, 16), 64, 32), ((2, 3, 5, 8), 32, 64),
                                             ((1, 4, 4, 8), 64, 64), ((1, 3, 5, 8), 128, 64),
                                             ((2, 3, 6, 4), 64, 128), ((1, 5, 4, 4), 128, 128)])
def test_wgrad_tiled_shapes(gpu, shape, cin, cout):
    """The z-column-tiled weight gradient (Z in {4, 8, 16, 32}, the network's channel pairs at those levels),
    with partial tiles at the x and y edges, against torch's conv3d weight and bias gradients."""
    from xadvcool_amd import surrogate as S
    print(f"[nn] tiled wgrad {shape} {cin}->{cout}", flush=True)
    N, X, Y, Z = shape
    gen = torch.Generator().manual_seed(cin + 3 * cout + Z)
    c = S.Conv(cin, cout, gen)
    x = torch.randn(*shape, cin, generator=gen)
    g = torch.randn(*shape, cout, generator=gen)
    xr = ncdhw(bf(x)).clone().requires_grad_(True)
    wr = torch_weight(bf(c.w.cpu()), cin, cout).clone().requires_grad_(True)
    br = torch.zeros(cout, requires_grad=True)
    F.conv3d(xr, wr, bias=br, padding=1).backward(ncdhw(bf(g)))
    for rep in range(2):  # accumulation across calls, partial buffers reused
        c.wgrad(x.to(torch.bfloat16).cuda(), g.to(torch.bfloat16).cuda(), shape)
    assert rel(torch_weight(c.dw.cpu(), cin, cout), 2 * wr.grad) < 1e-3
    assert rel(c.db.cpu(), 2 * br.grad) < 1e-3
    assert torch.all(c.dw.cpu()[:, 27 * cin:] == 0)


def test_pool_up_silu(gpu):
    from xadvcool_amd import surrogate as S
    gen = torch.Generator().manual_seed(3)
    x = torch.randn(2, 8, 4, 6, 16, generator=gen)
    xb = x.to(torch.bfloat16).cuda()
    p = S.pool_fwd(xb, (2, 4, 2, 3))
    pref = ndhwc(F.avg_pool3d(ncdhw(bf(x)), 2))
    assert rel(p.float().cpu(), pref) < 5e-3
    u = S.up_fwd(p, (2, 8, 4, 6))
    uref = p.float().cpu().repeat_interleave(2, 1).repeat_interleave(2, 2).repeat_interleave(2, 3)
    assert torch.equal(u.float().cpu(), uref)
    g = torch.randn(2, 8, 4, 6, 16, generator=gen).cuda()
    gc = S.up_bwd(g, (2, 4, 2, 3))
    gref = g.cpu().reshape(2, 4, 2, 2, 2, 3, 2, 16).sum(dim=(2, 4, 6))
    assert rel(gc.cpu(), gref) < 1e-6
    gf = S.pool_bwd(gc, (2, 8, 4, 6))
    assert rel(gf.cpu(), gc.cpu().repeat_interleave(2, 1).repeat_interleave(2, 2).repeat_interleave(2, 3) / 8) < 1e-6
    xf = x.cuda()
    y = S.silu_fwd(xf)
    assert rel(y.float().cpu(), bf(F.silu(x))) < 5e-3
    gy = torch.randn_like(x).cuda()
    gx = S.silu_bwd(xf, gy)
    xr = x.clone().requires_grad_(True)
    F.silu(xr).backward(gy.cpu())
    assert rel(gx.float().cpu(), xr.grad) < 5e-3


def _offset_copy(t):
    """The same values one element into a fresh buffer: the pointer is then not 16-byte aligned, so the
    launchers take the scalar elementwise kernels instead of the vectorised ones."""
    buf = torch.empty(t.numel() + 8, dtype=t.dtype, device=t.device)
    v = buf[1:1 + t.numel()].view(t.shape)
    v.copy_(t)
    return v


@pytest.mark.parametrize("shape", [(2, 8, 4, 6, 16), (1, 4, 8, 2, 32)])
def test_vectorised_elementwise_kernels_match_scalar(gpu, shape):
    """Pooling, upsampling, SiLU and the adds: the vectorised kernels give the scalar kernels' bits
    (also on element counts that leave a tail: 1001 and 4099 values)."""
    from xadvcool_amd import surrogate as S
    gen = torch.Generator().manual_seed(11)
    N, X, Y, Z, C = shape
    cs = (N, X // 2, Y // 2, Z // 2)
    xb = torch.randn(*shape, generator=gen).to(torch.bfloat16).cuda()
    p = S.pool_fwd(xb, cs)
    assert torch.equal(p.cpu(), S.pool_fwd(_offset_copy(xb), cs).cpu())
    assert torch.equal(S.up_fwd(p, shape[:4]).cpu(), S.up_fwd(_offset_copy(p), shape[:4]).cpu())
    g = torch.randn(*shape, generator=gen).cuda()
    gc = S.up_bwd(g, cs)
    assert torch.equal(gc.cpu(), S.up_bwd(_offset_copy(g), cs).cpu())
    assert torch.equal(S.pool_bwd(gc, shape[:4]).cpu(), S.pool_bwd(_offset_copy(gc), shape[:4]).cpu())
    # the fused skip-gradient sum: the bits of add_f32(acc, pool_bwd(...)), vectorised and scalar
    acc = torch.randn(*shape, generator=gen).cuda()
    two = S.add_f32(acc, S.pool_bwd(gc, shape[:4]))
    assert torch.equal(S.pool_bwd_add(gc, acc, shape[:4]).cpu(), two.cpu())
    assert torch.equal(S.pool_bwd_add(_offset_copy(gc), acc, shape[:4]).cpu(), two.cpu())
    for n in (g.numel(), 1001, 4099):
        x = torch.randn(n, generator=gen).cuda() * 4
        gy = torch.randn(n, generator=gen).cuda()
        assert torch.equal(S.silu_fwd(x).cpu(), S.silu_fwd(_offset_copy(x)).cpu())
        assert torch.equal(S.silu_bwd(x, gy).cpu(), S.silu_bwd(_offset_copy(x), gy).cpu())
        a, b = x.to(torch.bfloat16), gy.to(torch.bfloat16)
        assert torch.equal(S.add_bf16(a, b).cpu(), S.add_bf16(_offset_copy(a), b).cpu())
        assert torch.equal(S.add_f32(x, gy).cpu(), S.add_f32(_offset_copy(x), gy).cpu())


def _torch_net(net, x, sc):
    """CPU FP32 copy of UNet3D.forward with the same BF16 rounding points (autograd-capable).
    Returns (output, conv weights, conv biases, FiLM MLP parameters)."""
    Ws = {k: torch_weight(bf(c.w.cpu()), c.cin, c.cout).clone().requires_grad_(True) for k, c in net.conv.items()}
    Bs = {k: c.b.cpu().clone().requires_grad_(True) for k, c in net.conv.items()}
    D = {k: (getattr(net, k).w.cpu().clone().requires_grad_(True), getattr(net, k).b.cpu().clone().requires_grad_(True))
         for k in ("film1", "film2")}
    film = F.silu(sc @ D["film1"][0].T + D["film1"][1]) @ D["film2"][0].T + D["film2"][1]
    sites = {name: (o, C) for name, o, C in net.FILM}

    def cv(k, h, site=None):
        p = F.conv3d(h, Ws[k], bias=Bs[k], padding=1)
        if site is not None:  # FiLM: per-sample, per-channel scale and shift of the pre-activation
            o, C = sites[site]
            p = p * (1.0 + film[:, o:o + C, None, None, None]) + film[:, o + C:o + 2 * C, None, None, None]
        return p

    def act(h):  # SiLU output stored in BF16
        return bf(F.silu(h))

    h = ncdhw(bf(x))
    s1 = act(cv("c1b", act(cv("c1a", h))))
    q1 = bf(F.avg_pool3d(s1, 2))
    s2 = act(cv("c2b", act(cv("c2a", q1))))
    q2 = bf(F.avg_pool3d(s2, 2))
    s3 = act(cv("c3b", act(cv("c3a", q2))))
    q3 = bf(F.avg_pool3d(s3, 2))
    a4 = act(cv("c4b", act(cv("c4a", q3)), "p4b"))
    r4 = bf(a4 + act(cv("c4d", act(cv("c4c", a4)))))
    up = lambda t: t.repeat_interleave(2, 2).repeat_interleave(2, 3).repeat_interleave(2, 4)
    d3 = act(cv("d3b", bf(act(cv("d3a", up(r4))) + s3), "pd3b"))
    d2 = act(cv("d2b", bf(act(cv("d2a", up(d3))) + s2), "pd2b"))
    d1 = act(cv("d1b", bf(act(cv("d1a", up(d2))) + s1), "pd1b"))
    return ndhwc(cv("head", d1)), Ws, Bs, D


def randomize_zero_layers(net, seed=0):
    """Give the layers that start at zero (the residual block's last convolution, the FiLM MLP's output) small
    random values, so that every gradient of a test is non-trivial."""
    gen = torch.Generator().manual_seed(seed)
    c = net.conv["c4d"]
    w = torch.zeros(c.cout, c.Kp)
    w[:, :27 * c.cin] = torch.randn(c.cout, 27 * c.cin, generator=gen) * 0.02
    c.load({"w": w, "b": torch.randn(c.cout, generator=gen) * 0.02})
    net.film2.load({"w": torch.randn(net.film2.fout, net.film2.fin, generator=gen) * 0.05,
                    "b": torch.randn(net.film2.fout, generator=gen) * 0.05})


def _scalars(n, seed):
    from xadvcool_amd import surrogate as S
    rng = np.random.default_rng(seed)
    return torch.from_numpy(np.stack([S.scalars(rng.uniform(0.52, 0.56), rng.uniform(0.05, 0.2),
                                                rng.uniform(0.01, 0.05)) for _ in range(n)]))


@pytest.mark.parametrize("shape", [(2, 16, 8, 8), (1, 8, 32, 32)])
def test_network_forward_and_gradients(gpu, shape):
    """The whole network against torch.autograd.  The (1, 8, 32, 32) lattice takes the LDS-halo
    convolutions at the two finest levels (Z = 32 and 16) and the vectorised elementwise kernels."""
    from xadvcool_amd import surrogate as S
    print(f"[nn] network forward and gradients on {shape}", flush=True)
    torch.manual_seed(0)
    net = S.UNet3D(seed=1)
    randomize_zero_layers(net)
    x = torch.randn(*shape, 16) * 0.5
    sc = _scalars(shape[0], 3)
    tgt = torch.randn(*shape, 5)
    w = (torch.rand(*shape, 5) > 0.3).float()
    out = net.forward(x.to(torch.bfloat16).cuda(), sc.cuda(), keep=True)
    ref, Ws, Bs, D = _torch_net(net, x, sc)
    assert rel(out.cpu(), ref.detach()) < 2e-2
    loss, g = S.mse_loss(out, tgt.cuda(), w.cuda())
    net.zero_grad()
    net.backward(g)
    lref = ((ref[..., :5] - tgt) ** 2 * w).sum() / w.sum()
    lref.backward()
    assert abs(loss.item() - lref.item()) / lref.item() < 2e-2
    grads = [(k, torch_weight(c.dw.cpu(), c.cin, c.cout), Ws[k].grad, c.db.cpu(), Bs[k].grad)
             for k, c in net.conv.items()]
    grads += [(k, getattr(net, k).dw.cpu(), D[k][0].grad, getattr(net, k).db.cpu(), D[k][1].grad)
              for k in ("film1", "film2")]
    for k, gw, rw, gb, rb in grads:
        e = rel(gw, rw)
        cos = float(torch.nn.functional.cosine_similarity(gw.flatten(), rw.flatten(), dim=0))
        print(f"[nn] grad {k}: rel {e:.3e} cos {cos:.5f}", flush=True)
        assert e < 6e-2 and cos > 0.998, k
        assert rel(gb, rb) < 6e-2, k


def test_film_and_dense_kernels(gpu):
    """FiLM modulation and the MLP's dense layers against torch.autograd in FP32 (summation order only)."""
    from xadvcool_amd import surrogate as S
    gen = torch.Generator().manual_seed(5)
    N, Vn, C, L, off = 3, 5000, 32, 96, 16
    p = torch.randn(N, 10, 25, 20, C, generator=gen)
    film = torch.randn(N, L, generator=gen) * 0.3
    gy = torch.randn(p.shape, generator=gen)
    y = S.film_fwd(p.cuda(), film.cuda(), off)
    pr, fr = p.clone().requires_grad_(True), film.clone().requires_grad_(True)
    yr = pr * (1 + fr[:, None, None, None, off:off + C]) + fr[:, None, None, None, off + C:off + 2 * C]
    assert rel(y.cpu(), yr.detach()) < 1e-6
    yr.backward(gy)
    dfilm = torch.zeros(N, L, device="cuda")
    gx = S.film_bwd(p.cuda(), gy.cuda(), film.cuda(), dfilm, off)
    assert rel(gx.cpu(), pr.grad) < 1e-6
    assert rel(dfilm.cpu()[:, off:off + 2 * C], fr.grad[:, off:off + 2 * C]) < 1e-5
    assert float(dfilm.cpu()[:, :off].abs().max()) == 0.0  # other sites' entries untouched
    # the fused FiLM + SiLU kernels give the bits of the separate kernels
    pc, fc, gc = p.cuda(), film.cuda(), gy.cuda()
    assert torch.equal(S.film_silu_fwd(pc, fc, off).cpu(), S.silu_fwd(S.film_fwd(pc, fc, off)).cpu())
    d1, d2 = torch.zeros(N, L, device="cuda"), torch.zeros(N, L, device="cuda")
    fused = S.film_silu_bwd(pc, gc, fc, d1, off)
    sep = S.to_bf16(S.film_bwd(pc, S.silu_bwd_f32(S.film_fwd(pc, fc, off), gc), fc, d2, off))
    assert torch.equal(fused.cpu(), sep.cpu()) and torch.equal(d1.cpu(), d2.cpu())
    for act in (True, False):
        d = S.Dense(7, 40, gen, act=act)
        d.load({"w": torch.randn(40, 7, generator=gen), "b": torch.randn(40, generator=gen)})
        xi = torch.randn(5, 7, generator=gen)
        pre, post = d.fwd(xi.cuda())
        wr, br = d.w.cpu().clone().requires_grad_(True), d.b.cpu().clone().requires_grad_(True)
        xr = xi.clone().requires_grad_(True)
        pr2 = xr @ wr.T + br
        outr = F.silu(pr2) if act else pr2
        assert rel(post.cpu(), outr.detach()) < 1e-6 and rel(pre.cpu(), pr2.detach()) < 1e-6
        g = torch.randn(5, 40, generator=gen)
        outr.backward(g)
        d.zero_grad()
        gxi = d.bwd(xi.cuda(), pre, g.cuda())
        assert rel(d.dw.cpu(), wr.grad) < 1e-5 and rel(d.db.cpu(), br.grad) < 1e-5
        assert rel(gxi.cpu(), xr.grad) < 1e-5


@pytest.mark.parametrize("C,shape", [(16, (2, 9, 8, 32)), (32, (3, 5, 7, 16)), (64, (2, 3, 4, 8)),
                                     (128, (8, 3, 4, 4)), (128, (1, 1, 1, 1))])
def test_activation_backward_in_one_pass(gpu, C, shape):
    """xnn_act_bwd (SiLU, and FiLM then SiLU): the BF16 gradient is bit for bit that of the elementwise kernels,
    dfilm that of film_silu_bwd and the bias gradient that of xnn_bias_grad on the same BF16 gradient, up to FP32
    summation order; the voxel counts leave partial blocks."""
    from xadvcool_amd import surrogate as S
    print(f"[nn] activation backward C={C} {shape}", flush=True)
    gen = torch.Generator().manual_seed(C + shape[0])
    N, L, off = shape[0], 2 * C + 40, 24
    p = (torch.randn(*shape, C, generator=gen) * 2).cuda()
    gy = torch.randn(*shape, C, generator=gen).cuda()
    film = (torch.randn(N, L, generator=gen) * 0.3).cuda()
    gx, db = S.act_bwd(p, gy)
    assert torch.equal(gx, S.silu_bwd(p, gy))
    ref_db = gx.double().reshape(-1, C).sum(0).float()
    assert rel(db.cpu(), ref_db.cpu()) < 1e-5
    d1, d2 = torch.zeros(N, L, device="cuda"), torch.zeros(N, L, device="cuda")
    gxf, dbf = S.act_bwd(p, gy, film, d1, off)
    assert torch.equal(gxf, S.film_silu_bwd(p, gy, film, d2, off))
    assert rel(d1.cpu()[:, off:off + 2 * C], d2.cpu()[:, off:off + 2 * C]) < 1e-5
    assert float(d1[:, :off].abs().max()) == 0.0 and float(d1[:, off + 2 * C:].abs().max()) == 0.0
    assert rel(dbf.cpu(), gxf.double().reshape(-1, C).sum(0).float().cpu()) < 1e-5
    gx2, db2 = S.act_bwd(p, gy, want_db=False)
    assert db2 is None and torch.equal(gx2, gx)


@pytest.mark.parametrize("cin,cout,shape", [(64, 128, (1, 3, 5, 3)), (128, 128, (2, 5, 4, 4)),
                                             (128, 64, (1, 3, 5, 8)), (128, 128, (8, 19, 16, 4)),
                                             (128, 32, (1, 2, 3, 4)), (64, 16, (1, 2, 2, 2)),
                                             (128, 64, (4, 38, 32, 8))])
def test_conv_split_k_coarse_levels(gpu, cin, cout, shape):
    """The split-K convolution of the coarsest level (Z <= 4) and of 128 input channels, forward and data
    gradient, against PyTorch (rtol 1e-3), with partial 16-voxel tiles, the network's bottleneck at batch 8
    (two 16-voxel tiles per block) and a 128 -> 64 layer with four tiles per block."""
    from xadvcool_amd import surrogate as S
    print(f"[nn] split-K conv {cin}->{cout} {shape}", flush=True)
    gen = torch.Generator().manual_seed(cin + cout + sum(shape))
    c = S.Conv(cin, cout, gen)
    c.b.copy_(torch.randn(cout, generator=gen))
    x = torch.randn(*shape, cin, generator=gen)
    out = c.fwd(x.to(torch.bfloat16).cuda(), shape).cpu()
    W = torch_weight(bf(c.w.cpu()), cin, cout)
    xr = ncdhw(bf(x)).clone().requires_grad_(True)
    o = F.conv3d(xr, W, bias=c.b.cpu(), padding=1)
    assert rel(out, ndhwc(o.detach())) < 1e-3
    g = torch.randn(*shape, cout, generator=gen)
    gin = c.dgrad(g.to(torch.bfloat16).cuda(), shape).cpu()
    o.backward(ncdhw(bf(g)))
    assert rel(gin, ndhwc(xr.grad)) < 1e-3


@pytest.mark.parametrize("cin,cout,shape", [(16, 16, (2, 3, 16, 32)), (32, 16, (2, 3, 8, 32)),
                                             (16, 32, (2, 3, 16, 16)), (32, 32, (1, 3, 16, 16)),
                                             (64, 64, (1, 3, 8, 8)), (64, 128, (1, 3, 5, 4)),
                                             (16, 16, (1, 3, 5, 6))])
def test_conv_with_silu_epilogue_matches_conv_then_silu(gpu, cin, cout, shape):
    """xnn_conv_fwd_silu gives the bits of xnn_conv_fwd followed by xnn_silu_fwd: in the LDS-halo kernels'
    epilogue (the first five shapes) and through the separate pass (split-K and gather kernels)."""
    from xadvcool_amd import surrogate as S
    print(f"[nn] conv+SiLU {cin}->{cout} {shape}", flush=True)
    gen = torch.Generator().manual_seed(77 + cin + cout + shape[2])
    c = S.Conv(cin, cout, gen)
    c.b.copy_(torch.randn(cout, generator=gen))
    x = torch.randn(*shape, cin, generator=gen).to(torch.bfloat16).cuda()
    p_ref = c.fwd(x, shape)
    a_ref = S.silu_fwd(p_ref)
    p, a = c.fwd_silu(x, shape)
    assert torch.equal(p, p_ref)
    assert torch.equal(a, a_ref)


def test_mse_gradient_in_bf16(gpu):
    """mse_loss: the loss and the BF16 gradient (normaliser read on the device) against torch in FP32."""
    from xadvcool_amd import surrogate as S
    gen = torch.Generator().manual_seed(9)
    pred = torch.randn(2, 5, 6, 7, 16, generator=gen)
    tgt = torch.randn(2, 5, 6, 7, S.NOUT, generator=gen)
    w = (torch.rand(2, 5, 6, 7, S.NOUT, generator=gen) > 0.4).float()
    loss, g = S.mse_loss(pred.cuda(), tgt.cuda(), w.cuda())
    pr = pred.clone().requires_grad_(True)
    lr = ((pr[..., :S.NOUT] - tgt) ** 2 * w).sum() / w.sum()
    lr.backward()
    assert abs(loss.item() - lr.item()) / lr.item() < 1e-5
    assert g.dtype == torch.bfloat16
    assert torch.equal(g.float().cpu(), bf(pr.grad)) or rel(g.float().cpu(), pr.grad) < 4e-3
    assert float(g[..., S.NOUT:].abs().max()) == 0.0


def test_old_weights_load_as_the_same_function(gpu, tmp_path):
    """Weights saved without the residual block and the FiLM layers load with the block at identity and no
    modulation: the output equals that of a network whose new layers are explicitly zero."""
    from xadvcool_amd import surrogate as S
    net = S.UNet3D(seed=4)
    d = {k: v for k, v in net.state_dict().items() if not k.startswith(("c4c.", "c4d.", "film"))}
    np.savez(tmp_path / "old.npz", **d)
    old = S.UNet3D.load_file(str(tmp_path / "old.npz"))
    x = (torch.randn(1, 16, 8, 8, 16) * 0.5).to(torch.bfloat16).cuda()
    sc = _scalars(1, 4).cuda()
    a = old.forward(x, sc)
    net.conv["c4c"].load({"w": np.zeros_like(old.conv["c4c"].w.cpu().numpy()), "b": np.zeros(128, np.float32)})
    b = net.forward(x, sc)
    assert torch.equal(a, b)


def test_adam_lowers_the_loss(gpu):
    from xadvcool_amd import surrogate as S
    torch.manual_seed(1)
    net = S.UNet3D(seed=2)
    x = (torch.randn(2, 16, 8, 8, 16) * 0.5).to(torch.bfloat16).cuda()
    sc = _scalars(2, 6).cuda()
    tgt = torch.randn(2, 16, 8, 8, 5).cuda() * 0.3
    w = torch.ones(2, 16, 8, 8, 5).cuda()
    losses = []
    for _ in range(30):
        out = net.forward(x, sc, keep=True)
        loss, g = S.mse_loss(out, tgt, w)
        net.zero_grad()
        net.backward(g)
        net.adam_step(2e-3)
        losses.append(loss.item())
    print(f"[nn] adam losses {losses[0]:.4f} -> {losses[-1]:.4f}", flush=True)
    assert losses[-1] < 0.5 * losses[0]
