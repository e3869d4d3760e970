"""The surrogate at its training size: 152 x 128 x 32 inputs (148 x 128 x 32 lattices padded), batch 8.

* One forward, weighted MSE and backward of the shipped network on a real
  solver sample (a full-size design advanced by the batched solver), against
  torch.autograd on a CPU FP32 copy with the same BF16 rounding points: the
  output, the loss, and every layer's weight and bias gradients, at the
  bounds of the small-lattice network test.
* A few Adam steps at batch 8 on eight solver samples lower the loss.
"""
import numpy as np
import pytest
import torch

from test_gpu_surrogate import _torch_net, randomize_zero_layers, rel, torch_weight

pytestmark = pytest.mark.gpu


def _samples(n, steps=2000):
    from xadvcool_amd import warmstart as W
    from xadvcool_amd.batch import BatchedSolver
    designs = W.make_designs(W.COMPARE_SEED0 + 9000, n)
    bs = BatchedSolver(n, (148, 128, 32), tol_ke=0.0, tol_T=0.0, max_steps=steps, collect_fields=True,
                       metrics=False)
    try:
        res = {r["key"]: r for r in bs.run(designs)}
    finally:
        bs.close()
    return [W._sample(d, res[d["key"]]["fields"]) for d in designs]


def test_full_size_forward_and_gradients(gpu):
    from xadvcool_amd import surrogate as S
    from xadvcool_amd import warmstart as W
    net = S.UNet3D.load_file(W.DEFAULT_WEIGHTS)
    randomize_zero_layers(net)  # any layer still at its zero start gets values, so every gradient counts
    (x, t, w, sc), = _samples(1)
    assert tuple(x.shape) == (152, 128, 32, S.CIN)
    mean, std = (torch.tensor(a, device="cuda") for a in net.norm)
    tn = ((t - mean) / std).contiguous()
    print("[config4] full-size forward + backward, batch 1", flush=True)
    out = net.forward(x[None].contiguous(), sc[None].contiguous(), keep=True)
    loss, g = S.mse_loss(out, tn[None].contiguous(), w[None].contiguous())
    net.zero_grad()
    net.backward(g)
    ref, Ws, Bs, D = _torch_net(net, x[None].float().cpu(), sc[None].cpu())
    assert rel(out.cpu(), ref.detach()) < 2e-2
    tc, wc = tn[None].cpu(), w[None].cpu()
    lref = ((ref[..., :S.NOUT] - tc) ** 2 * wc).sum() / wc.sum()
    lref.backward()
    assert abs(loss.item() - lref.item()) / lref.item() < 2e-2
    grads = [(k, torch_weight(c.dw.cpu(), c.cin, c.cout), Ws[k].grad, c.db.cpu(), Bs[k].grad)
             for k, c in net.conv.items()]
    grads += [(k, getattr(net, k).dw.cpu(), D[k][0].grad, getattr(net, k).db.cpu(), D[k][1].grad)
              for k in ("film1", "film2")]
    for k, gw, rw, gb, rb in grads:
        e = rel(gw, rw)
        cos = float(torch.nn.functional.cosine_similarity(gw.flatten(), rw.flatten(), dim=0))
        print(f"[config4] grad {k}: rel {e:.3e} cos {cos:.5f}", flush=True)
        assert e < 6e-2 and cos > 0.998, k
        assert rel(gb, rb) < 6e-2, k


def test_batch8_adam_steps_lower_the_loss(gpu):
    from xadvcool_amd import surrogate as S
    from xadvcool_amd import warmstart as W
    net = S.UNet3D.load_file(W.DEFAULT_WEIGHTS)
    samples = _samples(8)
    mean, std = (torch.tensor(a, device="cuda") for a in net.norm)
    x, t, w, sc = W._stack(samples, range(len(samples)), mean, std)
    losses = []
    for _ in range(8):
        out = net.forward(x, sc, keep=True)
        loss, g = S.mse_loss(out, t, w)
        net.zero_grad()
        net.backward(g)
        net.adam_step(2e-4)
        losses.append(loss.item())
    print(f"[config4] batch-8 losses {np.round(losses, 5).tolist()}", flush=True)
    assert all(np.isfinite(losses)) and losses[-1] < losses[0]
