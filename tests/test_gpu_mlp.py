"""Surrogate network kernels on the GPU (BF16 matrix cores, FP32 accumulation)
against the float64 oracle with the same bfloat16 rounding points
(oracle/mlp_oracle.py, quantize=True) and a torch float32 reference; the AdamW
step against the oracle's; training drives the loss down."""
import numpy as np
import pytest

from mcluminescence_amd import surrogate as S
from mcluminescence_amd._abi import AdamW, MclError
from oracle import mlp_oracle as O

pytestmark = pytest.mark.gpu

FULL = O.Desc()   # 6 -> 512 x4 (SiLU) -> 256


def _net(ctx, d=FULL, max_batch=4096, seed=0):
    return S.Network(ctx, d.n_in, d.n_hidden, d.n_hidden_layers, d.n_out, max_batch,
                     params=O.init_params(d, seed=seed))


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def _torch_reference(p, d, x):
    """float32 torch forward with the kernels' bfloat16 rounding points."""
    import torch
    h = x.to(torch.bfloat16).float()
    layers = O.unflatten(p, d)
    for l, (W, b) in enumerate(layers):
        Wq = torch.tensor(W, dtype=torch.float32, device=x.device).to(torch.bfloat16).float()
        z = h @ Wq.T + torch.tensor(b, dtype=torch.float32, device=x.device)
        h = torch.nn.functional.silu(z).to(torch.bfloat16).float() if l + 1 < len(layers) else z
    return h


def test_forward_matches_oracle_and_torch(gpu_ctx):
    import torch
    print("forward 6-512x4-256, batch 256", flush=True)
    net = _net(gpu_ctx)
    p = net.get_params()
    x = np.random.default_rng(1).uniform(-1, 1, (256, 6)).astype(np.float32)
    y = net.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    yq, _ = O.forward(p, FULL, x, quantize=True)
    y64, _ = O.forward(p, FULL, x)
    yt = _torch_reference(p, FULL, torch.from_numpy(x).cuda()).cpu().numpy()
    print(f"rel-L2 vs bf16 oracle {_rel(y, yq):.2e}, vs torch fp32 {_rel(y, yt):.2e}, "
          f"vs exact fp64 {_rel(y, y64):.2e}", flush=True)
    assert _rel(y, yq) < 5e-3
    assert _rel(y, yt) < 5e-3
    assert _rel(y, y64) < 3e-2


def test_inference_batch_65536(gpu_ctx):
    import torch
    print("forward, batch 65536", flush=True)
    net = _net(gpu_ctx, max_batch=65536, seed=2)
    p = net.get_params()
    g = torch.Generator(device="cpu").manual_seed(3)
    x = (torch.rand((65536, 6), generator=g) * 2 - 1).cuda()
    y = net.forward(x)
    rows = torch.arange(0, 65536, 97, device="cuda")
    yt = _torch_reference(p, FULL, x[rows])
    assert _rel(y[rows].cpu().numpy(), yt.cpu().numpy()) < 5e-3
    # ragged batch: padded to a multiple of 64 inside the wrapper
    y2 = net.forward(x[:1000])
    torch.testing.assert_close(y2, y[:1000], rtol=0, atol=0)


def test_gradient_matches_oracle(gpu_ctx):
    import torch
    print("gradient, batch 256", flush=True)
    net = _net(gpu_ctx, seed=4)
    p = net.get_params()
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, (256, 6)).astype(np.float32)
    t = rng.uniform(0, 1, (256, 256)).astype(np.float32)
    loss, g = net.grad(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda())
    g = g.cpu().numpy()
    lq, gq = O.loss_grad(p, FULL, x, t, quantize=True)
    l64, g64 = O.loss_grad(p, FULL, x, t)
    assert abs(loss - lq) / lq < 1e-3
    G, Q, E = O.split_flat(g, FULL), O.split_flat(gq, FULL), O.split_flat(g64, FULL)
    for k in G:
        r, r64 = _rel(G[k], Q[k]), _rel(G[k], E[k])
        cos = float(np.dot(G[k].ravel(), Q[k].ravel()) /
                    (np.linalg.norm(G[k]) * np.linalg.norm(Q[k])))
        print(f"{k}: rel-L2 vs bf16 oracle {r:.2e} (vs fp64 {r64:.2e}), cos {cos:.6f}", flush=True)
        assert r < 2e-2, k
        assert cos > 0.9995, k
        assert r64 < 6e-2, k
    # deterministic: no atomics in the reductions
    loss2, g2 = net.grad(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda())
    assert loss2 == loss
    np.testing.assert_array_equal(g2.cpu().numpy(), g)


def test_gradient_rows_by_index(gpu_ctx):
    import torch
    net = _net(gpu_ctx, seed=6)
    rng = np.random.default_rng(7)
    x = torch.from_numpy(rng.uniform(-1, 1, (1000, 6)).astype(np.float32)).cuda()
    t = torch.from_numpy(rng.uniform(0, 1, (1000, 256)).astype(np.float32)).cuda()
    idx = torch.from_numpy(rng.permutation(1000)[:128].astype(np.int32)).cuda()
    l1, g1 = net.grad(x, t, idx)
    l2, g2 = net.grad(x[idx.long()].contiguous(), t[idx.long()].contiguous())
    assert l1 == l2
    torch.testing.assert_close(g1, g2, rtol=0, atol=0)


def test_train_step_is_adamw_on_the_gradient(gpu_ctx):
    import torch
    print("AdamW step", flush=True)
    net = _net(gpu_ctx, seed=8)
    p0 = net.get_params()
    rng = np.random.default_rng(9)
    x = torch.from_numpy(rng.uniform(-1, 1, (512, 6)).astype(np.float32)).cuda()
    t = torch.from_numpy(rng.uniform(0, 1, (512, 256)).astype(np.float32)).cuda()
    idx = torch.arange(256, dtype=torch.int32, device="cuda")
    _, g = net.grad(x, t, idx)
    g = g.cpu().numpy().astype(np.float64)
    hp = dict(lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.01)
    net.train_step(x, t, idx, AdamW(hp["lr"], hp["beta1"], hp["beta2"], hp["eps"],
                                    hp["weight_decay"], 1))
    p1 = net.get_params()
    ref, m, v = O.adamw(p0, g, np.zeros_like(g), np.zeros_like(g), step=1, **hp)
    np.testing.assert_allclose(p1, ref, rtol=1e-5, atol=1e-7)
    # second step from the moments the kernel kept
    _, g2 = net.grad(x, t, idx)
    g2 = g2.cpu().numpy().astype(np.float64)
    net.train_step(x, t, idx, AdamW(hp["lr"], hp["beta1"], hp["beta2"], hp["eps"],
                                    hp["weight_decay"], 2))
    ref2, _, _ = O.adamw(p1, g2, m, v, step=2, **hp)
    np.testing.assert_allclose(net.get_params(), ref2, rtol=1e-5, atol=1e-7)


def test_training_reduces_the_loss(gpu_ctx):
    import torch
    print("training on a smooth synthetic map, 400 steps", flush=True)
    d = FULL
    net = _net(gpu_ctx, seed=10)
    rng = np.random.default_rng(11)
    N = 8192
    x = rng.uniform(-1, 1, (N, 6)).astype(np.float32)
    grid = np.linspace(0, 1, 256)
    centre = 0.5 + 0.3 * x[:, :1]
    width = 0.08 + 0.04 * (x[:, 1:2] + 1)
    t = (0.2 + 0.6 * (x[:, 2:3] + 1) / 2) * np.exp(-0.5 * ((grid - centre) / width) ** 2)
    xt = torch.from_numpy(x).cuda()
    tt = torch.from_numpy(t.astype(np.float32)).cuda()
    cfg = S.TrainConfig(epochs=100, batch=2048, lr=2e-3, warmup_steps=20, weight_decay=0.0)
    hist = S.train(net, xt, tt, cfg, log=None)
    loss = hist["loss"]
    first, last = loss[:4].mean(), loss[-8:].mean()
    print(f"loss {first:.3e} -> {last:.3e} over {len(loss)} steps", flush=True)
    assert np.all(np.isfinite(loss))
    assert last < first / 20


def test_invalid_batches_are_rejected(gpu_ctx):
    import torch
    net = _net(gpu_ctx, max_batch=128)
    x = torch.zeros((100, 6), device="cuda")
    t = torch.zeros((100, 256), device="cuda")
    with pytest.raises(MclError):
        net.grad(x, t)                    # not a multiple of 64
    with pytest.raises(ValueError):
        net.forward(torch.zeros((200, 6), device="cuda"))   # beyond max_batch


def test_close_right_after_unsynchronised_training(gpu_ctx):
    # the destroy waits for the network's own stream, where its steps were queued:
    # freeing the weights under running kernels would fault or corrupt later work
    import torch
    print("close after queued AdamW steps, then a fresh network's forward", flush=True)
    rng = np.random.default_rng(12)
    x = torch.from_numpy(rng.uniform(-1, 1, (4096, 6)).astype(np.float32)).cuda()
    t = torch.from_numpy(rng.uniform(0, 1, (4096, 256)).astype(np.float32)).cuda()
    idx = torch.arange(4096, dtype=torch.int32, device="cuda")
    for k in range(3):
        net = _net(gpu_ctx, max_batch=4096, seed=20 + k)
        for step in range(1, 6):
            net.train_step(x, t, idx, AdamW(1e-3, 0.9, 0.999, 1e-8, 0.0, step))
        net.close()                       # no synchronisation before it
    fresh = _net(gpu_ctx, max_batch=4096, seed=30)
    ref = _net(gpu_ctx, max_batch=4096, seed=30)
    np.testing.assert_array_equal(fresh.forward(x).cpu().numpy(), ref.forward(x).cpu().numpy())
    fresh.close()
    ref.close()


def test_fused_inference_matches_the_layered_form(gpu_ctx):
    # one kernel for the whole forward (activations in LDS) against the per-layer
    # GEMM kernels: the same products in the same order, bit for bit; a batch that
    # ends in half a workgroup of rows, and after a training step (weights repacked)
    import torch
    print("fused vs layered inference, batch 65536 and 4160", flush=True)
    net = _net(gpu_ctx, max_batch=65536, seed=4)
    g = torch.Generator(device="cpu").manual_seed(5)
    x = (torch.rand((65536, 6), generator=g) * 2 - 1).cuda()
    t = torch.rand((4096, 256), generator=g).cuda()
    try:
        for step in range(2):
            outs = {}
            for fused in (1.0, 0.0):
                gpu_ctx.tune_set("mlp_fused", fused)
                outs[fused] = (net.forward(x).cpu(), net.forward(x[:4160]).cpu())
            for a, b in zip(outs[1.0], outs[0.0]):
                torch.testing.assert_close(a, b, rtol=0, atol=0)
            idx = torch.arange(4096, dtype=torch.int32, device="cuda")
            net.train_step(x, t, idx, AdamW(1e-3, 0.9, 0.999, 1e-8, 0.0, step + 1))
    finally:
        gpu_ctx.tune_set("mlp_fused", 1.0)
    net.close()
