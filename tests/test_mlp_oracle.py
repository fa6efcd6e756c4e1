"""Pins of the surrogate oracle (oracle/mlp_oracle.py) on the CPU: its gradient
against central finite differences and against torch.autograd in float64, its
AdamW step against torch.optim.AdamW, and its bfloat16 rounding against torch's
float32 -> bfloat16 conversion."""
import numpy as np
import pytest
import torch

from oracle import mlp_oracle as O

TINY = O.Desc(n_in=3, n_hidden=5, n_hidden_layers=2, n_out=4)


def _data(d, B, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1, 1, (B, d.n_in)), rng.uniform(-1, 1, (B, d.n_out))


def test_param_count_of_the_surrogate_shape():
    # 6 -> 512 x4 (SiLU) -> 256
    assert O.Desc().param_count() == 922_880


def test_gradient_matches_central_differences():
    p = O.init_params(TINY, seed=1).astype(np.float64)
    x, t = _data(TINY, 7, 2)
    _, g = O.loss_grad(p, TINY, x, t)
    h = 1e-6
    num = np.empty_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        num[i] = (O.loss_grad(p + e, TINY, x, t)[0] - O.loss_grad(p - e, TINY, x, t)[0]) / (2 * h)
    np.testing.assert_allclose(g, num, rtol=1e-6, atol=1e-9)


def _torch_net(p, d):
    layers = O.unflatten(p, d)
    params = []
    for W, b in layers:
        params += [torch.tensor(W, dtype=torch.float64, requires_grad=True),
                   torch.tensor(b, dtype=torch.float64, requires_grad=True)]
    return params


def _torch_forward(params, x):
    h = x
    n = len(params) // 2
    for l in range(n):
        h = h @ params[2 * l].T + params[2 * l + 1]
        if l + 1 < n:
            h = torch.nn.functional.silu(h)
    return h


def test_gradient_matches_torch_autograd_float64():
    d = O.Desc(n_in=6, n_hidden=32, n_hidden_layers=3, n_out=16)
    p = O.init_params(d, seed=3).astype(np.float64)
    x, t = _data(d, 64, 4)
    loss, g = O.loss_grad(p, d, x, t)
    params = _torch_net(p, d)
    y = _torch_forward(params, torch.tensor(x))
    L = torch.mean((y - torch.tensor(t)) ** 2)
    L.backward()
    gt = np.concatenate([q.grad.numpy().ravel() for q in params])
    assert abs(loss - L.item()) < 1e-12
    np.testing.assert_allclose(g, gt, rtol=1e-10, atol=1e-14)


def test_adamw_matches_torch_optim():
    rng = np.random.default_rng(5)
    p0 = rng.normal(size=50)
    grads = [rng.normal(size=50) for _ in range(4)]
    hp = dict(lr=3e-3, beta1=0.9, beta2=0.99, eps=1e-8, weight_decay=0.05)
    w = torch.tensor(p0, requires_grad=True)
    opt = torch.optim.AdamW([w], lr=hp["lr"], betas=(hp["beta1"], hp["beta2"]), eps=hp["eps"],
                            weight_decay=hp["weight_decay"])
    p, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    for k, g in enumerate(grads, start=1):
        w.grad = torch.tensor(g)
        opt.step()
        p, m, v = O.adamw(p, g, m, v, step=k, **hp)
    np.testing.assert_allclose(p, w.detach().numpy(), rtol=1e-13, atol=1e-15)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(6)
    a = np.concatenate([rng.normal(size=10000) * 10 ** rng.uniform(-30, 30, 10000),
                        [0.0, -0.0, 1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 65504.0]]).astype(np.float32)
    ref = torch.tensor(a).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(O.bf16(a), ref)


def test_quantized_forward_is_close_to_exact():
    # the bfloat16 restatement departs from float64 by bfloat16 rounding only
    d = O.Desc(n_in=6, n_hidden=128, n_hidden_layers=4, n_out=128)
    p = O.init_params(d, seed=7)
    x, _ = _data(d, 32, 8)
    y64, _ = O.forward(p, d, x)
    yq, _ = O.forward(p, d, x, quantize=True)
    rel = np.linalg.norm(yq - y64) / np.linalg.norm(y64)
    assert 1e-5 < rel < 3e-2


@pytest.mark.parametrize("seed", [0, 1])
def test_loss_is_mean_square(seed):
    p = O.init_params(TINY, seed=seed)
    x, t = _data(TINY, 5, seed)
    y, _ = O.forward(p, TINY, x)
    loss, _ = O.loss_grad(p, TINY, x, t)
    assert abs(loss - np.mean((y - t) ** 2)) < 1e-15
