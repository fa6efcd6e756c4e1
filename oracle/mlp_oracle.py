"""NumPy float64 restatement of the surrogate network: forward pass, mean-squared
loss, backward pass and the AdamW update.

TEST INFRASTRUCTURE ONLY — the checker for the HIP surrogate kernels, never the
product.  The reference has no surrogate (its fit, src/class/optimizer.py:66-111,
simulates every candidate), so nothing here follows a reference file; each
function states the mathematics it restates, and tests/test_mlp_oracle.py pins
it against central finite differences, torch.autograd (float64, CPU) and
torch.optim.AdamW.

Layout (the flat parameter vector of include/mcl_surrogate.h): for every linear
layer l, W_l [out][in] row-major then b_l [out]; y = W x + b; SiLU after every
layer but the last.

``quantize=True`` rounds to bfloat16 (round to nearest even) exactly where the
kernels hold bfloat16 data: the input rows, the weights, each layer's
pre-activation z and activation silu(z), the scaled output error and every
dL/dz.  Everything else is float64.  With ``quantize=False`` the restatement is
the exact float64 network, against which the bfloat16 error is stated.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, List, Tuple

import numpy as np


@dataclass(frozen=True)
class Desc:
    n_in: int = 6
    n_hidden: int = 512
    n_hidden_layers: int = 4
    n_out: int = 256

    def shapes(self) -> List[Tuple[int, int]]:
        """(out, in) of every linear layer."""
        dims = [self.n_in] + [self.n_hidden] * self.n_hidden_layers + [self.n_out]
        return [(dims[i + 1], dims[i]) for i in range(len(dims) - 1)]

    def param_count(self) -> int:
        return sum(o * i + o for o, i in self.shapes())


def bf16(a: np.ndarray) -> np.ndarray:
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unflatten(p: np.ndarray, d: Desc) -> List[Tuple[np.ndarray, np.ndarray]]:
    out, o = [], 0
    for n_out, n_in in d.shapes():
        W = p[o:o + n_out * n_in].reshape(n_out, n_in)
        o += n_out * n_in
        b = p[o:o + n_out]
        o += n_out
        out.append((W, b))
    assert o == p.size
    return out


def flatten(layers: List[Tuple[np.ndarray, np.ndarray]]) -> np.ndarray:
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers])


def init_params(d: Desc, seed: int = 0) -> np.ndarray:
    """Uniform(-1/sqrt(in), 1/sqrt(in)) for weights and biases (the default
    initialisation of a linear layer)."""
    rng = np.random.default_rng(seed)
    parts = []
    for n_out, n_in in d.shapes():
        lim = 1.0 / np.sqrt(n_in)
        parts.append(rng.uniform(-lim, lim, n_out * n_in))
        parts.append(rng.uniform(-lim, lim, n_out))
    return np.concatenate(parts).astype(np.float32)


def _sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(p: np.ndarray, d: Desc, x: np.ndarray, quantize: bool = False):
    """y = net(x) and the per-layer cache (layer input, pre-activation z)."""
    q = bf16 if quantize else (lambda a: np.asarray(a, np.float64))
    layers = unflatten(np.asarray(p, np.float64), d)
    h = q(x)
    cache = []
    for l, (W, b) in enumerate(layers):
        z = h @ q(W).T + b
        if l + 1 < len(layers):
            zq = q(z)
            cache.append((h, zq))
            h = q(z * _sigmoid(z))
        else:
            cache.append((h, None))
            h = z
    return h, cache


def loss_grad(p: np.ndarray, d: Desc, x: np.ndarray, t: np.ndarray, quantize: bool = False):
    """L = mean((net(x) - t)^2) over rows and outputs, and dL/dp (flat)."""
    q = bf16 if quantize else (lambda a: np.asarray(a, np.float64))
    layers = unflatten(np.asarray(p, np.float64), d)
    y, cache = forward(p, d, x, quantize)
    B, n_out = y.shape
    diff = y - np.asarray(t, np.float64)
    loss = float(np.mean(diff * diff))
    dz = q(diff * (2.0 / (B * n_out)))
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        h, _ = cache[l]
        grads[l] = (dz.T @ h, dz.sum(axis=0))
        if l == 0:
            break
        da = dz @ q(W)
        z = cache[l - 1][1]
        s = _sigmoid(z)
        dz = q(da * (s * (1.0 + z * (1.0 - s))))
    return loss, flatten(grads)


def adamw(p: np.ndarray, g: np.ndarray, m: np.ndarray, v: np.ndarray, *, lr: float,
          beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8,
          weight_decay: float = 0.0, step: int = 1):
    """One AdamW step (Loshchilov & Hutter, decoupled decay):
    p <- p (1 - lr wd); m <- b1 m + (1-b1) g; v <- b2 v + (1-b2) g^2;
    p <- p - lr / (1 - b1^t) * m / (sqrt(v) / sqrt(1 - b2^t) + eps)."""
    p = np.asarray(p, np.float64) * (1.0 - lr * weight_decay)
    m = beta1 * np.asarray(m, np.float64) + (1.0 - beta1) * g
    v = beta2 * np.asarray(v, np.float64) + (1.0 - beta2) * g * g
    bc1 = 1.0 - beta1 ** step
    bc2 = 1.0 - beta2 ** step
    p = p - (lr / bc1) * m / (np.sqrt(v) / np.sqrt(bc2) + eps)
    return p, m, v


def split_flat(p: np.ndarray, d: Desc) -> Dict[str, np.ndarray]:
    """Per-layer views 'W0', 'b0', ... of a flat vector (for per-layer checks)."""
    out = {}
    for l, (W, b) in enumerate(unflatten(np.asarray(p), d)):
        out[f"W{l}"] = W
        out[f"b{l}"] = b
    return out
