"""The network-scored differential evolution in device memory (mcl_de_*): its
objective against the same objective computed by torch from the network's
forward, greedy selection, the argmin, determinism, convergence on a target the
network can reproduce exactly, the fused one-kernel generation against the
layered one (bit-identical populations), and argument checks."""
import math

import numpy as np
import pytest

from mcluminescence_amd import fit as F
from mcluminescence_amd import fit_bench
from mcluminescence_amd import surrogate as S
from mcluminescence_amd._abi import MclError

pytestmark = pytest.mark.gpu

U_BETA = 0.1          # the fixed sixth input (encoded heating rate)
SCALE = 1.0e4


@pytest.fixture(scope="module")
def net(gpu_ctx):
    n = S.Network.load(gpu_ctx, fit_bench.MODEL, max_batch=4096)
    yield n
    n.close()


def _curves(net, u5):
    """The network's de-normalised curves for coordinates u5 [n, 5] (torch)."""
    import torch
    x = torch.cat([u5, torch.full((u5.shape[0], 1), U_BETA, device="cuda")], dim=1)
    y = net.forward(x)
    return torch.expm1(y * math.log1p(1.0 / net.c0)) * (net.c0 * SCALE)


def _objective(net, u5, target):
    d = _curves(net, u5) - target[None, :]
    return (d * d).mean(dim=1) / max(float(target.max()), 1.0) ** 2


def _target(net, u_star):
    import torch
    return _curves(net, torch.tensor([u_star], dtype=torch.float32, device="cuda"))[0]


def test_objective_matches_torch_and_selection_is_greedy(net):
    print("DE: objective vs torch, greedy selection, argmin (pop 300)", flush=True)
    import torch
    tgt = _target(net, [0.2, -0.3, 0.1, 0.4, -0.5])
    de = F.NetworkDE(net, tgt.cpu().numpy(), SCALE, [U_BETA], 5, 300, seed=11)
    try:
        x0, f0 = de.population()
        assert torch.all(x0.abs() <= 1.0)
        ref = _objective(net, x0, tgt)
        torch.testing.assert_close(f0, ref, rtol=2e-4, atol=1e-9)
        for _ in range(3):
            de.run(1)
            x1, f1 = de.population()
            assert torch.all(f1 <= f0)
            assert torch.all(x1.abs() <= 1.0)
            moved = (x1 != x0).any(dim=1)
            assert torch.equal(moved | (f1 < f0), moved)         # a member changes only by a trial
            torch.testing.assert_close(f1, _objective(net, x1, tgt), rtol=2e-4, atol=1e-9)
            xb, fb = de.best()
            i = int(torch.argmin(f1).item())
            assert fb == float(f1[i].item())
            np.testing.assert_array_equal(xb.astype(np.float32), x1[i].cpu().numpy())
            x0, f0 = x1, f1
    finally:
        de.close()


def test_search_is_deterministic_and_converges(net):
    print("DE: 150 generations twice on a network-consistent target (pop 512)", flush=True)
    u_star = [0.35, -0.6, 0.05, -0.2, 0.7]
    tgt = _target(net, u_star).cpu().numpy()
    res = []
    for _ in range(2):
        de = F.NetworkDE(net, tgt, SCALE, [U_BETA], 5, 512, seed=3)
        try:
            _, f_start = de.best()
            de.run(150)
            res.append((*de.best(), f_start))
        finally:
            de.close()
    (xa, fa, f_start), (xb, fb, _) = res
    np.testing.assert_array_equal(xa, xb)
    assert fa == fb
    assert fa < 1e-3 * f_start and fa < 1e-4, (fa, f_start)


@pytest.mark.parametrize("pop", [300, 1024, 2000])
def test_fused_generation_matches_layered(net, pop):
    # the surrogate's own shapes: the fused kernel's unrolled form
    print(f"DE: fused vs layered generations, 40 generations (pop {pop})", flush=True)
    _fused_vs_layered(net, pop)


def test_fused_generation_matches_layered_other_shapes(gpu_ctx):
    # 6 -> 128 x 2 -> 128: the fused kernel's form for any shape it takes
    print("DE: fused vs layered generations, 6 -> 128 x 2 -> 128, pop 300", flush=True)
    small = S.Network(gpu_ctx, n_hidden=128, n_hidden_layers=2, n_out=128, max_batch=4096, seed=3)
    try:
        _fused_vs_layered(small, 300)
    finally:
        small.close()


def _fused_vs_layered(net, pop):
    import torch
    tgt = _target(net, [-0.1, 0.25, 0.6, -0.45, 0.3]).cpu().numpy()
    res = []
    for layered in (False, True):
        de = F.NetworkDE(net, tgt, SCALE, [U_BETA], 5, pop, seed=21, layered=layered)
        try:
            de.run(15)
            de.run(25)
            x, f = de.population()
            res.append((x.clone(), f.clone(), de.best()))
        finally:
            de.close()
    (xa, fa, ba), (xb, fb, bb) = res
    assert torch.equal(xa, xb) and torch.equal(fa, fb)
    np.testing.assert_array_equal(ba[0], bb[0])
    assert ba[1] == bb[1]


def test_invalid_parameters_are_rejected(net):
    tgt = np.ones(256, np.float32)
    for kw in ({"pop": 3}, {"dim": 0}, {"dim": 9}, {"F": -1.0}, {"CR": 1.5}, {"pop": 5000}):
        args = dict(dim=5, pop=64, F=0.5, CR=0.3)
        args.update(kw)
        fixed = [U_BETA] if args["dim"] == 5 else [U_BETA] * max(0, 6 - args["dim"])
        with pytest.raises(MclError):
            F.NetworkDE(net, tgt, SCALE, fixed[:8], args["dim"], args["pop"], seed=1,
                        F=args["F"], CR=args["CR"])


def test_unknown_flags_are_rejected(net):
    import ctypes as C
    import torch
    from mcluminescence_amd._abi import DeParams
    p = DeParams()
    p.pop, p.dim, p.F, p.CR, p.seed = 64, 5, 0.5, 0.3, 1
    p.fixed[0] = U_BETA
    p.log_span, p.out_scale, p.inv_peak2 = 1.0, 1.0, 1.0
    p.flags = 4
    tgt = torch.ones(256, dtype=torch.float32, device="cuda")
    h = C.c_void_p()
    assert net.lib.mcl_de_create(net.handle, C.byref(p), tgt.data_ptr(), C.byref(h)) != 0
    assert not h.value
