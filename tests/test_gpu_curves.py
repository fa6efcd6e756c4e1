"""Per-set curves from per-step counts (mcl_bin_counts) against numpy's segment
sums, and the isothermal surrogate's Monte Carlo curves (the trapped fraction at
the end of each log-time bin) against the band-only closed form."""
import math

import numpy as np
import pytest

from mcluminescence_amd import surrogate as S
from mcluminescence_amd._abi import MCL_EINVAL, MCL_ERANGE, MclError

pytestmark = pytest.mark.gpu


def _ref_bins(counts, sims, edges):
    """numpy: sum over each set's sims, then over each bin's steps (np.add.reduceat
    on the steps up to the last edge)."""
    tot = counts.astype(np.int64).reshape(-1, sims, counts.shape[1]).sum(axis=1)
    return np.add.reduceat(tot[:, :edges[-1]], edges[:-1], axis=1)


@pytest.mark.parametrize("row,sims,nb,aligned", [(4000, 8, 256, True), (1001, 3, 17, False),
                                                 (64, 1, 64, True), (5000, 5, 512, True)])
def test_bin_counts_match_numpy(gpu_ctx, row, sims, nb, aligned):
    import torch
    print(f"bin_counts: rows {row}, sims {sims}, bins {nb}", flush=True)
    rng = np.random.default_rng(row + nb)
    n_sets = 37
    counts = rng.integers(0, 1000, (n_sets * sims, row), dtype=np.int64).astype(np.uint32)
    last = row if aligned else row - 3
    inner = np.sort(rng.choice(np.arange(1, last), nb - 1, replace=False))
    edges = np.concatenate([[int(rng.integers(0, inner[0]))], inner, [last]]).astype(np.int32)
    dev = torch.from_numpy(counts.view(np.int32)).cuda()
    out = torch.full((n_sets, nb), -1, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()              # the library runs on its own stream
    gpu_ctx.bin_counts(dev.data_ptr(), n_sets, sims, row, edges, out.data_ptr())
    gpu_ctx.sync()
    np.testing.assert_array_equal(out.cpu().numpy(), _ref_bins(counts, sims, edges))


def test_bin_counts_rejects_bad_edges(gpu_ctx):
    import torch
    dev = torch.zeros((2, 100), dtype=torch.int32, device="cuda")
    out = torch.zeros((2, 4), dtype=torch.int64, device="cuda")
    with pytest.raises(MclError) as e:
        gpu_ctx.bin_counts(dev.data_ptr(), 2, 1, 100, [0, 5, 5, 9, 10], out.data_ptr())
    assert e.value.code == MCL_EINVAL
    with pytest.raises(MclError) as e:
        gpu_ctx.bin_counts(dev.data_ptr(), 2, 1, 100, [0, 5, 7, 9, 101], out.data_ptr())
    assert e.value.code == MCL_ERANGE


def test_iso_curves_follow_the_band_only_closed_form(gpu_ctx):
    # band-only decay at constant T: each electron survives n steps with (1 - p)^n,
    # p = 1 - e^{-k dt}, so the trapped fraction after step e - 1 is f0 (1 - p)^e
    print("isothermal surrogate curves: 3 parameter sets x 512 clusters", flush=True)
    theta = np.array([[1.4, 1e13, 4e9, 1e23, 1.0, 500.0],
                      [1.2, 1e11, 3e9, 3e22, 0.6, 480.0],
                      [1.6, 1e14, 5e9, 1e24, 0.35, 560.0]])
    n, sims = 400, 512
    curves = S.ISO.mc_curves(gpu_ctx, theta, n=n, sims=sims)
    edges = np.asarray(S.ISO.edges)
    for th, c in zip(theta, curves):
        k = th[1] * math.exp(-th[0] / (8.617343e-5 * th[5]))
        q = math.exp(-k * S.ISO_DT)                    # 1 - p
        frac0 = int(n * th[4]) / n
        surv = q ** edges[1:]
        expect = frac0 * surv
        # binomial over the sims x n0 independent electrons, floored at one electron
        var = frac0 * surv * (1 - surv) / (n * sims)
        z = (c - expect) / np.sqrt(np.maximum(var, 1.0 / (n * sims) ** 2))
        assert np.abs(z).max() < 5, (th, np.abs(z).max())
        assert np.all(np.diff(c) <= 0) and c[0] <= frac0
