"""Ensemble parity of the GPU fixed-dt Monte Carlo with the reference's event-mode
algorithm: per-bin mean recombinations per cluster against the event-mode
oracle's ensembles (tests/golden/ensemble_event_mode.json, made by
tests/golden/make_ensemble_fixtures.py from oracle/event_mc.py, the numpy
restatement of simulate.py:36-92), within 3 standard errors per bin; glow peak
temperatures within 1 K of the oracle's and of the Randall-Wilkins closed form.

Statistics: z = (mean_gpu - mean_oracle) / sqrt(v_o / S_o + v_g / S_g) per bin,
with v = max(sample variance, min(pooled mean, 1)) on each side: in tail bins
where a side saw no event at all its sample variance is 0, and the Poisson
variance of the pooled mean (for bins of under one event per cluster) is the
floor that keeps the normal approximation honest.  With
~40-80 bins per case a chance |z| > 3 has probability ~0.3 % per bin, so each
case allows at most one such bin and none beyond 4.5.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from mcluminescence_amd._abi import RunParams
from oracle import event_mc as E

pytestmark = pytest.mark.gpu

FX = json.loads((Path(__file__).parent / "golden" / "ensemble_event_mode.json").read_text())
CASES = [k for k in FX if not k.startswith("_")]
GPU_SIMS = 4096


def _gpu_bins(ctx, name):
    f = FX[name]
    run = RunParams.make(**{**f["run"], "sims": GPU_SIMS})
    counts, _, _ = ctx.run_fixed_dt([run], n_bins=0)        # per-step counts
    per = f["steps_per_bin"]
    nb = f["n_bins"]
    coarse = counts[:, : nb * per].reshape(GPU_SIMS, nb, per).sum(axis=2).astype(np.float64)
    return run, coarse


def _peak(T, y):
    """Vertex of a parabola fitted to the 7 bins around the maximum."""
    i = int(np.argmax(y))
    lo, hi = max(i - 3, 0), min(i + 4, len(y))
    c = np.polyfit(T[lo:hi], y[lo:hi], 2)
    return -c[1] / (2 * c[0])


@pytest.mark.parametrize("name", CASES)
def test_mean_curve_within_three_sigma_per_bin(gpu_ctx, name):
    f = FX[name]
    print(f"ensemble {name}: {GPU_SIMS} GPU clusters vs {f['sims']} event-mode clusters",
          flush=True)
    run, coarse = _gpu_bins(gpu_ctx, name)
    mg, sg = coarse.mean(axis=0), coarse.std(axis=0, ddof=1)
    mo, so = np.array(f["mean"]), np.array(f["std"])
    pooled = (mo * f["sims"] + mg * GPU_SIMS) / (f["sims"] + GPU_SIMS)
    floor = np.minimum(pooled, 1.0)
    se = np.sqrt(np.maximum(so ** 2, floor) / f["sims"] + np.maximum(sg ** 2, floor) / GPU_SIMS)
    live = se > 0
    assert np.all(mg[~live] == mo[~live])          # bins empty on both sides
    z = (mg[live] - mo[live]) / se[live]
    print(f"  bins {live.sum()}, max |z| {np.abs(z).max():.2f}, |z|>3: {(np.abs(z) > 3).sum()}, "
          f"chi2/dof {np.mean(z ** 2):.2f}", flush=True)
    assert (np.abs(z) > 3).sum() <= 1
    assert np.abs(z).max() < 4.5
    # totals per cluster agree too
    tot_se = np.sqrt(np.sum(so ** 2) / f["sims"])
    assert abs(coarse.sum(axis=1).mean() - f["total_mean"]) < 4 * max(tot_se, 0.5)


def test_glow_peak_within_1K(gpu_ctx):
    name = "tl_glow_500"
    f = FX[name]
    run, coarse = _gpu_bins(gpu_ctx, name)
    per = f["steps_per_bin"]
    T = (run.T_start + 273.15) + run.T_rate * run.dt * per * (np.arange(f["n_bins"]) + 0.5)
    Tg = _peak(T, coarse.mean(axis=0))
    To = _peak(T, np.array(f["mean"]))
    # the fine-grained GPU curve against the closed form (band term, first order)
    counts, _, _ = gpu_ctx.run_fixed_dt([RunParams.make(**{**f["run"], "sims": GPU_SIMS})])
    Tf = (run.T_start + 273.15) + run.T_rate * run.dt * (np.arange(counts.shape[1]) + 0.5)
    sm = np.convolve(counts.sum(axis=0).astype(float), np.ones(21) / 21, "same")
    i = int(np.argmax(sm))
    c = np.polyfit(Tf[i - 60:i + 61], sm[i - 60:i + 61], 2)
    Tfine = -c[1] / (2 * c[0])
    T_rw = E.randall_wilkins_peak(run.E_cb, run.s, run.T_rate)
    print(f"glow peak: GPU {Tg:.2f} K (fine {Tfine:.2f} K), event-mode oracle {To:.2f} K, "
          f"Randall-Wilkins {T_rw:.2f} K", flush=True)
    assert abs(Tg - To) < 1.0
    assert abs(Tfine - T_rw) < 1.0
