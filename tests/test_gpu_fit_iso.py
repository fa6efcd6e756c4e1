"""The isothermal-decay surrogate (mcluminescence_amd/surrogate_iso.npz) on the GPU:
its curves against low-noise Monte Carlo, and the surrogate-assisted fit of a
synthetic isothermal lab curve against the same fit done with plain Monte Carlo
(the ISO half of the fit workload; the reference's exp == "iso" objective,
optimizer.py:73-74, on ISO_lab, tl_trap_lab.py:125-179)."""
import numpy as np
import pytest

from mcluminescence_amd import fit as F
from mcluminescence_amd import fit_bench
from mcluminescence_amd import surrogate as S
from mcluminescence_amd import train_surrogate as T

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def net(gpu_ctx):
    n = S.Network.load(gpu_ctx, fit_bench.MODELS["iso"], max_batch=4096)
    assert n.kind is S.ISO
    yield n
    n.close()


def test_iso_surrogate_curves_within_two_percent_nrmse(gpu_ctx, net):
    print("isothermal surrogate vs 1024-cluster Monte Carlo, 64 held-out resolved decays",
          flush=True)
    theta = T.sample(S.ISO, 64, np.random.default_rng(2024), resolved_only=True)
    mc = S.ISO.mc_curves(gpu_ctx, theta, n=400, sims=1024, seed=S.DEFAULT_SEED + 4243)
    pred = S.predict_curves(net, theta)
    e = S.nrmse(pred, mc)
    print(f"NRMSE median {np.median(e):.4f}, p90 {np.percentile(e, 90):.4f}, max {e.max():.4f}",
          flush=True)
    assert np.median(e) < 0.02
    assert np.percentile(e, 90) < 0.03


def test_iso_surrogate_fit_gains_over_plain_mc(gpu_ctx, net):
    print("one synthetic isothermal lab curve: surrogate-assisted fit vs plain Monte Carlo fit",
          flush=True)
    rng = np.random.default_rng(7)
    truth = fit_bench.draw_truth_iso(rng)
    y, scale = F.lab_curve(gpu_ctx, truth, rng=rng, kind=S.ISO)
    # one untimed pass of each method loads every kernel both use
    F.fit_surrogate(gpu_ctx, net, y, scale, truth[5], pop=1024, gens=2, seed=9)
    F.fit_mc(gpu_ctx, y, scale, truth[5], pop=8, gens=1, seed=9, kind=S.ISO)
    rs = F.fit_surrogate(gpu_ctx, net, y, scale, truth[5], pop=1024, gens=150, seed=1)
    rm = F.fit_mc(gpu_ctx, y, scale, truth[5], pop=128, gens=30, seed=1, kind=S.ISO)
    jo = lambda th: F.judge_objective(gpu_ctx, th, y, scale, kind=S.ISO)
    js, jm, jt = jo(rs.theta), jo(rm.theta), jo(truth)
    print(f"surrogate {rs.seconds:.3f} s (judge NRMSE {rs.judge_nrmse:.4f}, objective {js:.3e}); "
          f"plain MC {rm.seconds:.2f} s (objective {jm:.3e}); truth objective {jt:.3e}", flush=True)
    print("recovery surrogate", rs.recovery(truth), "mc", rm.recovery(truth), flush=True)
    assert rs.kind == "iso" and not rs.fallback and rs.judge_nrmse <= 0.02
    assert rs.mc_candidates == 1 and rm.mc_candidates == 128 * 31
    assert js <= 1.5 * max(jm, jt)
    assert rs.recovery(truth)["log10_k"] < 0.05
    assert rm.seconds > 5 * rs.seconds


def test_iso_fallback_runs_monte_carlo_generations(gpu_ctx, net):
    rng = np.random.default_rng(8)
    truth = fit_bench.draw_truth_iso(rng)
    y, scale = F.lab_curve(gpu_ctx, truth, rng=rng, kind=S.ISO)
    r = F.fit_surrogate(gpu_ctx, net, y, scale, truth[5], pop=256, gens=20, gate=0.0, seed=2,
                        fallback_pop=32, fallback_gens=3)
    assert r.fallback and r.kind == "iso"
    assert r.mc_candidates == 1 + 32 * 4
    assert np.isfinite(r.objective_mc)
