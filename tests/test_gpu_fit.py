"""The trained surrogate (mcluminescence_amd/surrogate_tl.npz) on the GPU: its
curves against low-noise Monte Carlo, and what it gains inside the fit loop
against the same fit done with plain Monte Carlo."""
import numpy as np
import pytest

from mcluminescence_amd import fit as F
from mcluminescence_amd import fit_bench
from mcluminescence_amd import surrogate as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def net(gpu_ctx):
    n = S.Network.load(gpu_ctx, fit_bench.MODEL, max_batch=4096)
    yield n
    n.close()


def test_surrogate_curves_within_two_percent_nrmse(gpu_ctx, net):
    print("surrogate vs 256-cluster Monte Carlo, 64 held-out parameter sets", flush=True)
    rng = np.random.default_rng(2024)
    theta = S.sample_params(64, rng)
    mc = S.mc_curves(gpu_ctx, theta, n=400, sims=256, seed=S.DEFAULT_SEED + 4242)
    pred = S.predict_curves(net, theta)
    e = S.nrmse(pred, mc)
    print(f"NRMSE median {np.median(e):.4f}, p90 {np.percentile(e, 90):.4f}, max {e.max():.4f}",
          flush=True)
    assert np.median(e) < 0.02
    assert np.percentile(e, 90) < 0.025


def test_surrogate_fit_gains_over_plain_mc(gpu_ctx, net):
    print("one synthetic lab curve: surrogate-assisted fit vs plain Monte Carlo fit", flush=True)
    rng = np.random.default_rng(7)
    truth = fit_bench.draw_truth(rng)
    y, scale = F.lab_curve(gpu_ctx, truth, rng=rng)
    # one untimed pass of each method loads every kernel both use: the timed fits
    # below are steady-state fits, as in a service fitting many curves
    F.fit_surrogate(gpu_ctx, net, y, scale, truth[5], pop=1024, gens=2, seed=9)
    F.fit_mc(gpu_ctx, y, scale, truth[5], pop=8, gens=1, seed=9)
    rs = F.fit_surrogate(gpu_ctx, net, y, scale, truth[5], pop=1024, gens=150, seed=1)
    rm = F.fit_mc(gpu_ctx, y, scale, truth[5], pop=128, gens=30, seed=1)
    js = F.judge_objective(gpu_ctx, rs.theta, y, scale)
    jm = F.judge_objective(gpu_ctx, rm.theta, y, scale)
    jt = F.judge_objective(gpu_ctx, truth, y, scale)
    print(f"surrogate {rs.seconds:.2f} s (judge NRMSE {rs.judge_nrmse:.4f}, objective {js:.3e}); "
          f"plain MC {rm.seconds:.2f} s ({rm.mc_candidates} simulated candidates, objective "
          f"{jm:.3e}); truth objective {jt:.3e}", flush=True)
    print("recovery surrogate", rs.recovery(truth), "mc", rm.recovery(truth), flush=True)
    assert not rs.fallback and rs.judge_nrmse <= 0.02
    assert rs.mc_candidates == 1 and rm.mc_candidates == 128 * 31
    # as good a fit: within 50 % of the plain fit's objective and of the truth's
    assert js <= 1.5 * max(jm, jt)
    assert rs.recovery(truth)["E"] < 0.05
    # in a fraction of the time
    assert rm.seconds > 5 * rs.seconds


def test_fallback_runs_monte_carlo_generations(gpu_ctx, net):
    rng = np.random.default_rng(8)
    truth = fit_bench.draw_truth(rng)
    y, scale = F.lab_curve(gpu_ctx, truth, rng=rng)
    # a gate no network meets forces the Monte Carlo fallback
    r = F.fit_surrogate(gpu_ctx, net, y, scale, truth[5], pop=256, gens=20, gate=0.0, seed=2,
                        fallback_pop=32, fallback_gens=3)
    assert r.fallback
    assert r.mc_candidates == 1 + 32 * 4
    assert np.isfinite(r.objective_mc)
