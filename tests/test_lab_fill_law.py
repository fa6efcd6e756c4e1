"""The dose-filling law of the lab loops, pinned outside the restatements:

1. Recombination off (every channel at 1000 eV, so every rate underflows to 0):
   TLTrapSim._filling_time (src/class/tl_trap_lab.py:53-60) fills at
   lambda = (D / D0)(N_e - n_e), i.e. each empty trap independently at rate
   D / D0.  So n_e(t) - n_e(0) ~ Binomial(N_e - n_e(0), 1 - e^{-(D/D0) t}), mean
   n_e / N_e = 1 - (1 - f0) e^{-(D/D0) t}.  The loops record the fraction after
   the first event at or past an observation time (tl_trap_lab.py:90-108,
   147-172): one fill more, even when every trap is full (then lambda = 1e-20
   still draws a fill, tl_trap_lab.py:56-60, and the fraction reads
   (N_e + 1) / N_e).  The recorded value's mean and variance follow from the
   binomial pmf (oracle.analytic.fill_observed_moments).  Checked for the replay harness (orc_run_lab, the lab
   kernel's spec) and for the numpy restatement of ISO_lab (oracle/lab_mc.py).
2. The reference's own fits: the ten-parameter rows of
   results/lab_sims/result_tl_clbr.csv:2-6 and their recorded MSEs.  Those rows
   were written by the legacy driver (src/est_params/estimate.py:60-79 puts new
   rows first) with the legacy loop sim_lab_TL_residuals
   (src/est_params/functions.py:270-360): the lab kernel's MCL_LAB_LEGACY
   options, on the TLlab / lab_TL configuration (N_e = 100, D = 0.092 Gy/s).  A
   recorded MSE is the best of many noisy evaluations (DE with popsize 15 over
   ten parameters: 150 candidates a generation), so the check is against the
   minimum of 150 independent evaluations at the recorded parameters: its
   central 99 % holds the recorded value.  The class-based loop
   (tl_trap_lab.py:65-123) cannot produce these values (none of 4000
   evaluations reaches them), which is how the rows' provenance is known.
   Rows 7-9 hold nine parameters (an older bound set without the retrap
   probability) and are not used.
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from mcluminescence_amd import config, engine, lab
from mcluminescence_amd._abi import LAB_LEGACY
from mcluminescence_amd.optimizer import cfg_with_params
from oracle import analytic as A
from oracle import orc

REF = json.loads((Path(__file__).parent / "golden" / "reference_outputs.json").read_text())
TL_TABLE = REF["lab_data"]["CLBR_IRSL50_0.25KperGy"]
FIT_ROWS = ("2", "3", "4", "5", "6")
TIMES = [30.0, 300.0, 1000.0, 3000.0, 10000.0, 30000.0]
DE_CANDIDATES = 150


def _base():
    return config.compose(["exp_type_fp=TLlab", "physics_fp=lab_TL"])


def _no_recombination(f0, sims):
    cfg = _base()
    cfg.physics_fp.E_cb = 1000.0
    cfg.physics_fp.E_loc_1 = 1000.0
    cfg.physics_fp.E_loc_2 = 1000.0
    run = engine.initialize_runs(cfg)[0]
    table = {"exp_no": [0] * len(TIMES), "time": TIMES, "temp": [20.0] * len(TIMES),
             "dose": [0.092] * len(TIMES), "e_ratio": [f0] * len(TIMES)}
    runs, obs, _ = lab.iso_lab_runs(run, table, sims=sims)
    return cfg, runs, obs


def _check_binomial(frac, f0, N, D, D0):
    """frac [sims, len(TIMES)] against the binomial law of the filling process."""
    n0 = int(N * f0)
    for j, t in enumerate(TIMES):
        mean, var = A.fill_observed_moments(n0, N, D, D0, t)
        # the recorded value is the closed-form fraction plus the crossing fill
        assert mean == pytest.approx(A.fill_fraction(n0 / N, D, D0, t) + 1.0 / N, rel=1e-9)
        x = frac[:, j]
        se = math.sqrt(max(var, 1e-8) / x.size)
        assert abs(x.mean() - mean) < 4 * se, (t, x.mean(), mean, se)
        if var > 1e-6:
            # sample variance of a binomial: relative SE about sqrt(2 / n)
            assert abs(x.var(ddof=1) / var - 1) < 4 * math.sqrt(2.0 / x.size) + 0.02, (t, x.var(), var)


@pytest.mark.parametrize("f0", [0.0, 0.3])
def test_replay_fill_only_is_binomial(f0):
    cfg, runs, obs = _no_recombination(f0, sims=4096)
    ratio, status = orc.run_lab(runs, obs, threads=8)
    assert np.all(status == 0)
    assert np.all(np.diff(ratio, axis=1) >= 0)          # nothing ever leaves
    _check_binomial(ratio, f0, int(cfg.exp_type_fp.N_e), 0.092, float(cfg.physics_fp.D0))


def test_restatement_fill_only_is_binomial():
    from oracle import lab_mc
    cfg, runs, _ = _no_recombination(0.3, sims=1)
    exp = dict(cfg.exp_type_fp.to_dict())
    phys = dict(cfg.physics_fp.to_dict())
    rng = np.random.default_rng(11)
    ex = {"e_ratio": 0.3, "temp": 20.0, "dose": 0.092, "times": TIMES}
    frac = np.array([lab_mc.iso_lab_fractions(exp, phys, [ex], rng)[0] for _ in range(400)])
    _check_binomial(frac, 0.3, int(exp["N_e"]), 0.092, float(phys["D0"]))


def _fit_row(key):
    row = REF["lab_fit_results"]["result_tl_clbr"]["rows"][key]
    params = np.array([float(v) for v in row[:10]])
    return params, float(row[10])


def _mse_distribution(params, sims, options, run_lab):
    run = engine.initialize_runs(cfg_with_params(_base(), params))[0]
    runs, obs = lab.tl_lab_runs(run, TL_TABLE, sims=sims)
    ratio, status = run_lab(runs, obs, options)
    assert np.all(status == 0)
    final = ratio[:, 0].reshape(len(runs), sims)
    err = final - np.asarray(TL_TABLE["Fill"], np.float64)[:, None]
    return np.mean(err ** 2, axis=0)              # one MSE per simulation (sims = 1 each)


def _check_recorded_fits(run_lab, sims):
    for key in FIT_ROWS:
        params, recorded = _fit_row(key)
        mse = _mse_distribution(params, sims, LAB_LEGACY, run_lab)
        F = float(np.mean(mse <= recorded))
        p_min = 1.0 - (1.0 - F) ** DE_CANDIDATES
        print(f"row {key}: recorded {recorded:.4e}, legacy-loop median {np.median(mse):.4e}, "
              f"F = {F:.4f}, P(min of {DE_CANDIDATES} <= recorded) = {p_min:.3f}", flush=True)
        assert 0.005 <= p_min <= 0.995, (key, recorded, F, p_min)
        assert recorded < np.median(mse)
        cls = _mse_distribution(params, sims, 0, run_lab)
        assert np.mean(cls <= recorded) < 1e-3, (key, np.median(cls))


def test_recorded_reference_fits_match_the_legacy_loop():
    _check_recorded_fits(lambda runs, obs, opt: orc.run_lab(runs, obs, threads=8, options=opt),
                         sims=4000)


@pytest.mark.gpu
def test_gpu_lab_kernel_matches_recorded_reference_fits(gpu_ctx):
    print("lab kernel, legacy options: recorded TL CLBR fits, 5 rows x 11 lab rows x 4000 sims",
          flush=True)

    def run_lab(runs, obs, opt):
        ratio, status, _ = gpu_ctx.run_lab(runs, obs, options=opt)
        return ratio, status

    _check_recorded_fits(run_lab, sims=4000)
