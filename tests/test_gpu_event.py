"""Event-mode kernel (the reference's one-event-per-step loop, simulate.py:46-92):
every row (lum, event time, electron ratio) bit-exact against the C replay of the
same spec, the ensemble against the numpy event-mode restatement of the
reference (tests/golden/ensemble_event_mode.json), and the driver contract of
simulate(mode=event)."""
import json
from pathlib import Path

import numpy as np
import pytest

from mcluminescence_amd import config
from mcluminescence_amd import workloads as W
from mcluminescence_amd._abi import MCL_EOVERFLOW, MclError, RunParams
from mcluminescence_amd.simulate import simulate
from oracle import orc

pytestmark = pytest.mark.gpu

FX = json.loads((Path(__file__).parent / "golden" / "ensemble_event_mode.json").read_text())
GOLDEN = json.loads((Path(__file__).parent / "golden" / "reference_outputs.json").read_text())


def _event(run: RunParams, max_dt: float, sims: int) -> RunParams:
    p = RunParams.make(**{**run.as_dict(), "sims": sims, "n_steps": 1})
    p.dt = max_dt
    return p


def _cases():
    tl12 = _event(W.tunnelling_tl12(sims=1, n=200, T_rate=1.0)[0], 1.0, 4)
    iso = _event(W.iso_decay_4096x400(sims=1)[0], 0.5, 16)
    two = W.tunnelling_tl12(sims=1, n=150, T_rate=2.0)[0]
    two.E_loc_2 = 1.25
    two.Retrap = 0.4
    two = _event(two, 1.0, 3)
    glow = _event(W.tl_glow_16x500(sims=1)[0], 0.1, 4)
    # every electron in the second channel (Retrap >= 1)
    ch2 = W.tunnelling_tl12(sims=1, n=150, T_rate=2.0)[0]
    ch2.E_loc_2 = 1.25
    ch2.Retrap = 1.0
    ch2 = _event(ch2, 1.0, 3)
    # the bench's event workload (0 -> 800 degC at 5 K/s, step cap 0.2 s) on 64 clusters:
    # most steps take the cap without forming K
    bench = _event(W.tunnelling_tl12(sims=1, n=400, T_rate=5.0)[0], 1.0, 64)
    return [("tl12_200", tl12, 800.0, 1200), ("iso_400", iso, 2000.0, 500),
            ("two_channel_150", two, 400.0, 1000), ("glow_500", glow, 400.0, 5000),
            ("always_channel_2_150", ch2, 400.0, 1000), ("tl12_5Ks_64x400", bench, 160.0, 4096)]


@pytest.mark.parametrize("name,run,duration,max_steps", _cases(), ids=[c[0] for c in _cases()])
def test_rows_bit_exact_against_replay(gpu_ctx, name, run, duration, max_steps):
    print(f"event mode {name}: {run.sims} clusters", flush=True)
    lum, t, er, steps, stats = gpu_ctx.run_event([run], [duration], max_steps)
    rl, rt, rer, rsteps = orc.run_event([run], [duration], max_steps, threads=8)
    np.testing.assert_array_equal(steps, rsteps)
    np.testing.assert_array_equal(lum, rl)
    np.testing.assert_array_equal(t, rt)
    np.testing.assert_array_equal(er, rer)
    assert stats.recombinations == int(lum.sum())
    assert np.all(steps > 0)
    for c in range(lum.shape[0]):
        n = steps[c]
        assert np.all(np.diff(t[c, :n]) > 0)                 # time only advances
        assert np.all(lum[c, n:] == 0) and np.all(t[c, n:] == 0)


@pytest.mark.parametrize("name", [k for k in FX if not k.startswith("_")])
def test_ensemble_matches_event_mode_restatement(gpu_ctx, name):
    f = FX[name]
    base = RunParams.make(**f["run"])
    sims = 2048
    dt = float(base.dt)
    max_dt = dt * (base.T_rate if base.T_rate else 1.0)
    run = _event(base, max_dt, sims)
    duration = base.n_steps * dt
    n0 = int(base.N_e * base.e_ratio_start)
    max_steps = n0 + int(base.n_steps) + 64
    lum, t, _, _, _ = gpu_ctx.run_event([run], [duration], max_steps)
    per = f["steps_per_bin"]
    nb = f["n_bins"]
    coarse = np.zeros((sims, nb))
    for c in range(sims):
        ev = t[c][lum[c] == 1]
        idx = np.floor(ev / dt).astype(np.int64) // per
        idx = idx[(idx >= 0) & (idx < nb)]
        np.add.at(coarse[c], idx, 1.0)
    mg, sg = coarse.mean(axis=0), coarse.std(axis=0, ddof=1)
    mo, so = np.array(f["mean"]), np.array(f["std"])
    pooled = (mo * f["sims"] + mg * sims) / (f["sims"] + sims)
    floor = np.minimum(pooled, 1.0)
    se = np.sqrt(np.maximum(so ** 2, floor) / f["sims"] + np.maximum(sg ** 2, floor) / sims)
    live = se > 0
    z = (mg[live] - mo[live]) / se[live]
    print(f"event ensemble {name}: max |z| {np.abs(z).max():.2f}, |z|>3: {(np.abs(z) > 3).sum()}, "
          f"chi2/dof {np.mean(z ** 2):.2f}", flush=True)
    assert (np.abs(z) > 3).sum() <= 1
    assert np.abs(z).max() < 4.5


def test_row_budget_overflow_is_reported(gpu_ctx):
    run = _event(W.iso_decay_4096x400(sims=1)[0], 0.5, 2)
    with pytest.raises(MclError) as ei:
        gpu_ctx.run_event([run], [2000.0], 100)
    assert ei.value.code == MCL_EOVERFLOW
    cfg = config.compose(["+mode=event", "exp_type_fp.steps=50"])
    with pytest.raises(IndexError):
        simulate(cfg, ctx=gpu_ctx, write_csv=False)


def test_simulate_event_mode_contract(gpu_ctx, tmp_path):
    cfg = config.compose(["+tag=gpuevent", "+mode=event"])
    print("simulate TL12 in event mode: 4 runs x 2 sims", flush=True)
    x_ax, lum, er, configs = simulate(cfg, ctx=gpu_ctx, results_dir=tmp_path)
    assert lum.shape == x_ax.shape == er.shape == (20000, 2, 4) and len(configs) == 4
    for r, rate in enumerate([0.1, 1.0, 5.0, 20.0]):
        cap = 1.0 / rate                       # max_dt 1 over T_rate
        for s in range(2):
            n = int(np.count_nonzero(x_ax[:, s, r]))
            assert set(np.unique(lum[:n, s, r])) <= {0.0, 1.0}
            assert np.all(np.diff(np.concatenate([[0.0], x_ax[:n, s, r]])) <= cap * (1 + 1e-12))
            assert x_ax[n - 1, s, r] >= 800.0 / rate
            left = 2000 - np.cumsum(lum[:n, s, r])
            np.testing.assert_allclose(er[:n, s, r], left / 2000)
            assert (lum[n:, s, r] == 0).all() and (er[n:, s, r] == 0).all()
        # totals against the reference's committed event-mode results (2 sims each)
        ref = [GOLDEN["exp_"][f"{r},{s}"]["recombinations"] for s in (0, 1)]
        got = lum[:, :, r].sum(axis=0)
        assert np.all(np.abs(got - np.mean(ref)) < 120), (r, got, ref)
    lines = (tmp_path / "exp_gpuevent.csv").read_text().splitlines()
    assert lines[0] == "run,sim,step,lum,electron_ratio" and len(lines) == 160001
