"""The fixed-dt replay harness pinned by what the mathematics and the reference fix:

- band-only isothermal decay: N(t) = N0 (1 - p)^n, p = 1 - e^{-k dt} (closed form);
- band-only linear ramp: the Randall-Wilkins first-order peak temperature;
- tunnelling-only TL12 + basicTL12 (the reference's shipped configuration): the
  per-run recombination totals of the reference's committed result CSVs;
- p = 1 everywhere: the whole run is deterministic, so a pure-Python restatement
  of the conflict rule (lowest index keeps a shared hole, losers re-search and
  retry, at most MCL_MAX_ROUNDS rounds) with nearest holes recomputed from
  scratch must give the same per-step counts;
- the float32 path against the all-float64 mode of the same harness.
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from mcluminescence_amd import workloads as W
from mcluminescence_amd._abi import NO_HOLE, RunParams
from oracle import event_mc as E
from oracle import orc

GOLDEN = json.loads((Path(__file__).parent / "golden" / "reference_outputs.json").read_text())
MAX_ROUNDS = 12


def _band_decay_check(counts, live, N0, p):
    """Band-only survival N0 (1 - p)^n at every step with survivors expected, and
    the mean life 1 / p, within 4 standard deviations."""
    gone = np.cumsum(counts.sum(axis=0))
    n = 1
    while N0 * (1 - p) ** n >= 10:
        surv = N0 - gone[n - 1]
        expect = N0 * (1 - p) ** n
        sd = math.sqrt(N0 * (1 - p) ** n * (1 - (1 - p) ** n))
        assert abs(surv - expect) < 4 * sd + 1, (n, surv, expect, sd)
        n += 1 if n < 10 else max(1, n // 4)
    mean_life = live.sum() / N0                    # E[steps alive] = 1/p
    sd = math.sqrt((1 - p) / p ** 2 / N0)
    assert abs(mean_life - 1 / p) < 4 * sd, (mean_life, 1 / p, sd)


def test_iso_band_decay_matches_closed_form():
    r = W.iso_decay_4096x400(sims=1024)[0]
    counts, live = orc.run_fixed_dt([r], threads=8)
    k = 1e13 * math.exp(-1.4 / (8.617343e-5 * 500.0))
    p = -math.expm1(-k * 0.5)
    assert k == pytest.approx(0.07738, rel=1e-3) and p == pytest.approx(0.03795, rel=1e-3)
    _band_decay_check(counts, live, 400 * 1024, p)


@pytest.mark.parametrize("T_K,p_expect", [(540.0, 0.349), (560.0, 0.716)])
def test_iso_band_decay_closed_form_where_conflicts_bind(T_K, p_expect):
    # at p = 0.35 / 0.71 a third to two thirds of the electrons fire each step and
    # many share a nearest hole: the survival is exact only if every conflict loser
    # gets its same-step retry (the T_new / T_old thinning on its new hole) and the
    # MCL_MAX_ROUNDS cap never binds
    r = W.iso_decay_4096x400(sims=1024)[0]
    r.T_start = T_K - 273.15
    counts, live = orc.run_fixed_dt([r], threads=8)
    p = -math.expm1(-1e13 * math.exp(-1.4 / (8.617343e-5 * T_K)) * 0.5)
    assert p == pytest.approx(p_expect, abs=0.001)
    _band_decay_check(counts, live, 400 * 1024, p)


def test_tl_band_peak_within_1K_of_randall_wilkins():
    r = W.tl_glow_16x500(sims=512)[0]
    counts, _ = orc.run_fixed_dt([r], threads=8)
    T = 300.0 + 0.1 * np.arange(4000)
    sm = np.convolve(counts.sum(axis=0).astype(float), np.ones(21) / 21, "same")
    i = int(np.argmax(sm))
    c = np.polyfit(T[i - 60:i + 61], sm[i - 60:i + 61], 2)
    Tm = -c[1] / (2 * c[0])
    T_rw = E.randall_wilkins_peak(1.3, 1e12, 1.0)
    assert T_rw == pytest.approx(495.88, abs=0.01)
    assert abs(Tm - T_rw) < 1.0


@pytest.mark.parametrize("run_idx,rate", [(0, 0.1), (1, 1.0), (2, 5.0), (3, 20.0)])
def test_tl12_totals_match_reference_csv(run_idx, rate):
    ref = [GOLDEN["exp_"][f"{run_idx},{s}"]["recombinations"] for s in (0, 1)]
    ref += [GOLDEN["exp_TLwithDose"][f"{run_idx},{s}"]["recombinations"] for s in (0, 1)]
    r = W.tunnelling_tl12(sims=32, n=2000, T_rate=rate)[0]
    counts, _ = orc.run_fixed_dt([r], threads=8)
    tot = counts.sum(axis=1).astype(float)
    # the reference's four sims against our 32: Welch-style 3 sigma with the pooled spread
    se = math.sqrt(tot.var(ddof=1) / len(tot) + tot.var(ddof=1) / len(ref))
    assert abs(tot.mean() - np.mean(ref)) < 3 * se, (tot.mean(), ref)


def test_tl12_emptystart_totals_match_reference_csv():
    for run_idx, rate in enumerate([0.1, 1.0, 5.0, 20.0]):
        ref = [GOLDEN["exp_TLwithDose_emptystart"][f"{run_idx},{s}"]["recombinations"]
               for s in (0, 1)]
        r = W.tunnelling_tl12(sims=64, n=2000, T_rate=rate)[0]
        r.e_ratio_start = 0.05                      # 100 electrons, 3455 holes
        counts, _ = orc.run_fixed_dt([r], threads=8)
        tot = counts.sum(axis=1).astype(float)
        se = math.sqrt(tot.var(ddof=1) / len(tot) + tot.var(ddof=1) / len(ref))
        assert abs(tot.mean() - np.mean(ref)) < 3 * se, (rate, tot.mean(), ref)


def _p1_run(n=40, steps=6):
    # E = 1 eV, s = 1e16 at 600 K: k dt ~ 4e7, p = 1 for every electron
    return W.single_trap(1.0, 1e16, 4e9, 1e23, n, T_start_K=600.0, T_rate=0.0, dt=1.0,
                         n_steps=steps, sims=1)


def _bruteforce_p1_counts(run: RunParams, run_index: int, sim: int):
    st = orc.init_state(run, run_index, sim)
    e_xyz, h_xyz = st["e_xyz"], st["h_xyz"]
    alive = np.ones(len(h_xyz), np.uint8)
    live = list(range(len(e_xyz)))
    nn = orc.bruteforce_nn(e_xyz, h_xyz, alive)
    assert np.array_equal(nn, st["nn"])
    counts = []
    for _ in range(run.n_steps):
        if not live:
            counts.append(0)
            continue
        count = 0
        cand = [e for e in live if nn[e] != NO_HOLE]
        for rnd in range(MAX_ROUNDS):
            taken, winners = set(), []
            for e in cand:                       # ascending electron index
                if nn[e] not in taken:
                    taken.add(int(nn[e]))
                    winners.append(e)
            losers = [e for e in cand if e not in set(winners)]
            for e in winners:
                alive[nn[e]] = 0
            live = [e for e in live if e not in set(winners)]
            count += len(winners)
            full = orc.bruteforce_nn(e_xyz, h_xyz, alive)   # from scratch
            for e in live:
                nn[e] = full[e]
            if rnd == MAX_ROUNDS - 1:
                break
            cand = [e for e in losers if nn[e] != NO_HOLE]  # T_new = T_old = 2^32: accepted
            if not cand:
                break
        counts.append(count)
    return np.array(counts, np.uint32)


@pytest.mark.parametrize("n,run_index,sim", [(40, 0, 0), (64, 3, 1), (150, 1, 7)])
def test_conflict_rounds_match_bruteforce_at_p1(n, run_index, sim):
    run = _p1_run(n)
    want = _bruteforce_p1_counts(run, run_index, sim)
    run.sims = sim + 1
    counts, _ = orc.run_fixed_dt([RunParams.make(**{**run.as_dict()})] * (run_index + 1))
    got = counts[run_index * (sim + 1) + sim]
    np.testing.assert_array_equal(got, want)
    assert want[0] > 0 and want.sum() <= n


def test_float32_path_equals_float64_mode():
    for runs in (W.tl_glow_16x500(sims=64), W.tunnelling_tl12(sims=8, n=1000, T_rate=5.0),
                 W.iso_decay_4096x400(sims=64)):
        c32, l32 = orc.run_fixed_dt(runs, threads=8)
        c64, l64 = orc.run_fixed_dt(runs, fp64=True, threads=8)
        # per-cluster cumulative curves; identical on these workloads
        diff = np.abs(np.cumsum(c32.astype(np.int64) - c64.astype(np.int64), axis=1)).max()
        assert diff <= 2


def test_deterministic_and_seed_sensitive():
    runs = W.tl_glow_16x500(sims=4)
    a, _ = orc.run_fixed_dt(runs, seed=1)
    b, _ = orc.run_fixed_dt(runs, seed=1)
    c, _ = orc.run_fixed_dt(runs, seed=2)
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(a, c)
    assert a.sum() == c.sum() == 4 * 500


def test_binned_counts_are_sums_of_step_counts():
    runs = W.tl_glow_16x500(sims=3)
    per_step, _ = orc.run_fixed_dt(runs)
    binned, _ = orc.run_fixed_dt(runs, n_bins=256)
    n = np.arange(4000)
    b = (n * 256) // 4000
    want = np.zeros((3, 256), np.uint64)
    for j in range(256):
        want[:, j] = per_step[:, b == j].sum(axis=1)
    np.testing.assert_array_equal(binned, want)


def _event_bins(name, sims=1024):
    fx = json.loads((Path(__file__).parent / "golden" / "ensemble_event_mode.json").read_text())
    f = fx[name]
    base = RunParams.make(**f["run"])
    dt = float(base.dt)
    p = RunParams.make(**{**base.as_dict(), "sims": sims, "n_steps": 1})
    p.dt = dt * (base.T_rate if base.T_rate else 1.0)               # max_dt
    n0 = int(base.N_e * base.e_ratio_start)
    lum, t, _, steps = orc.run_event([p], [base.n_steps * dt], n0 + int(base.n_steps) + 64,
                                     threads=8)
    assert np.all(steps > 0)
    per, nb = f["steps_per_bin"], f["n_bins"]
    coarse = np.zeros((sims, nb))
    for c in range(sims):
        idx = np.floor(t[c][lum[c] == 1] / dt).astype(np.int64) // per
        idx = idx[(idx >= 0) & (idx < nb)]
        np.add.at(coarse[c], idx, 1.0)
    return f, coarse


@pytest.mark.parametrize("name", ["iso_500K_400", "tl_glow_500", "tl12_tunnelling_400"])
def test_event_replay_matches_event_mode_restatement(name):
    # The event spec (total-rate wait, winner by k_e / K) against the numpy
    # restatement of the reference loop (per-electron exponential waits, argmin),
    # per bin within 3 standard errors (Poisson floor for rare bins).
    f, coarse = _event_bins(name)
    S = coarse.shape[0]
    mg, sg = coarse.mean(axis=0), coarse.std(axis=0, ddof=1)
    mo, so = np.array(f["mean"]), np.array(f["std"])
    pooled = (mo * f["sims"] + mg * S) / (f["sims"] + S)
    floor = np.minimum(pooled, 1.0)
    se = np.sqrt(np.maximum(so ** 2, floor) / f["sims"] + np.maximum(sg ** 2, floor) / S)
    z = (mg[se > 0] - mo[se > 0]) / se[se > 0]
    assert (np.abs(z) > 3).sum() <= 1 and np.abs(z).max() < 4.5, np.abs(z).max()


def test_event_replay_terminal_rows_follow_the_reference_loop():
    # an emptied isothermal run ends with one step of dt = min(dt_fill, 1e20)
    # (simulate.py:57-60 with dose 0): a fill row (ratio 1/N_e) or a no-event row
    iso = W.iso_decay_4096x400(sims=64)[0]
    iso.n_steps = 1
    lum, t, er, steps = orc.run_event([iso], [2000.0], 600, threads=8)
    assert np.all(steps == 401)
    assert np.all(lum[:, :400] == 1) and np.all(lum[:, 400] == 0)
    assert np.all(t[:, 400] > 1e10)                 # 1e20 x Exp(1) or the 1e20 cap
    assert set(np.unique(er[:, 400])) <= {0.0, 1.0 / 400}
