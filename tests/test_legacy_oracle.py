"""The legacy lab loops (src/est_params/functions.py:51-434): the numpy
restatement (oracle/legacy_mc.py) and the lab replay's MCL_LAB_* options
(oracle/mcl_oracle.c, the spec of the GPU lab kernel), pinned by:

  * numpy's own semantics for the recombination pick (the np.where call of
    recomber on a (1, n) mask returns row indices) and the box arithmetic worked
    by hand;
  * an invariant: with Retrap = 1 and no dose every recombination is followed by
    a new pair, so n_e never changes;
  * a closed form for the shared channel draw: with channel 2 frozen, every step
    freezes all electrons at once with probability 1/2, so the number of
    recombinations before the run stalls is geometric;
  * band-only decay n0 e^{-kt} (read one event after t) when the pick rule
    cannot matter;
  * the replay's ensembles against the restatement's (tests/golden/
    legacy_ensembles.json, tests/golden/make_legacy_fixtures.py).
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from mcluminescence_amd import config, engine, legacy
from mcluminescence_amd._abi import (LAB_FIRST_ELECTRON, LAB_LEGACY, LAB_RETRAP,
                                     LAB_SHARED_CHANNEL, RunParams)
from oracle import legacy_mc, orc

GOLD = Path(__file__).parent / "golden"
LAB = json.loads((GOLD / "reference_outputs.json").read_text())["lab_data"]


def run(phys, exp, *, T_C=27.0, T_rate=0.0, e_ratio=0.5, D=0.0, sims=1):
    return RunParams.make(alpha=phys["alpha"], b=phys["b"], s=phys["s"], E_cb=phys["E_cb"],
                          E_loc_1=phys["E_loc_1"], E_loc_2=phys["E_loc_2"], D0=phys["D0"],
                          Retrap=phys["Retrap"], k_b=phys["k_b"], D=D,
                          rho_prime=exp["rho_prime"], holes=exp["holes"],
                          boundary_factor=exp["boundary_factor"], N_e=exp["N_e"],
                          e_ratio_start=e_ratio, T_start=T_C, T_rate=T_rate, dt=1.0, n_steps=1,
                          sims=sims)


EXP = dict(N_e=100.0, holes=100.0, rho_prime=1e-4, boundary_factor=1.2)
BASE = dict(alpha=9e9, b=1e12, k_b=8.617343e-5, s=1e13, D=0.0, D0=400.35, E_cb=1.8,
            E_loc_1=1.4, E_loc_2=1.4, Retrap=0.5)


def test_recomber_mask_yields_row_indices():
    # recomber (functions.py:241-261) compares against waits + np.zeros([1, 1]): a
    # (1, n) mask whose np.where(...)[0] is the row index 0 once per expired wait
    rng = np.random.default_rng(1)
    for _ in range(20):
        waits = rng.exponential(1.0, size=rng.integers(2, 40))
        t = float(np.sort(waits)[rng.integers(0, waits.size)])
        idx = np.where(t >= waits + np.zeros([1, 1]))[0]
        assert idx.size == int((waits <= t).sum()) and np.all(idx == 0)
    # so the restatement removes the first electron whatever the waits are
    box = legacy_mc.LegacyBox(EXP, BASE, 0.3, np.random.default_rng(2))
    before = box.e.copy()
    box.waits = np.linspace(5.0, 0.1, box.e.shape[0])        # the last electron expires first
    assert box.recombine(0.1) == 0
    np.testing.assert_array_equal(box.e, before[1:])
    # and keeps nearest holes exact
    d = np.linalg.norm(box.e[:, None, :] - box.h[None, :, :], axis=2)
    np.testing.assert_array_equal(box.nn, d.argmin(axis=1))


def test_legacy_box_worked_example():
    # holes = 437.6, bf = 1.2: the legacy box keeps int(437.6) = 437 core holes and
    # adds int(437 * (1.2^3 - 1)) = int(318.136) = 318; the class box has
    # int(437.6 * 1.728) = int(756.1728) = 756 holes
    exp = dict(EXP, holes=437.6)
    n_e, n_h, d, db = legacy_mc.box_geometry(exp, BASE, 0.25)
    assert (n_e, n_h) == (25, 755)
    rho = 1e-4 * (3 / (4 * math.pi) * 9e9 ** 3)
    assert d == pytest.approx((437 / rho) ** (1 / 3), rel=1e-15)
    assert db == pytest.approx(1.2 * d, rel=1e-15)
    p = run(BASE, exp, e_ratio=0.25)
    gl, gc = orc.geometry_legacy(p), orc.geometry(p)
    assert (gl.n_e0, gl.n_h) == (25, 755) and gc.n_h == 756
    assert gl.side_m == pytest.approx(d, rel=1e-15)
    assert gc.side_m == pytest.approx((437.6 / rho) ** (1 / 3), rel=1e-15)


def test_full_retrapping_keeps_the_count():
    # Retrap = 1, D = 0, band only: each recombination is followed by a new pair
    phys = dict(BASE, b=0.0, E_cb=0.6, Retrap=1.0)
    rng = np.random.default_rng(3)
    for _ in range(3):
        row = dict(Duration=0.01, T_start=27.0, T_end=27.0, e_ratio=0.5)
        assert legacy_mc.tl_fraction(EXP, phys, row, rng) == 0.5
    p = run(phys, EXP, e_ratio=0.5, sims=64)
    ratio, status = orc.run_lab([p], [[0.005, 0.01]], options=LAB_LEGACY, threads=8)
    assert np.all(status == 0) and np.all(ratio == 0.5)
    # without retrapping the same runs decay
    ratio2, _ = orc.run_lab([p], [[0.005, 0.01]], options=LAB_LEGACY & ~LAB_RETRAP, threads=8)
    assert ratio2.mean() < 0.45


def freeze_law(max_m: int = 80, retrap: float = 0.5):
    """P(n_e - n0 = d) when each step stalls with probability 1/2 (then one spurious
    fill, +1) and otherwise recombines (-1) and retraps with probability `retrap`."""
    probs = {}
    for m in range(max_m):
        pm = 0.5 ** (m + 1)
        for k in range(m + 1):
            pk = math.comb(m, k) * retrap ** k * (1 - retrap) ** (m - k)
            d = 1 - m + k
            probs[d] = probs.get(d, 0.0) + pm * pk
    return probs


def test_shared_channel_freeze_law():
    # band off (E_cb = 100), channel 1 instant (E_loc_1 = 0, b = 1e20), channel 2
    # frozen (E_loc_2 = 100): a step with the shared draw on channel 2 has K = 0,
    # so the run waits for the Exp(1e20) fill and the observation at 1 s sees it
    phys = dict(BASE, b=1e20, E_cb=100.0, E_loc_1=0.0, E_loc_2=100.0, Retrap=0.5)
    law = freeze_law()
    n0 = 50
    p = run(phys, EXP, e_ratio=0.5, sims=4096)
    ratio, status = orc.run_lab([p], [[1.0]], options=LAB_LEGACY, threads=8)
    assert np.all(status == 0)
    delta = np.round(ratio[:, 0] * 100).astype(int) - n0
    _check_law(delta, law)
    # the numpy restatement follows the same law
    rng = np.random.default_rng(4)
    row = dict(Duration=1.0, T_start=27.0, T_end=27.0, e_ratio=0.5)
    d_np = np.array([round(legacy_mc.tl_fraction(EXP, phys, row, rng) * 100) - n0
                     for _ in range(300)])
    _check_law(d_np, law)
    # per-electron draws almost never freeze all 50 electrons: the run empties instead
    ratio_pe, _ = orc.run_lab([p], [[1.0]], options=LAB_LEGACY & ~LAB_SHARED_CHANNEL, threads=8)
    assert np.mean(np.round(ratio_pe[:, 0] * 100) - n0) < -20


def _check_law(delta, law):
    n = delta.size
    mean = sum(d * q for d, q in law.items())
    var = sum((d - mean) ** 2 * q for d, q in law.items())
    assert abs(delta.mean() - mean) < 4 * math.sqrt(var / n), (delta.mean(), mean)
    for d in range(-3, 2):
        q = law.get(d, 0.0)
        got = np.mean(delta == d)
        assert abs(got - q) < 4 * math.sqrt(q * (1 - q) / n) + 1e-9, (d, got, q)


def test_band_only_decay_closed_form():
    # no tunnelling, no dose, no retrapping (the isothermal loop): the pick rule
    # cannot matter and n(t) ~ Binomial(n0, e^{-kt}), k = s e^{-E_cb/kT}
    phys = dict(BASE, b=0.0, E_cb=0.8)
    T = 300.0
    k = 1e13 * math.exp(-0.8 / (8.617343e-5 * T))
    times = [0.5 / k, 1.0 / k, 2.0 / k]
    n0 = 60
    p = run(phys, EXP, T_C=T - 273.15, e_ratio=0.6, sims=4096)
    ratio, status = orc.run_lab([p], [times], options=LAB_LEGACY & ~LAB_RETRAP, threads=8)
    assert np.all(status == 0)
    rng = np.random.default_rng(5)
    ex = dict(e_ratio=0.6, temp=T - 273.15, dose=0.0, times=times)
    np_frac = np.array([legacy_mc.iso_fractions(EXP, phys, ex, rng) for _ in range(400)])
    for j, t in enumerate(times):
        # the fraction is read after the first event past t: one recombination later
        # than n(t) (an empty cluster, probability (1 - q)^60 < 2e-4, fills instead)
        q = math.exp(-k * t)
        want, sd = (n0 * q - 1) / 100, math.sqrt(n0 * q * (1 - q)) / 100
        assert abs(ratio[:, j].mean() - want) < 4 * sd / math.sqrt(4096)
        assert abs(np_frac[:, j].mean() - want) < 4 * sd / math.sqrt(400)


def _z(mean, sd, n, fx):
    mo, so = np.array(fx["mean"]), np.array(fx["std"])
    floor = np.full_like(mo, 1e-4)        # fractions are multiples of 1/N_e = 0.01
    se = np.sqrt(np.maximum(so ** 2, floor) / fx["sims"] + np.maximum(sd ** 2, floor) / n)
    return (mean - mo) / se


@pytest.mark.parametrize("kind", ["tl", "iso"])
def test_legacy_replay_matches_restatement(kind):
    fx = json.loads((GOLD / "legacy_ensembles.json").read_text())
    cfg = engine.initialize_runs(config.compose([
        "exp_type_fp=TLlab", "physics_fp=lab_TL", f"physics_fp.E_loc_2={fx['phys']['E_loc_2']}",
        f"exp_type_fp.holes={fx['exp']['holes']}"]))[0]
    sims = 2048
    if kind == "tl":
        runs, obs, _ = legacy._tl_setup(cfg, legacy.lab._as_frame(LAB["CLBR_IRSL50_0.25KperGy"]),
                                        sims)
        opts = legacy.OPT_TL
    else:
        runs, obs, _ = legacy._iso_setup(cfg, legacy.lab._as_frame(LAB["CLBR_IR50_ISO"]), sims)
        opts = legacy.OPT_ISO
    ratio, status = orc.run_lab(runs, obs, options=opts, threads=8)
    assert np.all(status == 0)
    means, sds = [], []
    for r, o in enumerate(obs):
        block = ratio[r * sims:(r + 1) * sims, :len(o)]
        means.append(block.mean(axis=0))
        sds.append(block.std(axis=0, ddof=1))
    z = _z(np.concatenate(means), np.concatenate(sds), sims, fx[kind])
    assert z.size == len(fx[kind]["mean"])
    assert (np.abs(z) > 3).sum() <= max(1, len(z) // 50)
    assert np.abs(z).max() < 4.5, np.abs(z).max()


def test_first_electron_option_changes_only_the_pick():
    # same stream, same timing draws: with tunnelling the pick changes the
    # fractions; with band-only physics and equal rates the counts are alike in law
    phys = dict(BASE, E_loc_1=0.9, E_loc_2=0.9)
    p = run(phys, EXP, T_C=200.0, e_ratio=0.8, sims=256)
    a, _ = orc.run_lab([p], [[100.0, 1000.0]], options=LAB_FIRST_ELECTRON, threads=8)
    b, _ = orc.run_lab([p], [[100.0, 1000.0]], options=0, threads=8)
    assert not np.array_equal(a, b)
    assert np.all((a >= 0) & (a <= 0.8))
