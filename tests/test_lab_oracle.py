"""The lab-experiment replay (oracle/mcl_oracle.c, orc_run_lab: the spec of the GPU
lab kernel) against the numpy restatement of the reference's TL_lab / ISO_lab
loops (tests/golden/lab_ensembles.json, made by tests/golden/make_lab_fixtures.py
from oracle/lab_mc.py): per-row / per-observation mean trapped fractions within
3 standard errors."""
import json
from pathlib import Path

import numpy as np
import pytest

from mcluminescence_amd import config, engine, lab
from oracle import orc

FX = json.loads((Path(__file__).parent / "golden" / "lab_ensembles.json").read_text())
LAB = json.loads((Path(__file__).parent / "golden" / "reference_outputs.json").read_text())["lab_data"]


def lab_cfg():
    return engine.initialize_runs(config.compose(["exp_type_fp=TLlab", "physics_fp=lab_TL"]))[0]


def _z(mean, sd, n, fx):
    mo, so = np.array(fx["mean"]), np.array(fx["std"])
    floor = np.full_like(mo, 1e-4)        # fractions are multiples of 1/N_e = 0.01
    se = np.sqrt(np.maximum(so ** 2, floor) / fx["sims"] + np.maximum(sd ** 2, floor) / n)
    return (mean - mo) / se


@pytest.mark.parametrize("kind", ["tl", "iso"])
def test_lab_replay_matches_reference_restatement(kind):
    cfg = lab_cfg()
    sims = 2048
    if kind == "tl":
        runs, obs = lab.tl_lab_runs(cfg, LAB["CLBR_IRSL50_0.25KperGy"], sims=sims)
    else:
        runs, obs, _ = lab.iso_lab_runs(cfg, LAB["CLBR_IR50_ISO"], sims=sims)
    ratio, status = orc.run_lab(runs, obs, threads=8)
    assert np.all(status == 0)
    means, sds = [], []
    for r, o in enumerate(obs):
        block = ratio[r * sims:(r + 1) * sims, :len(o)]
        means.append(block.mean(axis=0))
        sds.append(block.std(axis=0, ddof=1))
    z = _z(np.concatenate(means), np.concatenate(sds), sims, FX[kind])
    assert (np.abs(z) > 3).sum() <= max(1, len(z) // 50)
    assert np.abs(z).max() < 4.5, np.abs(z).max()


def test_exact_fill_only_changes_nearest_holes():
    # the exact update differs from the reference's stale one, but both keep every
    # fraction in [0, 1] on the 1/N_e grid and both fill the empty start
    cfg = lab_cfg()
    runs, obs = lab.tl_lab_runs(cfg, LAB["CLBR_IRSL50_0.25KperGy"], sims=64)
    a, sa = orc.run_lab(runs, obs, threads=8)
    b, sb = orc.run_lab(runs, obs, exact_fill=True, threads=8)
    assert np.all(sa == 0) and np.all(sb == 0)
    for x in (a, b):
        assert np.all((x >= 0) & (x <= 1))
        np.testing.assert_allclose(x * 100, np.round(x * 100), atol=1e-9)
    assert not np.array_equal(a, b)
