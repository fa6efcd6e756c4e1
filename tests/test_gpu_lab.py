"""The lab kernel (TL_lab / ISO_lab experiments with dose filling) on the GPU:
trapped fractions bit-exact against the C replay, its ensemble against the numpy
restatement of the reference loops, and TLTrapSim.TL_lab / ISO_lab on the
reference's lab tables read from data/processed like the reference."""
import json
from pathlib import Path

import numpy as np
import pytest

from mcluminescence_amd import config, engine, lab
from mcluminescence_amd.tl_trap_lab import TLTrapSim
from oracle import orc

pytestmark = pytest.mark.gpu

FX = json.loads((Path(__file__).parent / "golden" / "lab_ensembles.json").read_text())
LAB = json.loads((Path(__file__).parent / "golden" / "reference_outputs.json").read_text())["lab_data"]


def lab_cfg():
    return engine.initialize_runs(config.compose(["exp_type_fp=TLlab", "physics_fp=lab_TL"]))[0]


@pytest.mark.parametrize("kind,exact", [("tl", False), ("iso", False), ("tl", True), ("iso", True)])
def test_fractions_bit_exact_against_replay(gpu_ctx, kind, exact):
    print(f"lab {kind} exact_fill={exact}", flush=True)
    cfg = lab_cfg()
    if kind == "tl":
        runs, obs = lab.tl_lab_runs(cfg, LAB["CLBR_IRSL50_0.25KperGy"], sims=16)
    else:
        runs, obs, _ = lab.iso_lab_runs(cfg, LAB["CLBR_IR50_ISO"], sims=8)
    ratio, status, stats = gpu_ctx.run_lab(runs, obs, exact_fill=exact)
    ref, rstatus = orc.run_lab(runs, obs, exact_fill=exact, threads=8)
    np.testing.assert_array_equal(status, rstatus)
    np.testing.assert_array_equal(ratio, ref)
    assert stats.recombinations > 0


@pytest.mark.parametrize("kind", ["tl", "iso"])
def test_ensemble_matches_reference_restatement(gpu_ctx, kind):
    cfg = lab_cfg()
    sims = 4096
    if kind == "tl":
        runs, obs = lab.tl_lab_runs(cfg, LAB["CLBR_IRSL50_0.25KperGy"], sims=sims)
    else:
        runs, obs, _ = lab.iso_lab_runs(cfg, LAB["CLBR_IR50_ISO"], sims=sims)
    ratio, status, _ = gpu_ctx.run_lab(runs, obs)
    assert np.all(status == 0)
    means, sds = [], []
    for r, o in enumerate(obs):
        block = ratio[r * sims:(r + 1) * sims, :len(o)]
        means.append(block.mean(axis=0))
        sds.append(block.std(axis=0, ddof=1))
    mean, sd = np.concatenate(means), np.concatenate(sds)
    fx = FX[kind]
    mo, so = np.array(fx["mean"]), np.array(fx["std"])
    se = np.sqrt(np.maximum(so ** 2, 1e-4) / fx["sims"] + np.maximum(sd ** 2, 1e-4) / sims)
    z = (mean - mo) / se
    print(f"lab ensemble {kind}: {len(z)} values, max |z| {np.abs(z).max():.2f}", flush=True)
    assert (np.abs(z) > 3).sum() <= max(1, len(z) // 50)
    assert np.abs(z).max() < 4.5


def test_tl_lab_and_iso_lab_objectives(gpu_ctx, tmp_path, monkeypatch):
    import pandas as pd
    proc = tmp_path / "data" / "processed"
    proc.mkdir(parents=True)
    pd.DataFrame(LAB["CLBR_IRSL50_0.25KperGy"]).to_csv(proc / "CLBR_IRSL50_0.25KperGy.csv", index=False)
    pd.DataFrame(LAB["CLBR_IR50_ISO"]).to_csv(proc / "CLBR_IR50_ISO.csv", index=False)
    monkeypatch.setenv("MCL_PROJECT_ROOT", str(tmp_path))
    sim = TLTrapSim(lab_cfg(), ctx=gpu_ctx)
    mse_tl = sim.TL_lab("CLBR_IRSL50_0.25KperGy")
    mse_iso = sim.ISO_lab("CLBR_IR50_ISO")
    assert np.isfinite(mse_tl) and 0 <= mse_tl < 1
    assert np.isfinite(mse_iso) and 0 <= mse_iso < 1
    # averaging clusters lowers the objective's noise: the expectation of the
    # one-cluster MSE exceeds the many-cluster one by the simulation variance
    many = sim.TL_lab("CLBR_IRSL50_0.25KperGy", sims=256, verbose=False)
    fill = np.array(LAB["CLBR_IRSL50_0.25KperGy"]["Fill"])
    expect = float(np.mean((np.array(FX["tl"]["mean"]) - fill) ** 2))
    assert abs(many - expect) < 0.01
