"""The lab kernel's legacy options (MCL_LAB_*; the loops of
src/est_params/functions.py:270-434) on the GPU: trapped fractions bit-exact
against the C replay for every option alone and together, the legacy ensembles
against the numpy restatement (tests/golden/legacy_ensembles.json), and the
legacy objectives and DE driver of legacy.py."""
import json
from pathlib import Path

import numpy as np
import pandas as pd
import pytest

from mcluminescence_amd import config, engine, legacy
from mcluminescence_amd._abi import (LAB_EXACT_FILL, LAB_FIRST_ELECTRON, LAB_LEGACY,
                                     LAB_LEGACY_GEOMETRY, LAB_RETRAP, LAB_SHARED_CHANNEL)
from oracle import orc

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).parent / "golden"
FX = json.loads((GOLD / "legacy_ensembles.json").read_text())
LAB = json.loads((GOLD / "reference_outputs.json").read_text())["lab_data"]
TL = legacy.lab._as_frame(LAB["CLBR_IRSL50_0.25KperGy"])
ISO = legacy.lab._as_frame(LAB["CLBR_IR50_ISO"])


def legacy_cfg():
    return engine.initialize_runs(config.compose([
        "exp_type_fp=TLlab", "physics_fp=lab_TL", f"physics_fp.E_loc_2={FX['phys']['E_loc_2']}",
        f"exp_type_fp.holes={FX['exp']['holes']}"]))[0]


@pytest.mark.parametrize("kind,options", [
    ("tl", LAB_SHARED_CHANNEL), ("tl", LAB_RETRAP), ("tl", LAB_FIRST_ELECTRON),
    ("tl", LAB_LEGACY_GEOMETRY), ("tl", LAB_RETRAP | LAB_EXACT_FILL), ("tl", LAB_LEGACY),
    ("iso", LAB_LEGACY & ~LAB_RETRAP), ("iso", LAB_LEGACY)])
def test_legacy_options_bit_exact_against_replay(gpu_ctx, kind, options):
    print(f"legacy lab {kind} options=0x{options:x}", flush=True)
    cfg = legacy_cfg()
    if kind == "tl":
        runs, obs, _ = legacy._tl_setup(cfg, TL, 16)
    else:
        runs, obs, _ = legacy._iso_setup(cfg, ISO, 8)
    ratio, status, stats = gpu_ctx.run_lab(runs, obs, options=options)
    ref, rstatus = orc.run_lab(runs, obs, options=options, threads=8)
    np.testing.assert_array_equal(status, rstatus)
    np.testing.assert_array_equal(ratio, ref)
    assert stats.recombinations > 0


def test_unknown_option_bits_are_rejected(gpu_ctx):
    from mcluminescence_amd._abi import MCL_EINVAL, MclError
    runs, obs, _ = legacy._tl_setup(legacy_cfg(), TL, 1)
    with pytest.raises(MclError) as e:
        gpu_ctx.run_lab(runs, obs, options=0x40)
    assert e.value.code == MCL_EINVAL


@pytest.mark.parametrize("kind", ["tl", "iso"])
def test_legacy_ensemble_matches_restatement(gpu_ctx, kind):
    cfg = legacy_cfg()
    sims = 4096
    if kind == "tl":
        runs, obs, _ = legacy._tl_setup(cfg, TL, sims)
        opts = legacy.OPT_TL
    else:
        runs, obs, _ = legacy._iso_setup(cfg, ISO, sims)
        opts = legacy.OPT_ISO
    ratio, status, _ = gpu_ctx.run_lab(runs, obs, options=opts)
    assert np.all(status == 0)
    means, sds = [], []
    for r, o in enumerate(obs):
        block = ratio[r * sims:(r + 1) * sims, :len(o)]
        means.append(block.mean(axis=0))
        sds.append(block.std(axis=0, ddof=1))
    mean, sd = np.concatenate(means), np.concatenate(sds)
    fx = FX[kind]
    mo, so = np.array(fx["mean"]), np.array(fx["std"])
    assert mean.size == mo.size
    se = np.sqrt(np.maximum(so ** 2, 1e-4) / fx["sims"] + np.maximum(sd ** 2, 1e-4) / sims)
    z = (mean - mo) / se
    print(f"legacy ensemble {kind}: {len(z)} values, max |z| {np.abs(z).max():.2f}", flush=True)
    assert (np.abs(z) > 3).sum() <= max(1, len(z) // 50)
    assert np.abs(z).max() < 4.5


def test_legacy_objectives_and_batched_population(gpu_ctx):
    cfg = config.compose(["exp_type_fp=TLlab", "physics_fp=lab_TL", "exp_type_fp.sims=64"])
    run_cfg = engine.initialize_runs(cfg)[0]
    tl = legacy.sim_lab_TL_residuals(run_cfg, ctx=gpu_ctx, table=TL)
    iso = legacy.sim_lab_TL_residuals_iso(run_cfg, ctx=gpu_ctx, table=ISO)
    assert 0 <= tl < 1 and 0 <= iso < 1
    rng = np.random.default_rng(5)
    lo = np.array([b[0] for b in legacy.BOUNDS])
    hi = np.array([b[1] for b in legacy.BOUNDS])
    pop = lo + rng.uniform(size=(5, 10)) * (hi - lo)
    for exp, tab in (("tl_clbr", TL), ("iso", ISO)):
        f = legacy.batched_objective(cfg, exp, pop, ctx=gpu_ctx, table=tab)
        assert f.shape == (5,) and np.all(np.isfinite(f)) and np.all((f >= 0) & (f <= 1))
        one = legacy.sim_starter(cfg, pop[0], exp, ctx=gpu_ctx, table=tab)
        assert abs(one - f[0]) < 0.03


def test_legacy_train_prepends_and_replays(gpu_ctx, tmp_path):
    cfg = config.compose(["exp_type_fp=TLlab", "physics_fp=lab_TL"])
    path = tmp_path / "result_tl_clbr.csv"
    print("legacy DE: 2 fits x 2 generations x 40 candidates", flush=True)
    res = legacy.train(cfg, "tl_clbr", runs=2, gens=2, pop=4, ctx=gpu_ctx, table=TL, seed=3,
                       polish=False, csv_path=path, disp=False)
    df = pd.read_csv(path)
    assert len(df) == 2 and list(df.columns) == [f"param_{i}" for i in range(10)] + ["mse"]
    assert df.mse.iloc[0] == pytest.approx(float(res[1].fun))    # the newest row first
    mse = legacy.replay(cfg, "tl_clbr", ctx=gpu_ctx, table=TL, csv_path=path, plot=False)
    assert np.isfinite(mse)
