"""The optimizer drop-in on the GPU: a short DE run on TL_lab with the population
batched into one lab-kernel launch, its result file, and replay."""
import json
from pathlib import Path

import numpy as np
import pandas as pd
import pytest

from mcluminescence_amd import config, optimizer

pytestmark = pytest.mark.gpu
LAB = json.loads((Path(__file__).parent / "golden" / "reference_outputs.json").read_text())["lab_data"]


def test_batched_objective_matches_per_candidate_scale(gpu_ctx):
    cfg = config.compose(["exp_type_fp=TLlab", "physics_fp=lab_TL"])
    rng = np.random.default_rng(3)
    lo = np.array([b[0] for b in optimizer.DEFAULT_BOUNDS])
    hi = np.array([b[1] for b in optimizer.DEFAULT_BOUNDS])
    pop = lo + rng.uniform(size=(6, 10)) * (hi - lo)
    f = optimizer.batched_objective(cfg, "tl_clbr", pop, ctx=gpu_ctx, sims=64,
                                    table=LAB["CLBR_IRSL50_0.25KperGy"])
    assert f.shape == (6,) and np.all(np.isfinite(f)) and np.all((f >= 0) & (f <= 1))
    # one candidate alone through TLTrapSim.TL_lab gives the same objective up to noise
    one = optimizer.run_one_sim(optimizer.cfg_with_params(cfg, pop[0]), "tl_clbr", ctx=gpu_ctx,
                                sims=64, table=LAB["CLBR_IRSL50_0.25KperGy"])
    assert abs(one - f[0]) < 0.02


def test_train_and_replay(gpu_ctx, tmp_path, monkeypatch):
    monkeypatch.setenv("MCL_PROJECT_ROOT", str(tmp_path))      # the re-run's plot goes here
    cfg = config.compose(["exp_type_fp=TLlab", "physics_fp=lab_TL"])
    print("DE on TL_lab: 2 generations x 40 candidates, batched", flush=True)
    res = optimizer.train(cfg, "tl_clbr", gens=2, pop=4, sims=4, ctx=gpu_ctx,
                          table=LAB["CLBR_IRSL50_0.25KperGy"], seed=1, polish=False,
                          csv_path=tmp_path / "result_tl_clbr.csv", disp=False)
    assert np.isfinite(res.fun) and res.x.shape == (10,)
    # the best parameters were run again with the plot written (optimizer.py:119-121)
    assert np.isfinite(res.best_rerun_mse)
    assert (tmp_path / "results" / "lab_sims" / "CLBR_IRSL50_0.25KperGy.png").exists()
    for (lo, hi), v in zip(optimizer.DEFAULT_BOUNDS, res.x):
        assert lo <= v <= hi
    df = pd.read_csv(tmp_path / "result_tl_clbr.csv")
    assert list(df.columns) == [f"param_{i}" for i in range(10)] + ["mse"] and len(df) == 1
    mse = optimizer.replay(cfg, "tl_clbr", ctx=gpu_ctx, table=LAB["CLBR_IRSL50_0.25KperGy"],
                           csv_path=tmp_path / "result_tl_clbr.csv", plot=False)
    assert np.isfinite(mse)
