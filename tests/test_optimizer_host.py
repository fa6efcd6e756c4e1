"""Host side of the optimizer drop-in: bounds, parameter mapping, CLI parsing."""
import numpy as np

from mcluminescence_amd import config, engine, optimizer


def test_bounds_and_parameter_order_follow_the_reference():
    # optimizer.py:32-43
    assert len(optimizer.DEFAULT_BOUNDS) == 10
    assert optimizer.DEFAULT_BOUNDS[0] == (1e-8, 1e-3) and optimizer.DEFAULT_BOUNDS[-1] == (0.0, 1.0)


def test_cfg_with_params_writes_every_field_and_copies():
    base = config.compose(["exp_type_fp=TLlab", "physics_fp=lab_TL"])
    p = [2e-4, 2.0, 1.3, 1.1, 250.0, 5e12, 3e11, 2e10, 300.0, 0.25]
    cfg = optimizer.cfg_with_params(base, p)
    e, ph = cfg.exp_type_fp, cfg.physics_fp
    assert (e.rho_prime, e.holes) == (2e-4, 300.0)
    assert (ph.E_cb, ph.E_loc_1, ph.E_loc_2, ph.D0, ph.s, ph.b, ph.alpha, ph.Retrap) == \
        (2.0, 1.3, 1.1, 250.0, 5e12, 3e11, 2e10, 0.25)
    assert base.physics_fp.E_cb == 1.8                      # the base is untouched
    run = engine.initialize_runs(cfg)[0]
    rp = engine.run_params(run, sims=1, n_steps=1)
    assert rp.holes == 300.0 and rp.Retrap == 0.25


def test_lab_run_construction_for_the_tl_table():
    import json
    from pathlib import Path
    from mcluminescence_amd import lab
    t = json.loads((Path(__file__).parent / "golden" / "reference_outputs.json").read_text())
    tab = t["lab_data"]["CLBR_IRSL50_0.25KperGy"]
    run = engine.initialize_runs(config.compose(["exp_type_fp=TLlab", "physics_fp=lab_TL"]))[0]
    runs, obs = lab.tl_lab_runs(run, tab, sims=3)
    assert len(runs) == 11 and all(r.sims == 3 for r in runs)
    assert runs[0].D == 0.092 and runs[0].T_start == 300.0
    np.testing.assert_allclose(runs[0].T_rate, (275 - 300) / 1086.0)
    assert obs[3] == [4343.0]
