"""simulate() and TLTrapSim through the HIP kernels: the reference's driver contract
(simulate.py:11-112) with per-cluster counts identical to the replay harness."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from mcluminescence_amd import config, engine
from mcluminescence_amd._abi import RunParams
from mcluminescence_amd.simulate import simulate
from mcluminescence_amd.tl_trap_lab import TLTrapSim
from oracle import orc

pytestmark = pytest.mark.gpu
GOLDEN = json.loads((Path(__file__).parent / "golden" / "reference_outputs.json").read_text())


def test_simulate_tl12_contract_and_csv(gpu_ctx, tmp_path):
    cfg = config.compose(["+tag=gputest"])
    print("simulate TL12: 4 runs x 2 sims x 20000 steps", flush=True)
    x_ax, lum, er, configs = simulate(cfg, ctx=gpu_ctx, results_dir=tmp_path)
    assert lum.shape == x_ax.shape == er.shape == (20000, 2, 4) and len(configs) == 4
    # 800 steps of max_dt / T_rate each, zero-padded afterwards
    for r, dt in enumerate([10.0, 1.0, 0.2, 0.05]):
        np.testing.assert_allclose(x_ax[:800, 0, r], dt * np.arange(1, 801))
        assert (x_ax[800:, :, r] == 0).all() and (lum[800:, :, r] == 0).all()
        assert (er[800:, :, r] == 0).all()
        left = 2000 - np.cumsum(lum[:800, :, r], axis=0)
        np.testing.assert_allclose(er[:800, :, r], left / 2000)
    # per-run totals against the reference's committed results (loose: 2 sims)
    for r in range(4):
        ref = [GOLDEN["exp_"][f"{r},{s}"]["recombinations"] for s in (0, 1)]
        got = lum[:, :, r].sum(axis=0)
        assert np.all(np.abs(got - np.mean(ref)) < 120), (r, got, ref)
    lines = (tmp_path / "exp_gputest.csv").read_text().splitlines()
    assert lines[0] == "run,sim,step,lum,electron_ratio" and len(lines) == 160001
    assert lines[1] == "0,0,0," + repr(float(lum[0, 0, 0])) + "," + repr(float(er[0, 0, 0]))


def test_simulate_counts_equal_replay_harness(gpu_ctx, tmp_path):
    cfg = config.compose(["exp_type_fp.e_ratio_start=0.05"])
    _, lum, _, configs = simulate(cfg, ctx=gpu_ctx, write_csv=False)
    params = [engine.run_params(configs[r], sims=2) for r in range(4)]
    ref, _ = orc.run_fixed_dt(params, n_bins=0, threads=8)
    for r in range(4):
        np.testing.assert_array_equal(lum[:800, :, r].T, ref[2 * r:2 * r + 2, :800])


def test_step_overflow_raises_index_error(gpu_ctx):
    cfg = config.compose(["exp_type_fp.steps=100"])
    with pytest.raises(IndexError):
        simulate(cfg, ctx=gpu_ctx, write_csv=False)


def test_tltrapsim_box_view(gpu_ctx):
    run = engine.initialize_runs(config.compose())[0]
    sim = TLTrapSim(run, e_ratio_start=0.25, ctx=gpu_ctx)
    assert sim.n_e0 == 500 and sim.n_h == 3455
    e, h = sim.box.electrons, sim.box.holes
    assert e.shape == (500, 3) and h.shape == (3455, 3)
    assert (e >= 0).all() and (e < sim.d * (1 + 1e-6)).all()
    assert (h >= 0).all() and (h < sim.d * 1.2 * (1 + 1e-6)).all()
    d, nn = sim.box.nearest()
    full = np.linalg.norm(e[:, None, :] - h[None, :, :], axis=2)
    np.testing.assert_allclose(d, full.min(axis=1), rtol=1e-6)
    assert (full[np.arange(500), nn] <= full.min(axis=1) * (1 + 1e-6)).all()
    counts = sim.run()
    assert counts.shape == (800,) and 0 < counts.sum() <= 500
    # after 200 steps the view shows fewer electrons and holes, same count removed
    sim.box.advance_to(200)
    gone = int(counts[:200].sum())
    assert sim.box.electrons.shape[0] == 500 - gone
    assert sim.box.holes.shape[0] == 3455 - gone


def test_tltrapsim_run_is_the_sweeps_row(gpu_ctx):
    # cluster (run_index 2, sim 5) alone gives exactly its row of the full sweep
    run = engine.initialize_runs(config.compose())[0]
    sim = TLTrapSim(run, e_ratio_start=0.25, ctx=gpu_ctx, run_index=2, sim=5)
    counts = sim.run(n_steps=300)
    p = RunParams.make(**{**sim._params.as_dict(), "sims": 6, "n_steps": 300})
    full, _, _ = gpu_ctx.run_fixed_dt([p, p, p], n_bins=0)
    np.testing.assert_array_equal(counts, full[2 * 6 + 5])
    assert counts.sum() > 0
