"""Batched uneven sweeps: the persistent size-class launch and its hipGraph replay give
exactly the counts of the one-workgroup-per-cluster launch and of the replay harness."""
import numpy as np
import pytest

from mcluminescence_amd import autotune
from mcluminescence_amd import workloads as W
from mcluminescence_amd._abi import F_GRAPH, F_PERSISTENT, Context, RunParams
from oracle import orc

pytestmark = pytest.mark.gpu


def _sweep():
    return W.uneven_sweep(n_sets=240, n_steps=192, seed=11)


def test_persistent_equals_plain_and_oracle(gpu_ctx):
    runs = _sweep()
    print("uneven sweep: 240 sets, 64-2048 electrons, 192 steps, 48 bins", flush=True)
    plain, lp, _ = gpu_ctx.run_fixed_dt(runs, n_bins=48)
    pers, lq, _ = gpu_ctx.run_fixed_dt(runs, n_bins=48, flags=F_PERSISTENT)
    np.testing.assert_array_equal(plain, pers)
    np.testing.assert_array_equal(lp, lq)
    ref, lr = orc.run_fixed_dt(runs, n_bins=48, threads=8)
    np.testing.assert_array_equal(plain, ref)
    np.testing.assert_array_equal(lp, lr)


def test_graph_replays_match_eager(gpu_ctx):
    import torch
    runs = _sweep()
    eager, _, _ = gpu_ctx.run_fixed_dt(runs, n_bins=48)
    plan = gpu_ctx.plan(runs, n_bins=48, flags=F_PERSISTENT | F_GRAPH)
    n, row = plan.info.total_clusters, plan.info.row_len
    counts = torch.zeros((n, row), dtype=torch.int32, device="cuda")
    live = torch.zeros(n, dtype=torch.int64, device="cuda")
    for rep in range(4):
        counts.fill_(-1)                       # eager work between replays
        torch.cuda.synchronize()
        plan.launch(counts.data_ptr(), live.data_ptr())
        plan.sync()
        got = counts.cpu().numpy().view(np.uint32)
        np.testing.assert_array_equal(got, eager, err_msg=f"graph replay {rep}")
        assert plan.work().recombinations == int(eager.sum())
    plan.close()


def test_tuning_knobs_do_not_change_results(gpu_ctx):
    runs = _sweep()[:80] + W.tl_glow_16x500(sims=4)
    base, lb, _ = gpu_ctx.run_fixed_dt(runs, n_bins=48, flags=F_PERSISTENT)
    try:
        for hpc, pf, wsm, la in [(1.0, 0.5, 0, 0), (4.0, 2.0, 64, 64), (2.0, 1.0, 2, 65535)]:
            gpu_ctx.tune_set("holes_per_cell", hpc)
            gpu_ctx.tune_set("persist_frac", pf)
            gpu_ctx.tune_set("wave_search_max", wsm)
            gpu_ctx.tune_set("lookahead_max", la)
            got, lg, _ = gpu_ctx.run_fixed_dt(runs, n_bins=48, flags=F_PERSISTENT)
            np.testing.assert_array_equal(got, base)
            np.testing.assert_array_equal(lg, lb)
    finally:
        for k, v in autotune.DEFAULTS.items():
            gpu_ctx.tune_set(k, v)


def test_autotuner_picks_a_measured_point(gpu_ctx, tmp_path):
    runs = _sweep()[:60]
    entry = autotune.autotune(gpu_ctx, runs, n_bins=48,
                              space={"holes_per_cell": [2.0, 3.0], "persist_frac": [1.0]},
                              reps=1, table_path=tmp_path / "t.json", log=None)
    assert entry["holes_per_cell"] in (2.0, 3.0) and entry["ms"] > 0
    table = autotune.load_table(tmp_path / "t.json")
    assert autotune.signature(runs) in table
    knobs = autotune.apply(gpu_ctx, runs, table)
    assert knobs["holes_per_cell"] == entry["holes_per_cell"]
    gpu_ctx.tune_set("holes_per_cell", 2.0)


def test_build_choice_knobs_set_get_and_reject(gpu_ctx):
    from mcluminescence_amd._abi import MclError
    for key in ("fixed_thresholds", "quiet_window", "global_memory"):
        default = gpu_ctx.tune_get(key)
        try:
            for v in (0.0, 1.0):
                gpu_ctx.tune_set(key, v)
                assert gpu_ctx.tune_get(key) == v
            for bad in (2.0, -1.0, 0.5):
                with pytest.raises(MclError):
                    gpu_ctx.tune_set(key, bad)
        finally:
            gpu_ctx.tune_set(key, default)
    assert gpu_ctx.tune_get("fixed_thresholds") == 1.0
    assert gpu_ctx.tune_get("quiet_window") == 1.0
    assert gpu_ctx.tune_get("global_memory") == 0.0


def test_reused_device_buffers_give_fresh_context_results(gpu_ctx):
    """Runs take their device buffers from the context's pool: a run that reuses the
    blocks of larger runs (fixed-dt, persistent and event mode) gives exactly what a
    fresh context gives."""
    print("device block reuse: sweep 240 sets, then 80 sets, then an event run", flush=True)
    big, small = _sweep(), _sweep()[:80]
    ev = RunParams.make(**{**W.tunnelling_tl12(sims=1, n=200)[0].as_dict(), "sims": 3, "n_steps": 1})
    ev.dt = 1.0
    gpu_ctx.run_fixed_dt(big, n_bins=48)                       # leaves its blocks in the pool
    gpu_ctx.run_fixed_dt(big, n_bins=48, flags=F_PERSISTENT)
    got = [gpu_ctx.run_fixed_dt(small, n_bins=48)[:2],
           gpu_ctx.run_fixed_dt(small, n_bins=48, flags=F_PERSISTENT)[:2],
           gpu_ctx.run_event([ev], [800.0], 1200)[:4]]
    fresh = Context(0)
    try:
        want = [fresh.run_fixed_dt(small, n_bins=48)[:2],
                fresh.run_fixed_dt(small, n_bins=48, flags=F_PERSISTENT)[:2],
                fresh.run_event([ev], [800.0], 1200)[:4]]
    finally:
        fresh.close()
    for g, w in zip(got, want):
        for a, b in zip(g, w):
            np.testing.assert_array_equal(a, b)


def test_full_uneven_sweep_graph_launch_matches_replay(gpu_ctx):
    """The bench's sweep at its quoted size: 10,000 sets, 64-2048 electrons, 1024
    steps, 256 bins, one persistent graph launch with the tuning table's knobs,
    counts and live electron-steps identical to the replay harness."""
    import torch
    runs = W.uneven_sweep()
    knobs = autotune.apply(gpu_ctx, runs)
    print(f"uneven sweep, full size: {len(runs)} sets, 1024 steps, 256 bins, knobs {knobs}",
          flush=True)
    try:
        plan = gpu_ctx.plan(runs, n_bins=256, flags=F_PERSISTENT | F_GRAPH)
        n, row = plan.info.total_clusters, plan.info.row_len
        counts = torch.zeros((n, row), dtype=torch.int32, device="cuda")
        live = torch.zeros(n, dtype=torch.int64, device="cuda")
        plan.launch(counts.data_ptr(), live.data_ptr())
        plan.sync()
        got = counts.cpu().numpy().view(np.uint32)
        got_live = live.cpu().numpy().view(np.uint64)
        plan.close()
    finally:
        autotune.apply(gpu_ctx, [], {})
    print("replay harness on the same 10,000 sets", flush=True)
    ref, ref_live = orc.run_fixed_dt(runs, n_bins=256, threads=16)
    bad = np.nonzero((got != ref).any(axis=1))[0]
    assert bad.size == 0, f"{bad.size} sets differ, first {bad[:5]}"
    np.testing.assert_array_equal(got_live, ref_live)
    assert int(ref.sum()) > 0
