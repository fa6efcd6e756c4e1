"""Several wavefronts per cluster: the 2-, 4- and 8-wave kernels give exactly the counts
and live electron-steps of the one-wave kernel and of the CPU replay harness, on
one-workgroup-per-cluster, persistent and graph launches, and at the extremes of
the other tuning knobs.

The waves split a cluster's 64-slot rows and meet at workgroup barriers; every
per-electron decision (draw, threshold, lowest-index conflict rule, stable
compaction order) is the one-wave kernel's, so any difference is a bug.
"""
import numpy as np
import pytest

from mcluminescence_amd import autotune
from mcluminescence_amd import workloads as W
from mcluminescence_amd._abi import F_GRAPH, F_PERSISTENT, MclError
from oracle import orc

pytestmark = pytest.mark.gpu


def _run(ctx, waves, runs, **kw):
    ctx.tune_set("waves_per_cluster", waves)
    try:
        return ctx.run_fixed_dt(runs, **kw)
    finally:
        ctx.tune_set("waves_per_cluster", 0)


def _cases():
    two = W.tunnelling_tl12(sims=3, n=700, T_rate=5.0)[0]
    two.E_loc_1, two.E_loc_2, two.Retrap = 1.15, 1.25, 0.3      # two channels
    dense = W.tunnelling_tl12(sims=4, n=1500, T_rate=5.0)[0]
    dense.rho_prime, dense.n_steps = 3e-3, 400                   # 24 rows, many retries
    hot = W.iso_decay_4096x400(sims=16)[0]
    hot.T_start = 700.0 - 273.15                                 # p = 1: every hole contested
    small = W.single_trap(1.3, 1e12, 4e9, 1e23, 40, T_start_K=300.0, T_rate=1.0, dt=0.1,
                          n_steps=300, sims=3)                   # one row: other waves idle
    return {
        "tl_glow_16x500": (W.tl_glow_16x500(), None),
        "two_channel_700": ([two], 64),
        "dense_1500": ([dense], None),
        "iso_all_conflicts": ([hot], None),
        "small_and_uneven": ([small] + W.uneven_sweep(n_sets=24, n_steps=160, seed=5), 40),
    }


@pytest.mark.parametrize("name", list(_cases()))
def test_waves_give_identical_counts(gpu_ctx, name):
    runs, n_bins = _cases()[name]
    print(f"multiwave {name}: {sum(r.sims for r in runs)} clusters", flush=True)
    base, lb, _ = _run(gpu_ctx, 1, runs, n_bins=n_bins)
    for w in (2, 4, 8):
        print(f"  {w} waves", flush=True)
        got, lg, _ = _run(gpu_ctx, w, runs, n_bins=n_bins)
        np.testing.assert_array_equal(got, base, err_msg=f"{name}: {w} waves")
        np.testing.assert_array_equal(lg, lb, err_msg=f"{name}: {w} waves, live steps")


def test_four_waves_match_replay_harness(gpu_ctx):
    runs = W.tunnelling_tl12(sims=4, n=1000, T_rate=5.0)
    runs[0].n_steps = 600
    print("multiwave vs harness: TL12 4 x 1000, 600 steps", flush=True)
    g, gl, _ = _run(gpu_ctx, 4, runs)
    o, ol = orc.run_fixed_dt(runs, threads=8)
    np.testing.assert_array_equal(g, o)
    np.testing.assert_array_equal(gl, ol)
    assert g.sum() > 0


def test_knob_extremes_with_waves(gpu_ctx):
    runs = W.uneven_sweep(n_sets=40, n_steps=192, seed=7)
    print("multiwave knob extremes: 40 uneven sets", flush=True)
    base, lb, _ = _run(gpu_ctx, 1, runs, n_bins=48)
    try:
        for la, wsm in [(0.0, 0.0), (65535.0, 64.0), (8.0, 2.0)]:
            gpu_ctx.tune_set("lookahead_max", la)
            gpu_ctx.tune_set("wave_search_max", wsm)
            for w in (2, 4, 8):
                print(f"  lookahead_max {la:g} wave_search_max {wsm:g}: {w} waves", flush=True)
                got, lg, _ = _run(gpu_ctx, w, runs, n_bins=48)
                np.testing.assert_array_equal(got, base)
                np.testing.assert_array_equal(lg, lb)
    finally:
        for k, v in autotune.DEFAULTS.items():
            gpu_ctx.tune_set(k, v)


def test_persistent_and_graph_launches_with_waves(gpu_ctx):
    import torch
    runs = W.uneven_sweep(n_sets=240, n_steps=192, seed=11)
    print("multiwave persistent/graph: 240 uneven sets", flush=True)
    base, lb, _ = _run(gpu_ctx, 1, runs, n_bins=48)
    for w in (0, 2, 4, 8):                   # 0: each size class picks its own
        print(f"  persistent, waves_per_cluster {w}", flush=True)
        got, lg, _ = _run(gpu_ctx, w, runs, n_bins=48, flags=F_PERSISTENT)
        np.testing.assert_array_equal(got, base)
        np.testing.assert_array_equal(lg, lb)
    plan = gpu_ctx.plan(runs, n_bins=48, flags=F_PERSISTENT | F_GRAPH)
    n, row = plan.info.total_clusters, plan.info.row_len
    counts = torch.zeros((n, row), dtype=torch.int32, device="cuda")
    live = torch.zeros(n, dtype=torch.int64, device="cuda")
    for rep in range(3):
        counts.fill_(-1)
        torch.cuda.synchronize()
        plan.launch(counts.data_ptr(), live.data_ptr())
        plan.sync()
        np.testing.assert_array_equal(counts.cpu().numpy().view(np.uint32), base,
                                      err_msg=f"graph replay {rep}")
        np.testing.assert_array_equal(live.cpu().numpy().view(np.uint64), lb)
    plan.close()


def test_waves_knob_is_validated(gpu_ctx):
    for bad in (3.0, -1.0, 16.0):
        with pytest.raises(MclError):
            gpu_ctx.tune_set("waves_per_cluster", bad)
    assert gpu_ctx.tune_get("waves_per_cluster") == 0.0
