"""Clusters beyond one workgroup's LDS: the global-memory form of the fixed-dt and
event-mode kernels.

A cluster whose holes, cell starts and electron arrays need more LDS than a
workgroup has keeps its holes and cell starts in a global scratch slice (and its
electron arrays too once they no longer fit LDS).  The reference has no size
limit (a dense electron x hole distance matrix, engine.py:113-119); here the
limit is the 16-bit electron and hole indices (65535 / 65534).  Every result is
compared bit for bit with the CPU replay harness, which has no size limit
either.
"""
import numpy as np
import pytest

from mcluminescence_amd import workloads as W
from mcluminescence_amd._abi import F_GRAPH, F_PERSISTENT, MCL_ERANGE, MclError, RunParams
from oracle import orc

pytestmark = pytest.mark.gpu

SEED = 0x4D434C554D000001


def _tl(n, sims, n_steps=4000):
    # glow curve 300 -> 700 K at 1 K/s: every electron recombines by the end
    return W.single_trap(1.3, 1e12, 4e9, 1e23, n, T_start_K=300.0, T_rate=1.0, dt=0.1,
                         n_steps=n_steps, sims=sims)


def _iso(n, sims, n_steps, T_K=540.0):
    # p = 0.35 per step: many shared-hole conflicts and retry rounds
    return W.single_trap(1.4, 1e13, 4e9, 1e23, n, T_start_K=T_K, T_rate=0.0, dt=0.5,
                         n_steps=n_steps, sims=sims)


def _check(ctx, runs, n_bins=None, flags=0):
    print(f"large clusters: {[(orc.geometry(r).n_e0, orc.geometry(r).n_h, r.sims) for r in runs]}"
          f" n_bins={n_bins} flags={flags:#x}", flush=True)
    g, gl, stats = ctx.run_fixed_dt(runs, seed=SEED, n_bins=n_bins, flags=flags)
    o, ol = orc.run_fixed_dt(runs, seed=SEED, n_bins=n_bins, threads=8)
    bad = np.nonzero((g != o).any(axis=1))[0]
    assert bad.size == 0, (f"{bad.size} clusters differ; first {bad[0]}: steps "
                           f"{np.nonzero(g[bad[0]] != o[bad[0]])[0][:5]}")
    np.testing.assert_array_equal(gl, ol)
    assert stats.live_electron_steps == int(ol.sum())
    return g


def _lds_fits(ctx, run):
    e, h = ctx.cluster_limits()
    geo = orc.geometry(run)
    return geo.n_e0 <= e and geo.n_h <= h


def test_global_holes_electrons_in_lds(gpu_ctx):
    # 4000 electrons / 6911 holes: 189 KB in LDS form, so holes and cell starts go
    # to global memory and the 64 KB of electron arrays stay in LDS
    run = _tl(4000, 8)
    assert not _lds_fits(gpu_ctx, run)
    g = _check(gpu_ctx, [run])
    assert g.sum() == 8 * 4000


def test_everything_global_conflict_heavy(gpu_ctx):
    # 12000 electrons / 20736 holes: the electron arrays (192 KB) leave LDS too
    _check(gpu_ctx, [_iso(12000, 3, 400)])


def test_largest_clusters(gpu_ctx):
    # 32768 electrons (8 waves of 64 rows of 64) / 56623 holes
    run = _iso(32768, 2, 150)
    assert orc.geometry(run).n_e0 == 32768
    _check(gpu_ctx, [run], n_bins=50)


def test_tunnelling_large_bit_exact(gpu_ctx):
    # per-electron tunnelling thresholds, two channels, binned output
    r = W.tunnelling_tl12(sims=2, n=10000, T_rate=5.0)[0]
    r.E_loc_1, r.E_loc_2, r.Retrap = 1.15, 1.25, 0.3
    _check(gpu_ctx, [r], n_bins=100)


@pytest.mark.parametrize("waves", [1, 2, 4, 8])
def test_every_wave_count(gpu_ctx, waves):
    # 4000 electrons: 63 rows, within one wave's 64
    gpu_ctx.tune_set("waves_per_cluster", waves)
    try:
        _check(gpu_ctx, [_iso(4000, 2, 300)])
    finally:
        gpu_ctx.tune_set("waves_per_cluster", 0)


def test_knob_below_the_rows_a_wave_takes(gpu_ctx):
    # one wave asked for 9000 electrons (141 rows): the launch takes four waves
    gpu_ctx.tune_set("waves_per_cluster", 1)
    try:
        _check(gpu_ctx, [_iso(9000, 2, 200)])
    finally:
        gpu_ctx.tune_set("waves_per_cluster", 0)


@pytest.mark.parametrize("flags", [F_PERSISTENT, F_PERSISTENT | F_GRAPH])
def test_mixed_sizes_persistent(gpu_ctx, flags):
    # size classes side by side: LDS classes for the small clusters, the
    # global-memory form for the large ones (scratch slices per class)
    runs = [_tl(300, 64, 1000), _tl(900, 16, 1000), _tl(4500, 4, 1000), _iso(11000, 2, 200)]
    _check(gpu_ctx, runs, n_bins=100, flags=flags)


def test_plan_relaunch_and_info(gpu_ctx):
    import torch
    run = _iso(6000, 5, 200)
    plan = gpu_ctx.plan([run], seed=SEED, n_bins=20)
    try:
        info = plan.info
        e_max, _ = gpu_ctx.cluster_limits()
        assert info.max_electrons == 6000 > e_max
        assert 1 <= info.grid <= 5
        counts = torch.zeros((5, info.row_len), dtype=torch.int32, device="cuda")
        live = torch.zeros(5, dtype=torch.int64, device="cuda")
        torch.cuda.synchronize()
        outs = []
        for _ in range(2):   # the scratch slices and the ticket are reused
            plan.launch(counts.data_ptr(), live.data_ptr())
            plan.sync()
            outs.append((counts.cpu().numpy().view(np.uint32).copy(),
                         live.cpu().numpy().view(np.uint64).copy()))
    finally:
        plan.close()
    o, ol = orc.run_fixed_dt([run], seed=SEED, n_bins=20, threads=8)
    for c, lv in outs:
        np.testing.assert_array_equal(c, o)
        np.testing.assert_array_equal(lv, ol)


def test_single_cluster_and_state(gpu_ctx):
    run = _iso(4000, 3, 50)
    row, live = gpu_ctx.run_cluster(run, 0, 2)
    o, ol = orc.run_fixed_dt([run], seed=SEED, threads=8)
    np.testing.assert_array_equal(row, o[2])
    assert live == ol[2]
    # initial nearest holes and w, then the state after 5 steps (488 electrons left)
    g = gpu_ctx.cluster_state(run, 0, 1)
    ref = orc.init_state(run, 0, 1)
    np.testing.assert_array_equal(g["e_xyz"], ref["e_xyz"])
    np.testing.assert_array_equal(g["h_xyz"], ref["h_xyz"])
    np.testing.assert_array_equal(g["nn"], ref["nn"])
    np.testing.assert_array_equal(g["w"].view(np.uint32), ref["w"].view(np.uint32))
    g = gpu_ctx.cluster_state(run, 0, 1, max_steps=5)
    ref = orc.final_state(run, 0, 1, 5)
    np.testing.assert_array_equal(g["e_live"], ref["e_live"])
    np.testing.assert_array_equal(g["h_alive"], ref["h_alive"])
    lv = g["e_live"].astype(bool)
    assert 0 < lv.sum() < 4000
    np.testing.assert_array_equal(g["nn"][lv], ref["nn"][lv])


def test_beyond_the_limits_refused(gpu_ctx):
    # 69120 holes: beyond the 16-bit hole index
    with pytest.raises(MclError) as e:
        gpu_ctx.run_fixed_dt([_iso(40000, 1, 10)], seed=SEED)
    assert e.value.code == MCL_ERANGE
    # 33000 electrons among 20000 core holes: more rows than 8 waves take
    r = _iso(20000, 1, 10)
    r.N_e = 33000.0
    with pytest.raises(MclError) as e:
        gpu_ctx.run_fixed_dt([r], seed=SEED)
    assert e.value.code == MCL_ERANGE
    assert "32768" in str(e.value)


@pytest.fixture
def global_form(gpu_ctx):
    gpu_ctx.tune_set("global_memory", 1)
    try:
        yield gpu_ctx
    finally:
        gpu_ctx.tune_set("global_memory", 0)


def test_global_form_on_clusters_that_fit_lds(global_form):
    # the same kernels with the holes in global memory on the decay, glow,
    # conflict-heavy and two-channel workloads the LDS form is tested on
    iso = W.iso_decay_4096x400(sims=256)[0]
    hot = W.iso_decay_4096x400(sims=64)[0]
    hot.T_start = 560.0 - 273.15
    two = W.tunnelling_tl12(sims=3, n=300, T_rate=5.0)[0]
    two.E_loc_1, two.E_loc_2, two.Retrap = 1.15, 1.25, 0.3
    _check(global_form, [iso])
    _check(global_form, W.tl_glow_16x500())
    _check(global_form, [hot])
    _check(global_form, [two], n_bins=40)


def test_global_form_uneven_sweep_persistent(global_form):
    # 240 sets of 64-2048 electrons: five size classes, every one in global form
    runs = W.uneven_sweep(n_sets=240, n_steps=192, seed=11)
    _check(global_form, runs, n_bins=48, flags=F_PERSISTENT | F_GRAPH)


def _event_run(n, sims):
    # the glow run in event mode: one event per step, at most 0.1 s a step
    p = RunParams.make(**{**_tl(n, 1, 1).as_dict(), "sims": sims, "n_steps": 1})
    p.dt = 0.1
    return p


def _check_event(ctx, run, duration, max_steps):
    print(f"event mode, large clusters: {orc.geometry(run).n_e0} electrons x {run.sims}",
          flush=True)
    lum, t, er, steps, stats = ctx.run_event([run], [duration], max_steps)
    rl, rt, rer, rsteps = orc.run_event([run], [duration], max_steps, threads=8)
    np.testing.assert_array_equal(steps, rsteps)
    np.testing.assert_array_equal(lum, rl)
    np.testing.assert_array_equal(t, rt)
    np.testing.assert_array_equal(er, rer)
    assert stats.recombinations == int(lum.sum())


@pytest.mark.parametrize("n,sims", [(5000, 3), (16000, 2)])
def test_event_mode_beyond_lds(gpu_ctx, n, sims):
    # 5000: holes in global memory, electrons in LDS; 16000: both in global memory
    run = _event_run(n, sims)
    assert not _lds_fits(gpu_ctx, run)
    _check_event(gpu_ctx, run, 400.0, 4000 + n + 200)


def test_event_mode_global_form_small(global_form):
    _check_event(global_form, _event_run(500, 4), 400.0, 5000)
    iso = RunParams.make(**{**W.iso_decay_4096x400(sims=1)[0].as_dict(), "sims": 16,
                            "n_steps": 1})
    _check_event(global_form, iso, 2000.0, 500)
