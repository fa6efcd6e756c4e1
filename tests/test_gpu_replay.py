"""Bit-exact replay: the HIP kernel against the CPU replay harness (oracle/mcl_oracle.c).

Both consume the same Philox4x32-10 stream per (run, sim, electron, step) and the
same float32/float64 arithmetic, so per-step recombination counts must agree
exactly, cluster by cluster.
"""
import numpy as np
import pytest

from mcluminescence_amd import workloads as W
from mcluminescence_amd._abi import F_GRAPH, F_PERSISTENT, RunParams
from oracle import orc

pytestmark = pytest.mark.gpu


def _replay(ctx, runs, n_bins=None, seed=0x4D434C554D000001):
    print(f"replay: {len(runs)} runs, {sum(r.sims for r in runs)} clusters, seed {seed:#x}",
          flush=True)
    g_counts, g_live, stats = ctx.run_fixed_dt(runs, seed=seed, n_bins=n_bins)
    o_counts, o_live = orc.run_fixed_dt(runs, seed=seed, n_bins=n_bins, threads=8)
    return g_counts, g_live, o_counts, o_live, stats


def _assert_same(g_counts, g_live, o_counts, o_live):
    bad = np.nonzero((g_counts != o_counts).any(axis=1))[0]
    assert bad.size == 0, (
        f"{bad.size} clusters differ; first {bad[0]}: first step "
        f"{np.nonzero(g_counts[bad[0]] != o_counts[bad[0]])[0][:5]}")
    np.testing.assert_array_equal(g_live, o_live)


def test_initial_state_matches_oracle(gpu_ctx):
    run = W.tl_glow_16x500(1)[0]
    for run_index, sim in [(0, 0), (3, 7)]:
        g = gpu_ctx.cluster_state(run, run_index, sim)
        o = orc.init_state(run, run_index, sim)
        np.testing.assert_array_equal(g["e_xyz"], o["e_xyz"])
        np.testing.assert_array_equal(g["h_xyz"], o["h_xyz"])
        np.testing.assert_array_equal(g["nn"], o["nn"])
        np.testing.assert_array_equal(g["w"].view(np.uint32), o["w"].view(np.uint32))
        assert g["e_live"].all() and g["h_alive"].all()


def test_tl_glow_per_step_bit_exact(gpu_ctx):
    runs = W.tl_glow_16x500()
    g, gl, o, ol, stats = _replay(gpu_ctx, runs)
    _assert_same(g, gl, o, ol)
    assert g.sum() == 16 * 500
    assert stats.live_electron_steps == int(ol.sum())


def test_iso_decay_per_step_bit_exact(gpu_ctx):
    runs = W.iso_decay_4096x400(sims=256)
    g, gl, o, ol, _ = _replay(gpu_ctx, runs)
    _assert_same(g, gl, o, ol)


@pytest.mark.parametrize("fixed_thresholds,quiet_window", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_kernel_builds_bit_exact(gpu_ctx, fixed_thresholds, quiet_window):
    # each build of the fixed-dt kernels: with or without each electron's fire
    # threshold kept in its w slot (isothermal runs), with or without the quiet
    # window (ramped runs); the isothermal decay alone, a mixed launch (a ramped run
    # beside it) and the persistent graph of the mixed launch
    runs = W.iso_decay_4096x400(sims=128) + W.tl_glow_16x500(sims=16)
    gpu_ctx.tune_set("fixed_thresholds", fixed_thresholds)
    gpu_ctx.tune_set("quiet_window", quiet_window)
    try:
        print(f"fixed_thresholds {fixed_thresholds}, quiet_window {quiet_window}", flush=True)
        g, gl, o, ol, _ = _replay(gpu_ctx, runs[:1])
        _assert_same(g, gl, o, ol)
        g, gl, o, ol, _ = _replay(gpu_ctx, runs)
        _assert_same(g, gl, o, ol)
        gp, glp, _ = gpu_ctx.run_fixed_dt(runs, seed=0x4D434C554D000001,
                                          flags=F_PERSISTENT | F_GRAPH)
        _assert_same(gp, glp, o, ol)
    finally:
        gpu_ctx.tune_set("fixed_thresholds", 1)
        gpu_ctx.tune_set("quiet_window", 1)


def test_tunnelling_per_step_bit_exact(gpu_ctx):
    # reference TL12/basicTL12 physics: every electron-step takes the per-electron
    # threshold path (the band term is zero)
    runs = W.tunnelling_tl12(sims=4, n=600, T_rate=5.0)
    g, gl, o, ol, _ = _replay(gpu_ctx, runs)
    _assert_same(g, gl, o, ol)
    assert g.sum() > 0


def test_two_channel_and_binned_bit_exact(gpu_ctx):
    r = W.tunnelling_tl12(sims=3, n=300, T_rate=5.0)[0]
    r.E_loc_1, r.E_loc_2, r.Retrap = 1.15, 1.25, 0.3
    r2 = W.tl_glow_16x500(sims=5)[0]
    r2.n_steps = r.n_steps
    g, gl, o, ol, _ = _replay(gpu_ctx, [r, r2], n_bins=40)
    _assert_same(g, gl, o, ol)


def test_state_after_steps_matches_oracle_and_bruteforce(gpu_ctx):
    run = W.tunnelling_tl12(sims=1, n=500, T_rate=5.0)[0]
    for steps in (40, 100):
        g = gpu_ctx.cluster_state(run, 0, 0, max_steps=steps)
        o = orc.final_state(run, 0, 0, steps)
        np.testing.assert_array_equal(g["e_live"], o["e_live"])
        np.testing.assert_array_equal(g["h_alive"], o["h_alive"])
        live = g["e_live"].astype(bool)
        np.testing.assert_array_equal(g["nn"][live], o["nn"][live])
        bf = orc.bruteforce_nn(o["e_xyz"][live], o["h_xyz"], o["h_alive"])
        np.testing.assert_array_equal(g["nn"][live], bf)


@pytest.mark.parametrize("T_K", [540.0, 560.0, 700.0])
def test_conflict_heavy_iso_bit_exact(gpu_ctx, T_K):
    # p = 0.35, 0.71 and 1: many shared-hole conflicts and multi-round retries per step
    runs = W.iso_decay_4096x400(sims=64)
    runs[0].T_start = T_K - 273.15
    g, gl, o, ol, _ = _replay(gpu_ctx, runs)
    _assert_same(g, gl, o, ol)


def test_dense_tunnelling_retries_bit_exact(gpu_ctx):
    # high rho': short distances, retries decided by the T_new / T_old thinning draw
    r = W.tunnelling_tl12(sims=8, n=800, T_rate=5.0)[0]
    r.rho_prime = 3e-3
    r2 = W.tunnelling_tl12(sims=8, n=800, T_rate=5.0)[0]
    r2.rho_prime, r2.E_loc_1, r2.E_loc_2, r2.Retrap = 3e-3, 1.1, 1.3, 0.6
    g, gl, o, ol, _ = _replay(gpu_ctx, [r, r2])
    _assert_same(g, gl, o, ol)


def test_grid_edge_cases_bit_exact(gpu_ctx):
    runs = []
    # more electrons than holes: holes run out, electrons end with no partner
    r = W.single_trap(1.0, 1e16, 4e9, 1e23, 120, T_start_K=600.0, T_rate=0.0, dt=1.0,
                      n_steps=50, sims=3)
    r.N_e = 400.0
    r.e_ratio_start = 1.0
    runs.append(r)
    # holes confined to a smaller cube than the electrons (bf < 1): searches from
    # outside the hole grid fall back to the full scan
    r = W.tunnelling_tl12(sims=3, n=300, T_rate=5.0)[0]
    r.boundary_factor = 0.6
    r.n_steps = 50
    runs.append(r)
    # a handful of holes: one grid cell
    r = W.single_trap(1.3, 1e12, 4e9, 1e23, 6, T_start_K=300.0, T_rate=1.0, dt=0.1,
                      n_steps=50, sims=4)
    runs.append(r)
    # sparse, strongly tunnelling cluster at low density
    r = W.tunnelling_tl12(sims=3, n=500, T_rate=5.0)[0]
    r.rho_prime, r.n_steps = 3e-6, 50
    runs.append(r)
    g, gl, o, ol, _ = _replay(gpu_ctx, runs)
    _assert_same(g, gl, o, ol)
    # the nearest holes maintained under each of these geometries
    for run in runs:
        st = gpu_ctx.cluster_state(run, 0, 1, max_steps=7)
        ref = orc.final_state(run, 0, 1, 7)
        np.testing.assert_array_equal(st["e_live"], ref["e_live"])
        live = st["e_live"].astype(bool)
        np.testing.assert_array_equal(st["nn"][live], ref["nn"][live])


@pytest.mark.parametrize("T_K", [540.0, 560.0])
def test_conflict_heavy_iso_4096_closed_form(gpu_ctx, T_K):
    # the full decay workload where conflicts bind (p = 0.35 / 0.72): bit-exact
    # against the replay, and band-only survival N0 (1 - p)^n with mean life 1 / p
    from test_oracle_kinetics import _band_decay_check
    import math
    runs = W.iso_decay_4096x400()
    runs[0].T_start = T_K - 273.15
    g, gl, o, ol, _ = _replay(gpu_ctx, runs)
    _assert_same(g, gl, o, ol)
    p = -math.expm1(-1e13 * math.exp(-1.4 / (8.617343e-5 * T_K)) * 0.5)
    _band_decay_check(g, gl, 400 * 4096, p)
