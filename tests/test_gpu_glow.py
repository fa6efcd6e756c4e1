"""GPU glow curves (mcl_glow_curve) against the numpy restatement of the
reference's plots.py:39-66 (oracle/glow.py).

Tolerance: the kernel sums counts exactly in 64-bit integers and divides once
per point; np.histogram sums float weights through cumulative sums and
np.convolve multiplies by a rounded 1/k, so the two agree to rounding:
|gpu - oracle| <= 1e-12 * max|oracle| per point, same lengths.
"""
import numpy as np
import pytest

from mcluminescence_amd import config, plots, workloads
from mcluminescence_amd._abi import MCL_ERANGE, MclError, RunParams
from mcluminescence_amd.simulate import run_sweep_event, simulate
from oracle import glow

pytestmark = pytest.mark.gpu
RTOL = 1e-12


def assert_curves(got, want):
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert g.shape == w.shape, (g.shape, w.shape)
        scale = max(float(np.abs(w).max()) if w.size else 0.0, 1e-300)
        np.testing.assert_allclose(g, w, rtol=0, atol=RTOL * scale)


def ramp(T_start, T_rate, sims, n_steps=1, dt=1.0):
    return RunParams.make(alpha=4e9, b=1e12, s=1e12, E_cb=1.3, E_loc_1=1.3, E_loc_2=1.3,
                          D0=400.35, Retrap=0.5, k_b=8.617343e-5, D=0.0, rho_prime=1e-4,
                          holes=100.0, boundary_factor=1.0, N_e=100.0, e_ratio_start=1.0,
                          T_start=T_start, T_rate=T_rate, dt=dt, n_steps=n_steps, sims=sims)


def oracle_from_rows(runs, counts, t_axis, bw=1.0, win=50.0):
    """[clusters, steps] rows -> the [steps, sims, runs] arrays of the oracle, per run."""
    out, c = [], 0
    for r in runs:
        s = int(r.sims)
        lum = counts[c:c + s].T.astype(np.float64)[:, :, None]
        x = t_axis[c:c + s].T[:, :, None]
        out += glow.tl_curves(x, lum, [(r.T_start, r.T_rate)], bw, win)
        c += s
    return out


def test_tl12_sweep_matches_oracle_both_axes(gpu_ctx):
    cfg = config.compose(["exp_type_fp.e_ratio_start=0.25"])
    print("glow: TL12 fixed-dt 4 runs x 2 sims", flush=True)
    x_ax, lum, er, configs = simulate(cfg, ctx=gpu_ctx, write_csv=False)
    got = plots.glow_curves(x_ax, lum, configs, ctx=gpu_ctx)
    want = glow.tl_curves(x_ax, lum, [(float(configs[r].exp_type_fp.T_start),
                                       float(configs[r].exp_type_fp.T_rate)) for r in range(4)])
    assert_curves(got, want)
    assert all(len(c) > 700 for c in got)
    # the fixed-dt axis computed in the kernel equals the explicit x_ax rows bit for bit
    from mcluminescence_amd.engine import run_params
    runs = [run_params(configs[r], sims=2) for r in range(4)]
    counts = lum.transpose(2, 1, 0).reshape(8, -1).astype(np.uint32)
    closed, _ = gpu_ctx.glow_curve(runs, counts, None)
    for a, b in zip(closed, got):
        np.testing.assert_array_equal(a, b)


def test_baseline_glow_16x500_matches_oracle(gpu_ctx):
    runs = workloads.tl_glow_16x500()
    print("glow: 16 x 500 TL glow, per-step counts", flush=True)
    counts, _, _ = gpu_ctx.run_fixed_dt(runs, n_bins=0)
    r = runs[0]
    n = int(r.n_steps)
    t = np.broadcast_to(np.arange(1, n + 1, dtype=np.float64) * r.dt, counts.shape)
    got, stats = gpu_ctx.glow_curve(runs, counts, None)
    assert_curves(got, oracle_from_rows(runs, counts, t))
    assert stats.kernel_ms > 0


def test_event_mode_u8_rows_and_odd_stride(gpu_ctx):
    runs = workloads.tunnelling_tl12(sims=4, n=400, T_rate=5.0)
    for p in runs:
        p.dt = 1.0                                   # max_dt: step cap 1 / T_rate
    dur = [160.0] * len(runs)
    print("glow: event mode TL12 4 x 400", flush=True)
    lum, t, er, steps, _ = gpu_ctx.run_event(runs, dur, max_steps=2048)
    assert lum.shape[1] % 4 == 0                      # four-step loads
    got, _ = gpu_ctx.glow_curve(runs, lum, t)
    want = oracle_from_rows(runs, lum, t)
    assert_curves(got, want)
    odd = 2047                                        # one-step loads
    got_odd, _ = gpu_ctx.glow_curve(runs, np.ascontiguousarray(lum[:, :odd]),
                                    np.ascontiguousarray(t[:, :odd]))
    assert_curves(got_odd, oracle_from_rows(runs, lum[:, :odd], t[:, :odd]))


def test_event_mode_through_plots(gpu_ctx):
    cfg = config.compose(["exp_type_fp.e_ratio_start=0.05", "+mode=event"])
    print("glow: event-mode sweep through plots.glow_curves", flush=True)
    configs = __import__("mcluminescence_amd.engine", fromlist=["x"]).initialize_runs(cfg)
    x_ax, lum, er, _ = run_sweep_event(configs, 20000, 2, ctx=gpu_ctx)
    got = plots.glow_curves(x_ax, lum, configs, ctx=gpu_ctx)
    want = glow.tl_curves(x_ax, lum, [(float(configs[r].exp_type_fp.T_start),
                                       float(configs[r].exp_type_fp.T_rate)) for r in range(4)])
    assert_curves(got, want)


@pytest.mark.parametrize("bw,win", [(1.0, 50.0), (0.5, 3.0), (2.5, 10.0), (1.0, 0.3), (1.0, 500.0)])
def test_synthetic_edges_and_widths(gpu_ctx, bw, win):
    rng = np.random.default_rng(int(bw * 10 + win))
    # integer temperatures (values on edges, the max on the last edge), a negative
    # start (values below 0 dropped), cooling, padding, a run with no steps
    runs = [ramp(0.0, 1.0, 3), ramp(-20.0, 0.7, 2), ramp(300.0, -1.5, 4), ramp(12.5, 3.0, 1),
            ramp(5.0, 1.0, 2)]
    steps = 256
    total = sum(r.sims for r in runs)
    counts = rng.integers(0, 7, size=(total, steps)).astype(np.uint32)
    t = np.tile(np.arange(1, steps + 1, dtype=np.float64), (total, 1))
    t[9:10, 200:] = 0.0                               # run 3's only sim: padded after step 200
    t[10:12, :] = 0.0                                 # run 4: no step with t > 0
    got, _ = gpu_ctx.glow_curve(runs, counts, t, bin_width=bw, win=win)
    want = oracle_from_rows(runs[:4], counts[:10], t[:10], bw, win)
    assert_curves(got[:4], want)
    assert got[4].size == 0


def test_sums_beyond_32_bits(gpu_ctx):
    rng = np.random.default_rng(9)
    runs = [ramp(0.0, 2.0, 8)]
    counts = rng.integers(2 ** 31, 2 ** 32 - 1, size=(8, 400), dtype=np.uint64).astype(np.uint32)
    t = np.tile(np.arange(1, 401, dtype=np.float64) * 0.5, (8, 1))
    got, _ = gpu_ctx.glow_curve(runs, counts, t)
    assert_curves(got, oracle_from_rows(runs, counts, t))


def test_too_many_bins_is_an_error(gpu_ctx):
    runs = [ramp(0.0, 100.0, 1)]
    counts = np.ones((1, 200), np.uint32)
    with pytest.raises(MclError) as e:
        gpu_ctx.glow_curve(runs, counts, np.arange(1, 201, dtype=np.float64)[None])
    assert e.value.code == MCL_ERANGE


def test_plot_all_writes_the_reference_files(gpu_ctx, tmp_path, monkeypatch):
    monkeypatch.setenv("MCL_PROJECT_ROOT", str(tmp_path))
    cfg = config.compose(["exp_type_fp.e_ratio_start=0.05", "+tag=glowplot",
                          "+plots=TL,TL_norm,iso,bg,nope"])
    print("glow: plot_all TL12", flush=True)
    written = plots.plot_all(cfg, ctx=gpu_ctx)
    names = sorted(p.name for p in written)
    assert names == ["TL_norm_run_glowplot.png", "TL_run_glowplot.png", "bg_run_glowplot.png",
                     "iso_run_glowplot.png"]
    assert (tmp_path / "results" / "simulations" / "exp_glowplot_x_ax.csv").exists()
    again = plots.plot_all(cfg.copy(), from_csv=True, ctx=gpu_ctx)
    assert sorted(p.name for p in again) == names
