#!/usr/bin/env python
"""bench.py — MC electron-steps/s of the fixed-dt trap-kinetics Monte Carlo on one MI355X.

Workload (one "step" = one launch over the whole batch): isothermal decay at
500 K for 2000 s with dt = 0.5 s (4000 steps), 4096 clusters x 400 electrons /
400 core holes (691 with the 1.2 boundary shell), rho = 1e23 m^-3, E = 1.4 eV,
s = 1e13 1/s, alpha = 4e9 1/m; per-step recombination counts out
(4096 x 4000 uint32).  Inputs are synthetic (Philox-generated coordinates)
and device-resident before the timed region.

value      = live electron-steps per second over the K timed launches
             (whole-job wall time between two device synchronisations).
roofline   = the MC kernel against the VALU issue rate of the chip: algorithmic
             lane-slots per launch (per live electron-step 1/4 Philox4x32-10
             block + fire test = 28 slots; per electron-hole distance 9 slots)
             divided by the kernel's mean launch time from HIP events recorded
             on the launch stream.  roofline_hbm = the same kernel against HBM
             (the counts it must write), which is not what binds it.
cpu_baseline = the numpy event-mode restatement of the reference loop
             (simulate.py:36-92, one recombination per step, dense distance
             matrix) on a bounded sample of the same clusters, one core,
             expressed in the same unit (fixed-dt-equivalent live electron-steps);
             with the extras also on every host core (one process per cluster)
             and the C fixed-dt replay harness on one and on every core.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# MI355X: 256 CUs x 64 lanes x 2.4 GHz VALU issue slots (non-packed)
PEAK_VALU_LANE_SLOTS = 256 * 64 * 2.4e9
PEAK_HBM_GBPS = 8000.0
SLOTS_PER_LIVE_ESTEP = 28.0     # 1/4 x (20 v_mad_u64_u32 x 4 + 20 v_xor3) + fma/cmp/select
SLOTS_PER_DISTANCE = 9.0        # 3 sub, mul, 2 fma, cmp, 2 cndmask
# HBM bytes per launch of the decay workload, from rocprofv3 FETCH_SIZE (616.0 KiB,
# x2 by a plain-copy calibration) and WRITE_SIZE (64,260.6 KiB, x1)
TRAFFIC_BYTES_DECAY = (616.0 * 2 + 64260.6) * 1024


def cpu_baseline(seconds: float) -> dict:
    """Time the numpy event-mode port on whole clusters until ~seconds have elapsed."""
    import numpy as np
    from oracle import event_mc as E
    from mcluminescence_amd import workloads as W

    run = W.iso_decay_4096x400(sims=1)[0]
    exp, phys = E.run_params_to_cfg(run)
    n0 = int(run.N_e * run.e_ratio_start)
    done = 0
    esteps = 0.0
    t0 = time.perf_counter()
    while True:
        x, lum, _ = E.simulate_run(exp, phys, sims=1, steps=n0 + 16, seed=0xC0FFEE, first_sim=done)
        esteps += float(E.live_electron_steps(x, lum, n0, run.dt, run.n_steps).sum())
        done += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    return {"value": esteps / el, "unit": "electron-steps/s", "cores": 1, "kind": "port",
            "sample": f"{done} of the 4096 clusters (400 electrons, 2000 s at 500 K) in "
                      f"{el:.1f} s, numpy event-mode restatement of simulate.py:36-92",
            "curves_per_s": done / el}


def _port_cluster(k: int) -> float:
    """Live electron-steps of decay cluster k by the numpy port (a pool worker)."""
    from oracle import event_mc as E
    from mcluminescence_amd import workloads as W
    run = W.iso_decay_4096x400(sims=1)[0]
    exp, phys = E.run_params_to_cfg(run)
    n0 = int(run.N_e * run.e_ratio_start)
    x, lum, _ = E.simulate_run(exp, phys, sims=1, steps=n0 + 16, seed=0xC0FFEE, first_sim=k)
    return float(E.live_electron_steps(x, lum, n0, run.dt, run.n_steps).sum())


def _port_warm(_: int) -> int:
    from oracle import event_mc  # noqa: F401
    from mcluminescence_amd import workloads  # noqa: F401
    return 0


def cpu_baseline_pool(seconds: float) -> dict:
    """The numpy port on every host core (at most 16): one process per cluster,
    OMP_NUM_THREADS=1, clusters handed out until ~seconds have elapsed.  Called
    before the process touches the GPU, so the forked workers carry no GPU state."""
    import multiprocessing as mp
    cores = min(16, os.cpu_count() or 1)
    old = os.environ.get("OMP_NUM_THREADS")
    os.environ["OMP_NUM_THREADS"] = "1"
    try:
        esteps, done = 0.0, 0
        with mp.get_context("fork").Pool(cores) as pool:
            pool.map(_port_warm, range(cores))        # imports, outside the timed region
            t0 = time.perf_counter()
            el = 0.0
            for v in pool.imap_unordered(_port_cluster, range(4096)):
                esteps += v
                done += 1
                el = time.perf_counter() - t0
                if el >= seconds:
                    break
            pool.terminate()
    finally:
        if old is None:
            os.environ.pop("OMP_NUM_THREADS", None)
        else:
            os.environ["OMP_NUM_THREADS"] = old
    return {"value": esteps / el, "unit": "electron-steps/s", "cores": cores, "kind": "port",
            "sample": f"{done} of the 4096 decay clusters in {el:.1f} s, numpy event-mode "
                      f"restatement of simulate.py:36-92, one process per cluster"}


def cpu_event_lab(seconds: float) -> dict:
    """The numpy restatements of the event loop (simulate.py:36-92) and of TL_lab
    (tl_trap_lab.py:65-107) on the workloads of event_lab_extras, one core, each for
    about `seconds`."""
    import numpy as np
    from oracle import event_mc as E
    from oracle import lab_mc
    run = event_workload(sims=1)[0]
    exp, phys = E.run_params_to_cfg(run)
    exp["duration"] = EVENT_DURATION
    exp["max_dt"] = run.dt
    done, steps = 0, 0
    t0 = time.perf_counter()
    while True:
        x, _, _ = E.simulate_run(exp, phys, sims=1, steps=4096, seed=0xE7E47, first_sim=done)
        steps += int((x[:, 0] > 0).sum())
        done += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    event = {"value": steps / el, "unit": "event steps/s", "cores": 1, "kind": "port",
             "sample": f"{done} clusters of the event workload in {el:.1f} s, numpy "
                       "restatement of simulate.py:36-92"}
    # the 16 x 500 TL glow (300 -> 700 K at 1 K/s) through the same loop, step cap 0.1 s
    from mcluminescence_amd import workloads as W
    grun = W.tl_glow_16x500(sims=1)[0]
    gexp, gphys = E.run_params_to_cfg(grun)
    gexp["duration"] = grun.n_steps * grun.dt
    gexp["max_dt"] = grun.dt * grun.T_rate
    done = 0
    t0 = time.perf_counter()
    while True:
        E.simulate_run(gexp, gphys, sims=1, steps=8192, seed=0x61077, first_sim=done)
        done += 1
        el = time.perf_counter() - t0
        if el >= seconds or done >= 16:
            break
    tl_glow = {"value": el / done * 16 * 1e3, "unit": "ms per run", "cores": 1, "kind": "port",
               "sample": f"{done} of the 16 clusters of the 16 x 500 TL glow in {el:.1f} s "
                         "(the run's time scaled to 16 clusters), numpy restatement of "
                         "simulate.py:36-92 with the step cap 0.1 s"}
    lexp, lphys, rows = lab_workload()
    rng = np.random.default_rng(11)
    done = 0
    t0 = time.perf_counter()
    while True:
        lab_mc.tl_lab_fractions(lexp, lphys, rows, rng)
        done += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    lab = {"value": done * len(rows) / el, "unit": "lab experiments/s", "cores": 1,
           "kind": "port", "sample": f"{done} passes over the {len(rows)} TL rows in {el:.1f} s, "
                                     "numpy restatement of TL_lab (tl_trap_lab.py:65-107)"}
    return {"event": event, "lab": lab, "tl_glow_16x500": tl_glow}


EVENT_DURATION = 160.0      # s: 0 -> 800 degC at 5 K/s


def event_workload(sims: int = 4096):
    """The reference's TL12 physics (tunnelling only) at 400 electrons, 5 K/s, step
    cap max_dt / T_rate = 0.2 s: simulate.py's own loop, one event per step."""
    from mcluminescence_amd import workloads as W
    runs = W.tunnelling_tl12(sims=sims, n=400, T_rate=5.0)
    for p in runs:
        p.dt = 1.0                               # max_dt (the event kernel caps dt at it / T_rate)
    return runs


def lab_workload():
    """TL_lab on the CLBR_IRSL50 table (11 rows) with the shipped lab configs."""
    import json as _json
    from mcluminescence_amd import config, engine
    from oracle import event_mc as E
    t = _json.loads((ROOT / "tests" / "golden" / "reference_outputs.json").read_text())[
        "lab_data"]["CLBR_IRSL50_0.25KperGy"]
    cfg = engine.initialize_runs(config.compose(["exp_type_fp=TLlab", "physics_fp=lab_TL"]))[0]
    p = engine.run_params(cfg, sims=1, n_steps=1)
    exp, phys = E.run_params_to_cfg(p)
    phys["D"] = float(cfg.physics_fp.D)
    rows = [dict(Duration=t["Duration"][k], T_start=t["T_start"][k], T_end=t["T_end"][k],
                 e_ratio=t["e_ratio"][k]) for k in range(len(t["Duration"]))]
    return exp, phys, rows


def event_lab_extras(ctx) -> dict:
    """Event mode (4096 clusters of the event workload) and the lab kernel (TL_lab on
    the CLBR table, 4096 simulations per row): kernel time from HIP events, steps
    simulated per second."""
    import numpy as np
    from mcluminescence_amd import config, engine, lab
    runs = event_workload()
    out = {}
    ms, st = [], None
    for _ in range(4):
        _, _, _, steps, st = ctx.run_event(runs, [EVENT_DURATION], max_steps=4096)
        ms.append(st.kernel_ms)
    m = float(np.median(ms[1:]))
    out["event"] = {"workload": "TL12 physics, 4096 clusters x 400 electrons, 0 -> 800 degC at 5 K/s",
                    "kernel_ms": m, "event_steps": int(st.cluster_steps),
                    "event_steps_per_s": st.cluster_steps / (m * 1e-3),
                    "recombinations": int(st.recombinations)}
    cfg = engine.initialize_runs(config.compose(["exp_type_fp=TLlab", "physics_fp=lab_TL"]))[0]
    import json as _json
    table = _json.loads((ROOT / "tests" / "golden" / "reference_outputs.json").read_text())[
        "lab_data"]["CLBR_IRSL50_0.25KperGy"]
    lruns, obs = lab.tl_lab_runs(cfg, table, sims=4096)
    ms = []
    for _ in range(4):
        _, _, st = ctx.run_lab(lruns, obs)
        ms.append(st.kernel_ms)
    m = float(np.median(ms[1:]))
    n_exp = sum(int(r.sims) for r in lruns)
    out["lab"] = {"workload": "TL_lab, CLBR_IRSL50 table (11 rows) x 4096 simulations, lab_TL physics",
                  "kernel_ms": m, "experiments": n_exp, "experiments_per_s": n_exp / (m * 1e-3),
                  "event_steps": int(st.cluster_steps),
                  "event_steps_per_s": st.cluster_steps / (m * 1e-3)}
    return out


PEAK_BF16_TFLOPS = 2500.0       # dense BF16 matrix-core peak


def tl_extras(ctx, stream) -> dict:
    """Glow-curve throughput: the 16 x 500 TL glow workload (300 -> 700 K at 1 K/s,
    dt 0.1 s, 4000 steps) and batched 256-bin curves of the surrogate's window
    (1024 parameter sets x 8 clusters x 400 electrons, 2048 steps)."""
    import numpy as np
    import torch
    from mcluminescence_amd import surrogate as S
    from mcluminescence_amd import workloads as W
    plan = ctx.plan(W.tl_glow_16x500(), n_bins=4000)
    counts = torch.zeros((16, 4000), dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    plan.launch(counts.data_ptr())
    plan.sync()
    ms = []
    for _ in range(5):
        plan.launch(counts.data_ptr())
        plan.sync()
        ms.append(plan.last_ms())
    plan.close()
    glow = glow_extras(ctx, W.tl_glow_16x500(), counts)
    theta = S.sample_params(1024, np.random.default_rng(3))
    S.mc_curves(ctx, theta[:64], n=400, sims=8)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    S.mc_curves(ctx, theta, n=400, sims=8)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"tl_glow_16x500_ms": float(np.median(ms)),
            "glow_curve": glow,
            "tl_curves_per_s": 1024 / dt,
            "tl_curves_note": "256-bin glow curves of 8 clusters x 400 electrons each, 2048 steps, "
                              "1024 parameter sets per call, host round trip included"}


def sweep_extras(ctx) -> dict:
    """The uneven batched sweep: 10,000 parameter sets (half TL at 0.5-5 K/s over
    300-750 K, half isothermal at 400-600 K; 64-2048 electrons; 1024 steps) as one
    persistent, graph-captured launch with the tuning table's knobs, 256-bin curves
    out; and the same sets as one workgroup per cluster for comparison.  Device time
    of a launch from HIP events around it."""
    import numpy as np
    import torch
    from mcluminescence_amd import autotune
    from mcluminescence_amd import workloads as W
    from mcluminescence_amd._abi import F_GRAPH, F_PERSISTENT
    runs = W.uneven_sweep()
    knobs = autotune.apply(ctx, runs)
    out = {"workload": "uneven sweep: 10,000 sets (5,000 TL, 5,000 isothermal), 64-2048 "
                       "electrons, 1024 steps, 256-bin curves",
           "sets": len(runs), "knobs": knobs,
           "tuned": autotune.signature(runs) in autotune.load_table()}
    ref = None
    for label, flags in (("persistent_graph", F_PERSISTENT | F_GRAPH), ("one_wg_per_cluster", 0)):
        plan = ctx.plan(runs, n_bins=256, flags=flags)
        n, row = plan.info.total_clusters, plan.info.row_len
        counts = torch.zeros((n, row), dtype=torch.int32, device="cuda")
        live = torch.zeros(n, dtype=torch.int64, device="cuda")
        plan.launch(counts.data_ptr(), live.data_ptr())
        plan.sync()
        ms = []
        for _ in range(3):
            plan.launch(counts.data_ptr(), live.data_ptr())
            plan.sync()
            ms.append(plan.last_ms())
        w = plan.work()
        got = counts.cpu().numpy()
        if ref is None:
            ref = got
        else:
            assert np.array_equal(ref, got), "persistent and plain launches disagree"
        plan.close()
        m = float(np.median(ms))
        slots = w.live_electron_steps * SLOTS_PER_LIVE_ESTEP + w.distance_evals * SLOTS_PER_DISTANCE
        out[label] = {"ms": m, "curves_per_s": n / (m * 1e-3),
                      "live_electron_steps_per_s": w.live_electron_steps / (m * 1e-3),
                      "valu_frac": slots / (m * 1e-3) / PEAK_VALU_LANE_SLOTS}
        out["live_electron_steps"] = int(w.live_electron_steps)
        out["distance_evals"] = int(w.distance_evals)
        out["recombinations"] = int(w.recombinations)
    autotune.apply(ctx, [], {})                  # back to the default knobs
    return out


def large_cluster_extras(ctx) -> dict:
    """Clusters beyond one workgroup's LDS: TL glow curves (E = 1.3 eV, s = 1e12 1/s,
    alpha = 4e9 1/m, rho = 1e23 m^-3, 300 -> 700 K at 1 K/s, dt 0.1 s, 4000 steps) at
    about 2 M electrons per launch, from 3,000 electrons a cluster (LDS) to 32,768
    (holes, cell starts and electrons in a global scratch slice per resident
    workgroup).  Device time of a launch from HIP events around it (median of 3)."""
    import numpy as np
    import torch
    from mcluminescence_amd import workloads as W
    out = {"workload": "TL glow 300 -> 700 K, 4000 steps of 0.1 s, about 2 M electrons per "
                       "launch", "cases": []}
    for n in (3000, 4000, 8000, 16000, 32768):
        sims = max(8, 2_000_000 // n)
        runs = [W.single_trap(1.3, 1e12, 4e9, 1e23, n, T_start_K=300.0, T_rate=1.0, dt=0.1,
                              n_steps=4000, sims=sims)]
        plan = ctx.plan(runs)
        info = plan.info
        counts = torch.zeros((info.total_clusters, info.row_len), dtype=torch.int32,
                             device="cuda")
        live = torch.zeros(info.total_clusters, dtype=torch.int64, device="cuda")
        plan.launch(counts.data_ptr(), live.data_ptr())
        plan.sync()
        ms = []
        for _ in range(3):
            plan.launch(counts.data_ptr(), live.data_ptr())
            plan.sync()
            ms.append(plan.last_ms())
        w = plan.work()
        m = float(np.median(ms))
        out["cases"].append({"electrons": n, "clusters": sims, "ms": m,
                             "live_electron_steps_per_s": w.live_electron_steps / (m * 1e-3),
                             "lds_bytes": int(info.lds_bytes), "grid": int(info.grid),
                             "block": int(info.block)})
        print(f"large clusters: {n} electrons x {sims}: {m:.2f} ms", file=sys.stderr, flush=True)
        plan.close()
    return out


def cpu_harness(seconds: float) -> dict:
    """The C replay harness (the fixed-dt algorithm the kernels run, OpenMP over
    clusters) on bounded samples: the decay clusters on one core and on every host
    core (at most 16), and the first sets of the uneven sweep on every core."""
    from oracle import orc
    from mcluminescence_amd import workloads as W
    cores = min(16, os.cpu_count() or 1)
    out = {}
    for label, th in (("decay_1core", 1), ("decay_allcores", cores)):
        sims, el, live = 16, 0.0, 0
        while True:
            t0 = time.perf_counter()
            _, lv = orc.run_fixed_dt(W.iso_decay_4096x400(sims=sims), threads=th)
            el = time.perf_counter() - t0
            live = int(lv.sum())
            if el >= seconds / 4 or sims >= 4096:
                break
            sims = min(4096, sims * 4)
        out[label] = {"value": live / el, "unit": "electron-steps/s", "cores": th,
                      "sample": f"{sims} of the 4096 decay clusters in {el:.2f} s"}
    # the 16 x 500 TL glow (all 16 clusters, 4000 steps), whole run
    for label, th in (("tl_glow_16x500_1core", 1), ("tl_glow_16x500_allcores", cores)):
        t0 = time.perf_counter()
        _, lv = orc.run_fixed_dt(W.tl_glow_16x500(), threads=th)
        el = time.perf_counter() - t0
        out[label] = {"value": el * 1e3, "unit": "ms per run", "cores": th,
                      "electron_steps_per_s": int(lv.sum()) / el,
                      "sample": "the whole 16 x 500 TL glow (16 clusters, 4000 steps)"}
    runs = W.uneven_sweep()
    k = 64
    while True:
        t0 = time.perf_counter()
        _, lv = orc.run_fixed_dt(runs[:k], n_bins=256, threads=cores)
        el = time.perf_counter() - t0
        if el >= seconds / 4 or k >= len(runs):
            break
        k = min(len(runs), k * 4)
    out["sweep_allcores"] = {"value": k / el, "unit": "curves/s", "cores": cores,
                             "electron_steps_per_s": int(lv.sum()) / el,
                             "sample": f"the first {k} of the 10,000 sweep sets in {el:.2f} s"}
    out["kind"] = "port"
    out["what"] = "oracle/mcl_oracle.c fixed-dt replay harness (gcc -O2, OpenMP)"
    return out


def glow_extras(ctx, runs, counts, batch: int = 1024) -> dict:
    """The glow-curve kernel (1-degree bins, 50-degree boxcar) on the 16 x 500 glow's
    per-step counts, alone and as a batch of `batch` copies of that run (HBM-bound:
    algorithmic bytes = clusters x steps x 4 B of counts + the curves)."""
    import numpy as np
    import torch
    out = {}
    for label, reps in (("one_run", 1), (f"batch_{batch}", batch)):
        c = counts.repeat(reps, 1).contiguous()
        rr = runs * reps
        curve = torch.zeros((reps, 512), dtype=torch.float64, device="cuda")
        lens = torch.zeros(reps, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        ms = []
        for _ in range(6):
            st = ctx.glow_curve_device(rr, c.data_ptr(), 4, c.shape[1], None, 1.0, 50.0,
                                       curve.data_ptr(), 512, lens.data_ptr())
            ms.append(st.kernel_ms)
        m = float(np.median(ms[1:]))
        nbytes = c.numel() * 4 + int(lens.sum().item()) * 8
        out[label] = {"runs": reps, "clusters": int(c.shape[0]), "steps": int(c.shape[1]),
                      "points": int(lens[0].item()), "kernel_ms": m,
                      "hbm_gbps": nbytes / (m * 1e-3) / 1e9,
                      "frac": nbytes / (m * 1e-3) / 1e9 / PEAK_HBM_GBPS}
    return out


def iso_fit_extras(ctx) -> dict:
    """The isothermal half of the fit workload: the ISO network's held-out NRMSE
    (64 resolved decays against 1024-cluster Monte Carlo) and one surrogate-assisted
    fit of a synthetic isothermal lab curve against the plain Monte Carlo fit."""
    import numpy as np
    from mcluminescence_amd import fit as F
    from mcluminescence_amd import fit_bench
    from mcluminescence_amd import surrogate as S
    from mcluminescence_amd import train_surrogate as T
    net = S.Network.load(ctx, fit_bench.MODELS["iso"], max_batch=4096)
    theta = T.sample(S.ISO, 64, np.random.default_rng(2024), resolved_only=True)
    mc = S.ISO.mc_curves(ctx, theta, n=400, sims=1024, seed=S.DEFAULT_SEED + 4243)
    e = S.nrmse(S.predict_curves(net, theta), mc)
    rng = np.random.default_rng(7)
    truth = fit_bench.draw_truth_iso(rng)
    y, scale = F.lab_curve(ctx, truth, rng=rng, kind=S.ISO)
    F.fit_surrogate(ctx, net, y, scale, truth[5], pop=1024, gens=2, seed=9)   # loads kernels
    F.fit_mc(ctx, y, scale, truth[5], pop=8, gens=1, seed=9, kind=S.ISO)
    rs = F.fit_surrogate(ctx, net, y, scale, truth[5], pop=1024, gens=150, seed=1)
    rm = F.fit_mc(ctx, y, scale, truth[5], pop=128, gens=40, seed=1, kind=S.ISO)
    jo = lambda th: F.judge_objective(ctx, th, y, scale, kind=S.ISO)
    net.close()
    return {"heldout_nrmse_median": float(np.median(e)),
            "heldout_nrmse_p90": float(np.percentile(e, 90)),
            "heldout_note": "64 resolved decays (1/k in [10^-1, 10^4.5] s) vs 1024-cluster MC",
            "fit_seconds_surrogate": rs.seconds, "fit_seconds_mc": rm.seconds,
            "fit_speedup": rm.seconds / rs.seconds, "fit_judge_nrmse": rs.judge_nrmse,
            "fit_fallback": rs.fallback,
            "fit_objective_surrogate": jo(rs.theta), "fit_objective_mc": jo(rm.theta),
            "fit_objective_truth": jo(truth),
            "recovery_surrogate": rs.recovery(truth), "recovery_mc": rm.recovery(truth),
            "fit_note": "isothermal decay at T_iso = %.1f K, 1/k = %.1f s: surrogate DE 1024 x 150 "
                        "generations + one 64 x 1000 Monte Carlo judge vs plain Monte Carlo DE "
                        "128 x 40 generations" % (truth[5], 1.0 / F.iso_rate(truth)[0])}


def surrogate_extras(ctx, fits: bool = True) -> dict:
    """Surrogate network: one AdamW training step at batch 4096 and one inference
    batch of 65,536 (device time, HIP events), and (fits) one surrogate-assisted fit
    of a synthetic lab curve against the plain Monte Carlo fit of the same curve."""
    import numpy as np
    import torch
    from mcluminescence_amd import fit as F
    from mcluminescence_amd import fit_bench
    from mcluminescence_amd import surrogate as S
    from mcluminescence_amd._abi import AdamW
    net = S.Network.load(ctx, fit_bench.MODEL, max_batch=65536)
    flop = 2 * (6 * 512 + 3 * 512 * 512 + 512 * 256)      # per sample, unpadded
    g = torch.Generator(device="cpu").manual_seed(0)
    x = (torch.rand((65536, 6), generator=g) * 2 - 1).cuda()
    t = torch.rand((65536, 256), generator=g).cuda()
    for _ in range(3):
        net.forward(x)
    inf = []
    for _ in range(10):
        net.forward(x)
        inf.append(net.last_ms())
    idx = torch.arange(4096, dtype=torch.int32, device="cuda")
    params = net.get_params()
    hp = AdamW(1e-6, 0.9, 0.999, 1e-8, 0.0, 1)
    tr = []
    for k in range(20):
        hp.step = k + 1
        net.train_step(x, t, idx, hp)
        tr.append(net.last_ms())
    net.set_params(params)
    inf_ms, tr_ms = float(np.median(inf)), float(np.median(tr[3:]))
    timing = {"inference_b65536_ms": inf_ms,
              "inference_tflops": 65536 * flop / (inf_ms * 1e-3) / 1e12,
              "inference_frac_bf16_peak": 65536 * flop / (inf_ms * 1e-3) / 1e12 / PEAK_BF16_TFLOPS,
              "train_step_b4096_ms": tr_ms,
              "train_tflops": 3 * 4096 * flop / (tr_ms * 1e-3) / 1e12}
    if not fits:
        net.close()
        return timing
    rng = np.random.default_rng(7)
    truth = fit_bench.draw_truth(rng)
    y, scale = F.lab_curve(ctx, truth, rng=rng)
    F.fit_surrogate(ctx, net, y, scale, truth[5], pop=1024, gens=2, seed=9)   # loads kernels
    F.fit_mc(ctx, y, scale, truth[5], pop=8, gens=1, seed=9)
    rs = F.fit_surrogate(ctx, net, y, scale, truth[5], pop=1024, gens=150, seed=1)
    rm = F.fit_mc(ctx, y, scale, truth[5], pop=128, gens=40, seed=1)
    net.close()
    iso = iso_fit_extras(ctx)
    return {**timing, "iso": iso,
            "fit_seconds_surrogate": rs.seconds, "fit_seconds_mc": rm.seconds,
            "fit_speedup": rm.seconds / rs.seconds, "fit_judge_nrmse": rs.judge_nrmse,
            "fit_fallback": rs.fallback,
            "fit_objective_surrogate": F.judge_objective(ctx, rs.theta, y, scale),
            "fit_objective_mc": F.judge_objective(ctx, rm.theta, y, scale),
            "fit_note": "surrogate DE 1024 x 150 generations + one 64 x 1000 Monte Carlo judge "
                        "vs plain Monte Carlo DE 128 x 40 generations (64 x 1000 per candidate)"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sims", type=int, default=4096)
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the glow-curve and surrogate figures")
    ap.add_argument("--surrogate-only", action="store_true",
                    help="only time the surrogate network (training step, inference batch)")
    ap.add_argument("--sweep-only", action="store_true",
                    help="only time the uneven 10,000-set sweep (and its CPU harness leg "
                         "unless --no-cpu)")
    ap.add_argument("--large-clusters", action="store_true",
                    help="only time TL glow launches of clusters from 3,000 to 32,768 "
                         "electrons (the global-memory form beyond LDS)")
    ap.add_argument("--copy-gib", type=float, default=0.0,
                    help="only run a plain device copy of this many GiB three times (the "
                         "reference point for hardware byte counters) and exit")
    args = ap.parse_args()
    if args.surrogate_only:
        from mcluminescence_amd._abi import Context
        ctx = Context(0)
        print(json.dumps(surrogate_extras(ctx, fits=False)), flush=True)
        ctx.close()
        return 0
    if args.sweep_only:
        from mcluminescence_amd._abi import Context
        ctx = Context(0)
        res = {"sweep": sweep_extras(ctx)}
        ctx.close()
        if not args.no_cpu:
            res["cpu_harness"] = cpu_harness(args.cpu_seconds)
        print(json.dumps(res), flush=True)
        return 0
    if args.large_clusters:
        from mcluminescence_amd._abi import Context
        ctx = Context(0)
        print(json.dumps(large_cluster_extras(ctx)), flush=True)
        ctx.close()
        return 0
    if args.copy_gib > 0:
        import torch
        n = int(args.copy_gib * 2 ** 30) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": 4 * n, "copies": 3}), flush=True)
        return 0

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: the Monte Carlo shards clusters with no "
              f"exchange, and this build measures one GPU's share (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        return 2

    pool_leg = None
    if not args.no_cpu and not args.no_extras:
        pool_leg = cpu_baseline_pool(args.cpu_seconds / 2)

    import torch
    from mcluminescence_amd import autotune
    from mcluminescence_amd import workloads as W
    from mcluminescence_amd._abi import Context

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    # a dedicated stream (default priority): the library launches on it and the
    # HIP events below are recorded on it, so they bracket exactly the MC kernel
    stream = torch.cuda.Stream(device=dev)
    ctx = Context(0)
    ctx.set_stream(stream.cuda_stream)
    runs = W.iso_decay_4096x400(sims=args.sims)
    knobs = autotune.apply(ctx, runs)         # the tuning table's knobs (defaults if untuned)
    plan = ctx.plan(runs)                     # per-step counts: n_bins = 0, rows of n_steps
    n, nb = plan.info.total_clusters, plan.info.row_len
    counts = torch.zeros((n, nb), dtype=torch.int32, device=dev)
    live = torch.zeros(n, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        plan.launch(counts.data_ptr(), live.data_ptr())
    torch.cuda.synchronize()
    work = plan.work() if args.warmup > 0 else None

    evs = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    evs[0].record(stream)
    for k in range(args.steps):
        plan.launch(counts.data_ptr(), live.data_ptr())
        evs[k + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [evs[k].elapsed_time(evs[k + 1]) for k in range(args.steps)]
    mean_ms = sum(kernel_ms) / len(kernel_ms)
    w = plan.work()
    if work is not None:
        assert w.live_electron_steps == work.live_electron_steps, "non-deterministic work"

    value = args.steps * w.live_electron_steps / wall
    slots = w.live_electron_steps * SLOTS_PER_LIVE_ESTEP + w.distance_evals * SLOTS_PER_DISTANCE
    achieved = slots / (mean_ms * 1e-3)
    out_bytes = n * nb * 4 + n * 8
    result = {
        "metric": "MC electron-steps/s",
        "value": value,
        "unit": "electron-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (Philox-generated geometry, seed 0x4D434C554D000001)",
        "config": {"workload": "iso_decay_500K_4096x400", "clusters": n, "electrons": 400,
                   "holes": 691, "T_K": 500.0, "dt_s": 0.5, "n_steps": 4000, "E_eV": 1.4,
                   "s_per_s": 1e13, "alpha_per_m": 4e9, "rho_per_m3": 1e23,
                   "kernel": "fixed_dt_kernel<false, true, false> (1 wave64 per cluster, clusters in LDS, fire thresholds kept in the w slots, no quiet window)", "knobs": knobs},
        "work_per_step": {"live_electron_steps": w.live_electron_steps,
                          "distance_evals": w.distance_evals,
                          "recombinations": w.recombinations,
                          "cluster_steps": w.cluster_steps},
        "curves_per_s": args.steps * n / wall,
        "roofline": {"bound": "valu", "achieved": achieved / 1e12,
                     "peak": PEAK_VALU_LANE_SLOTS / 1e12, "unit": "Tslot/s",
                     "frac": achieved / PEAK_VALU_LANE_SLOTS,
                     "traffic": TRAFFIC_BYTES_DECAY if args.sims == 4096 else None,
                     "kernel_ms": mean_ms,
                     "hbm_gbps_algorithmic": out_bytes / (mean_ms * 1e-3) / 1e9,
                     "hbm_peak_gbps": PEAK_HBM_GBPS},
        # the same kernel against the HBM roof: algorithmic bytes (the counts and
        # live-step words it must write) over the same mean launch time
        "roofline_hbm": {"bound": "hbm", "achieved": out_bytes / (mean_ms * 1e-3) / 1e9,
                         "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                         "frac": out_bytes / (mean_ms * 1e-3) / 1e9 / PEAK_HBM_GBPS,
                         "traffic": TRAFFIC_BYTES_DECAY if args.sims == 4096 else None},
    }
    plan.close()
    if not args.no_extras:
        ctx.set_stream(None)
        result["tl"] = tl_extras(ctx, stream)
        result["sweep"] = sweep_extras(ctx)
        result.update(event_lab_extras(ctx))
        result["surrogate"] = surrogate_extras(ctx)
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
        if not args.no_extras:
            result["cpu_baseline"].update(cpu_event_lab(args.cpu_seconds / 3))
            result["cpu_baseline"]["harness"] = cpu_harness(args.cpu_seconds / 3)
            result["cpu_baseline"]["all_cores"] = pool_leg
        result["cpu_baseline"]["host_cpu"] = _cpu_model()
        result["cpu_baseline"]["os_cpu_count"] = os.cpu_count()
        # no published number exists for this metric (vs_baseline stays null); the
        # ratio to the CPU leg measured here is reported beside it
        cb = result["cpu_baseline"]
        result["vs_cpu_baseline"] = {"ratio": value / cb["value"], "cores": cb["cores"],
                                     "kind": cb["kind"]}
    print(json.dumps(result), flush=True)
    ctx.close()
    return 0


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    sys.exit(main())
