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
             on the launch stream.
cpu_baseline = the numpy event-mode restatement of the reference loop
             (simulate.py:460-516, one recombination per step, dense distance
             matrix) on a bounded sample of the same clusters, one core,
             expressed in the same unit (fixed-dt-equivalent live electron-steps).
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
                      f"{el:.1f} s, numpy event-mode restatement of simulate.py:460-516",
            "curves_per_s": done / el}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sims", type=int, default=4096)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: the Monte Carlo shards clusters with no "
              f"exchange, and this build measures one GPU's share (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        return 2

    import torch
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
        "dtype": "fp32+fp64",
        "data": "synthetic (Philox-generated geometry, seed 0x4D434C554D000001)",
        "config": {"workload": "iso_decay_500K_4096x400", "clusters": n, "electrons": 400,
                   "holes": 691, "T_K": 500.0, "dt_s": 0.5, "n_steps": 4000, "E_eV": 1.4,
                   "s_per_s": 1e13, "alpha_per_m": 4e9, "rho_per_m3": 1e23,
                   "kernel": "fixed_dt_kernel (1 wave64 per cluster)"},
        "work_per_step": {"live_electron_steps": w.live_electron_steps,
                          "distance_evals": w.distance_evals,
                          "recombinations": w.recombinations,
                          "cluster_steps": w.cluster_steps},
        "curves_per_s": args.steps * n / wall,
        "roofline": {"bound": "valu", "achieved": achieved / 1e12,
                     "peak": PEAK_VALU_LANE_SLOTS / 1e12, "unit": "Tslot/s",
                     "frac": achieved / PEAK_VALU_LANE_SLOTS, "traffic": None,
                     "kernel_ms": mean_ms,
                     "hbm_gbps_algorithmic": out_bytes / (mean_ms * 1e-3) / 1e9,
                     "hbm_peak_gbps": PEAK_HBM_GBPS},
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
        result["cpu_baseline"]["host_cpu"] = _cpu_model()
        result["cpu_baseline"]["os_cpu_count"] = os.cpu_count()
    print(json.dumps(result), flush=True)
    plan.close()
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
