"""Ensemble fixtures of the event-mode oracle (oracle/event_mc.py, the numpy
restatement of simulate.py:36-92) for the GPU ensemble test.

For each case the oracle runs `sims` independent clusters in event mode; events
are binned by time onto the fixed-dt step grid (oracle.event_mc.bin_events) and
then into coarser bins; the fixture keeps the parameter set, the bin width and,
per bin, the mean and standard deviation over the clusters of the
recombinations per cluster.  Writes tests/golden/ensemble_event_mode.json.

Run: python tests/golden/make_ensemble_fixtures.py   (a few minutes on 8 cores)
"""
import json
import sys
from concurrent.futures import ProcessPoolExecutor
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from mcluminescence_amd import workloads as W  # noqa: E402
from oracle import event_mc as E  # noqa: E402

OUT = Path(__file__).resolve().parent / "ensemble_event_mode.json"
SEED = 20240611


def cases():
    iso = W.iso_decay_4096x400(sims=1)[0]
    iso.n_steps = 400                                   # 200 s: the decay is over by then
    tl = W.tl_glow_16x500(sims=1)[0]
    tun = W.tunnelling_tl12(sims=1, n=400, T_rate=1.0)[0]   # TL12 + basicTL12, 400 electrons
    return {
        # name: (run, sims, steps per coarse bin)
        "iso_500K_400": (iso, 2048, 10),
        "tl_glow_500": (tl, 128, 50),
        "tl12_tunnelling_400": (tun, 128, 20),
    }


def _one(args):
    name, run, first, count = args
    exp, phys = E.run_params_to_cfg(run)
    n0 = int(run.N_e * run.e_ratio_start)
    # event steps: every recombination plus the dt-capped steps of the ramp
    steps = n0 + int(run.n_steps) + 64
    x, lum, _ = E.simulate_run(exp, phys, sims=count, steps=steps, seed=SEED, first_sim=first)
    return E.bin_events(x, lum, run.dt, int(run.n_steps))


def main():
    fx = {}
    with ProcessPoolExecutor(max_workers=8) as pool:
        for name, (run, sims, per_bin) in cases().items():
            chunks = [(name, run, f, min(16, sims - f)) for f in range(0, sims, 16)]
            per_step = np.concatenate(list(pool.map(_one, chunks)), axis=1)   # (n_steps, sims)
            nb = int(run.n_steps) // per_bin
            coarse = per_step[: nb * per_bin].reshape(nb, per_bin, sims).sum(axis=1)
            fx[name] = {
                "run": {k: getattr(run, k) for k, _ in type(run)._fields_},
                "sims": sims, "steps_per_bin": per_bin, "n_bins": nb,
                "mean": coarse.mean(axis=1).tolist(),
                "std": coarse.std(axis=1, ddof=1).tolist(),
                "total_mean": float(per_step.sum(axis=0).mean()),
            }
            print(name, "total recombinations per cluster", fx[name]["total_mean"], flush=True)
    fx["_about"] = ("event-mode oracle (oracle/event_mc.py) ensembles binned onto the fixed-dt "
                    f"grid; seed {SEED}; made by tests/golden/make_ensemble_fixtures.py")
    OUT.write_text(json.dumps(fx, indent=1) + "\n")
    print("wrote", OUT)


if __name__ == "__main__":
    main()
