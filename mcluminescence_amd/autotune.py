"""Autotuner for the Monte Carlo kernels' launch schedule.

The kernels' results do not depend on the tuning knobs (mcl_tune_set):

  holes_per_cell   density of each cluster's nearest-hole cell grid (search work
                   per lookup against the rate of inconclusive 3x3x3 searches)
  wave_search_max  queue length up to which a step's nearest-hole searches run
                   wave-cooperatively (four electrons at a time) rather than
                   lane-parallel
  lookahead_max    live electrons up to which a cluster tests 64 steps at once and
                   skips the quiet ones
  persist_frac     persistent workgroups per size class, as a fraction of what
                   fits on the device at once (load balance against contention
                   between concurrently running class kernels)
  waves_per_cluster  wavefronts per cluster (1, 2, 4, 8; 0 lets each launch or size
                   class pick the count that keeps the most waves resident)

so the tuner times every candidate on the workload itself (a plan launch on
device buffers, HIP-event time of the launch), keeps the fastest, and records the
result in a JSON table keyed by a signature of the workload (cluster-size
histogram, step counts, TL/isothermal mix).  ``apply`` looks a signature up and
sets the context's knobs; unknown workloads keep the defaults.
"""
from __future__ import annotations

import itertools
import json
import math
from pathlib import Path
from typing import Dict, Iterable, List, Optional, Sequence

from . import workloads as W
from ._abi import DEFAULT_SEED, F_GRAPH, F_PERSISTENT, Context, RunParams, geometry

TABLE = Path(__file__).resolve().parent / "tuning" / "table.json"
DEFAULTS = {"holes_per_cell": 2.0, "persist_frac": 1.0, "wave_search_max": 32.0,
            "lookahead_max": 8.0, "waves_per_cluster": 0.0}
SPACE = {"holes_per_cell": [2.0, 2.5, 3.0, 4.0], "persist_frac": [1.0, 2.0],
         "wave_search_max": [8.0, 32.0, 64.0], "lookahead_max": [0.0, 4.0, 8.0, 32.0],
         "waves_per_cluster": [0.0, 1.0]}


def signature(runs: Sequence[RunParams]) -> str:
    """Workload key: clusters per electron-count class (powers of two from 128),
    log2 of the mean steps, and the share of ramped (TL) runs."""
    hist: Dict[int, int] = {}
    steps = 0
    n = 0
    tl = 0
    for r in runs:
        if r.sims <= 0:
            continue
        e = geometry(r).n_e0
        cls = 0
        while (128 << cls) < e:
            cls += 1
        hist[cls] = hist.get(cls, 0) + int(r.sims)
        steps += int(r.n_steps) * int(r.sims)
        n += int(r.sims)
        tl += int(r.sims) if r.T_rate else 0
    if n == 0:
        return "empty"
    # cluster counts rounded to a power of two: nearby workloads share an entry
    parts = [f"c{k}:{2 ** round(math.log2(v))}" for k, v in sorted(hist.items())]
    return "|".join(parts) + f"|s{round(math.log2(steps / n))}|tl{round(10 * tl / n)}"


def load_table(path: Path = TABLE) -> Dict[str, Dict[str, float]]:
    if path.exists():
        return json.loads(path.read_text())
    return {}


def save_table(table: Dict[str, Dict[str, float]], path: Path = TABLE) -> None:
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_text(json.dumps(table, indent=1, sort_keys=True) + "\n")


def apply(ctx: Context, runs: Sequence[RunParams], table: Optional[Dict] = None) -> Dict[str, float]:
    """Set the context's knobs for this workload (defaults if it was never tuned)."""
    table = load_table() if table is None else table
    knobs = dict(DEFAULTS)
    knobs.update({k: v for k, v in table.get(signature(runs), {}).items() if k in DEFAULTS})
    for k, v in knobs.items():
        ctx.tune_set(k, v)
    return knobs


def time_config(ctx: Context, runs: Sequence[RunParams], knobs: Dict[str, float], n_bins: int,
                flags: int, reps: int = 3, seed: int = DEFAULT_SEED) -> float:
    """Best-of-reps device milliseconds of one plan launch with these knobs."""
    import torch
    for k, v in knobs.items():
        ctx.tune_set(k, v)
    plan = ctx.plan(runs, seed=seed, n_bins=n_bins, flags=flags)
    try:
        n, row = plan.info.total_clusters, plan.info.row_len
        counts = torch.empty((n, row), dtype=torch.int32, device="cuda")
        plan.launch(counts.data_ptr())
        plan.sync()
        best = math.inf
        for _ in range(reps):
            plan.launch(counts.data_ptr())
            plan.sync()
            best = min(best, plan.last_ms())
        return best
    finally:
        plan.close()


def autotune(ctx: Context, runs: Sequence[RunParams], n_bins: int = 0,
             flags: int = F_PERSISTENT | F_GRAPH, space: Optional[Dict[str, List[float]]] = None,
             reps: int = 3, table_path: Optional[Path] = TABLE, log=print) -> Dict[str, float]:
    """Time every point of the search space on this workload; store and apply the best."""
    space = space or SPACE
    keys = sorted(space)
    results = []
    for values in itertools.product(*(space[k] for k in keys)):
        knobs = dict(zip(keys, values))
        ms = time_config(ctx, runs, knobs, n_bins, flags, reps)
        results.append((ms, knobs))
        if log:
            log(f"autotune {knobs}: {ms:.3f} ms")
    ms, best = min(results, key=lambda t: t[0])
    entry = dict(best)
    entry["ms"] = ms
    if table_path is not None:
        table = load_table(table_path)
        table[signature(runs)] = entry
        save_table(table, table_path)
    for k, v in best.items():
        ctx.tune_set(k, v)
    return entry


# named workloads for the command line: (run list, n_bins)
WORKLOADS = {
    "uneven_sweep": (W.uneven_sweep, 256),
    "iso_decay_4096x400": (W.iso_decay_4096x400, 0),
    "tl_glow_16x500": (W.tl_glow_16x500, 4000),
}


def main(argv: Optional[Iterable[str]] = None) -> int:
    """python -m mcluminescence_amd.autotune --workload uneven_sweep [--table PATH]"""
    import argparse
    ap = argparse.ArgumentParser(description="time the launch knobs on a named workload and "
                                             "record the fastest in the tuning table")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="uneven_sweep")
    ap.add_argument("--table", type=Path, default=TABLE)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--plain", action="store_true",
                    help="tune the one-workgroup-per-cluster launch instead of the persistent, "
                         "graph-captured one")
    args = ap.parse_args(None if argv is None else list(argv))
    make, n_bins = WORKLOADS[args.workload]
    runs = make()
    flags = 0 if args.plain else F_PERSISTENT | F_GRAPH
    ctx = Context(0)
    try:
        base = time_config(ctx, runs, DEFAULTS, n_bins, flags, args.reps)
        print(f"defaults {DEFAULTS}: {base:.3f} ms", flush=True)
        entry = autotune(ctx, runs, n_bins=n_bins, flags=flags, reps=args.reps,
                         table_path=args.table, log=lambda m: print(m, flush=True))
        print(json.dumps({"workload": args.workload, "signature": signature(runs),
                          "flags": flags, "defaults_ms": base, **entry}), flush=True)
    finally:
        ctx.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
