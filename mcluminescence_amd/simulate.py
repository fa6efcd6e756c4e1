"""The Monte Carlo driver: the drop-in for the reference's src/class/simulate.py.

``simulate(cfg)`` keeps the reference's contract (simulate.py:11-112):

  * runs come from ``initialize_runs(cfg)`` (zip expansion of list-valued keys);
  * every run has ``exp_type_fp.sims`` clusters (the reference uses the first
    run's ``sims`` and ``steps`` for all, simulate.py:24-27);
  * it returns ``(x_ax, Lum, electron_ratio, configs)`` with arrays of shape
    ``[steps, sims, n_runs]`` (float64), zero-padded after each run ends;
  * it writes ``results/simulations/exp_{tag}.csv`` with columns
    ``run,sim,step,lum,electron_ratio``, run-major then sim then step.

The reference's CSV has no time axis (x_ax is only returned in memory,
simulate.py:112), so a side-car ``exp_{tag}_x_ax.csv`` with columns
``run,sim,step,t`` in the same row order is written beside it;
``read_results(tag)`` rebuilds (x_ax, Lum, electron_ratio) from the pair.

What runs underneath is one GPU launch for every (run, sim) cluster, in one of
two modes (config key ``mode``, default ``fixed_dt``):

  fixed_dt  the fixed-dt step of the kernels (several recombinations per step):
            ``Lum`` is the number of recombinations in each step, ``x_ax`` the end
            time of each step, ``electron_ratio`` the trapped fraction after it.  The
            step is the reference's temperature-step cap ``max_dt / T_rate`` for
            ramps (simulate.py:44-45) and ``dt`` (or ``max_dt``) for isothermal runs.
  event     the reference's own loop, one event per step (``Lum`` in {0, 1}, the
            step cap ``max_dt / T_rate``, ``x_ax`` the event times), on the event
            kernel.

A run that needs more than ``steps`` steps raises IndexError, as the reference
does when its step buffer overflows.

Command line, like ``python src/class/simulate.py [overrides]``:

    python -m mcluminescence_amd.simulate +tag=TLwithDose exp_type_fp.e_ratio_start=0.05
    python -m mcluminescence_amd.simulate +tag=TL12event +mode=event
"""
from __future__ import annotations

import argparse
import sys
from pathlib import Path
from typing import Dict, Optional, Sequence, Tuple

import numpy as np

from . import paths
from ._abi import DEFAULT_SEED, MCL_EOVERFLOW, Context, MclError
from .config import Config, compose
from .engine import initialize_runs, run_params, schedule

_CTX: Optional[Context] = None


def default_context() -> Context:
    """Process-wide libmcl context on HIP device 0 (created on first use)."""
    global _CTX
    if _CTX is None:
        _CTX = Context(0)
    return _CTX


def run_sweep(configs: Dict[int, Config], steps: int, sims: int, *, ctx: Optional[Context] = None,
              seed: int = DEFAULT_SEED):
    """One kernel launch over every run x sim; returns x_ax, Lum, electron_ratio
    ([steps, sims, n_runs] float64) and the per-run schedules."""
    ctx = ctx or default_context()
    n_runs = len(configs)
    scheds = [schedule(configs[r]) for r in range(n_runs)]
    for r, sch in enumerate(scheds):
        if sch["n_steps"] > steps:
            raise IndexError(f"run {r} needs {sch['n_steps']} steps of {sch['dt']:g} s to cover "
                             f"{sch['duration']:g} s but exp_type_fp.steps = {steps}")
    params = [run_params(configs[r], sims=sims) for r in range(n_runs)]
    counts, _, _ = ctx.run_fixed_dt(params, seed=seed, n_bins=0)
    x_ax = np.zeros((steps, sims, n_runs))
    lum = np.zeros((steps, sims, n_runs))
    e_ratio = np.zeros((steps, sims, n_runs))
    for r in range(n_runs):
        p = params[r]
        n = int(p.n_steps)
        n0 = int(p.N_e * p.e_ratio_start)
        c = counts[r * sims:(r + 1) * sims, :n].astype(np.float64)      # [sims, n]
        left = n0 - np.cumsum(c, axis=1)
        x_ax[:n, :, r] = (np.arange(1, n + 1, dtype=np.float64) * p.dt)[:, None]
        lum[:n, :, r] = c.T
        e_ratio[:n, :, r] = (left / float(p.N_e)).T
    return x_ax, lum, e_ratio, scheds


def run_sweep_event(configs: Dict[int, Config], steps: int, sims: int, *,
                    ctx: Optional[Context] = None, seed: int = DEFAULT_SEED):
    """The reference's event loop for every run x sim (simulate.py:36-92), one
    kernel launch; returns x_ax, Lum, electron_ratio ([steps, sims, n_runs])."""
    ctx = ctx or default_context()
    n_runs = len(configs)
    params, durations = [], []
    for r in range(n_runs):
        mc = configs[r].exp_type_fp
        max_dt = float(mc.get("max_dt", 0.5))                 # simulate.py:44
        params.append(run_params(configs[r], sims=sims, n_steps=1, dt=max_dt))
        durations.append(float(mc.get("duration", 0.0)))
    try:
        lum, t, er, used, _ = ctx.run_event(params, durations, max_steps=steps, seed=seed)
    except MclError as e:
        if e.code == MCL_EOVERFLOW:
            raise IndexError(f"a run needs more than exp_type_fp.steps = {steps} steps") from e
        raise
    shape = (n_runs, sims, steps)
    x_ax = t.reshape(shape).transpose(2, 1, 0).copy()
    lm = lum.reshape(shape).transpose(2, 1, 0).astype(np.float64)
    e_ratio = er.reshape(shape).transpose(2, 1, 0).copy()
    return x_ax, lm, e_ratio, used.reshape(n_runs, sims)


def results_frame(lum: np.ndarray, e_ratio: np.ndarray):
    """The reference's CSV layout (simulate.py:95-110): run-major, then sim, then step."""
    import pandas as pd
    steps, sims, runs = lum.shape
    return pd.DataFrame({
        "run": np.repeat(np.arange(runs), sims * steps),
        "sim": np.tile(np.repeat(np.arange(sims), steps), runs),
        "step": np.tile(np.arange(steps), runs * sims),
        "lum": lum.transpose(2, 1, 0).ravel(),
        "electron_ratio": e_ratio.transpose(2, 1, 0).ravel(),
    })


def write_results_csv(path: Path, lum: np.ndarray, e_ratio: np.ndarray) -> Path:
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    results_frame(lum, e_ratio).to_csv(path, index=False)
    return path


def x_ax_path(results_csv: Path) -> Path:
    """The time-axis side-car of results/simulations/exp_{tag}.csv."""
    results_csv = Path(results_csv)
    return results_csv.with_name(results_csv.stem + "_x_ax.csv")


def write_x_ax_csv(path: Path, x_ax: np.ndarray) -> Path:
    """Columns run,sim,step,t, the row order of the results CSV."""
    import pandas as pd
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    steps, sims, runs = x_ax.shape
    pd.DataFrame({
        "run": np.repeat(np.arange(runs), sims * steps),
        "sim": np.tile(np.repeat(np.arange(sims), steps), runs),
        "step": np.tile(np.arange(steps), runs * sims),
        "t": x_ax.transpose(2, 1, 0).ravel(),
    }).to_csv(path, index=False)
    return path


def read_results(tag: str, *, results_dir: Optional[Path] = None):
    """(x_ax, Lum, electron_ratio) as [steps, sims, runs] float64 from
    exp_{tag}.csv and its side-car (x_ax is zero without one)."""
    import pandas as pd
    base = Path(results_dir) if results_dir is not None else paths.simulations_dir()
    main = base / f"exp_{tag}.csv"
    df = pd.read_csv(main, float_precision="round_trip")
    runs, sims, steps = int(df.run.max()) + 1, int(df.sim.max()) + 1, int(df.step.max()) + 1
    if len(df) != runs * sims * steps:
        raise ValueError(f"{main} does not hold a full run x sim x step grid")
    shape = (runs, sims, steps)
    lum = df.lum.to_numpy(np.float64).reshape(shape).transpose(2, 1, 0).copy()
    e_ratio = df.electron_ratio.to_numpy(np.float64).reshape(shape).transpose(2, 1, 0).copy()
    side = x_ax_path(main)
    x_ax = np.zeros_like(lum)
    if side.exists():
        tf = pd.read_csv(side, float_precision="round_trip")
        if len(tf) != len(df):
            raise ValueError(f"{side} does not match {main}")
        x_ax = tf.t.to_numpy(np.float64).reshape(shape).transpose(2, 1, 0).copy()
    return x_ax, lum, e_ratio


def simulate(cfg: Config, *, ctx: Optional[Context] = None, seed: int = DEFAULT_SEED,
             write_csv: bool = True, results_dir: Optional[Path] = None
             ) -> Tuple[np.ndarray, np.ndarray, np.ndarray, Dict[int, Config]]:
    configs = initialize_runs(cfg)
    mc0 = configs[0].exp_type_fp
    steps = int(mc0.steps)
    sims = int(mc0.sims)
    mode = str(cfg.get("mode", "fixed_dt"))
    if mode == "event":
        x_ax, lum, e_ratio, _ = run_sweep_event(configs, steps, sims, ctx=ctx, seed=seed)
    elif mode == "fixed_dt":
        x_ax, lum, e_ratio, _ = run_sweep(configs, steps, sims, ctx=ctx, seed=seed)
    else:
        raise ValueError(f"mode must be 'fixed_dt' or 'event' (got {mode!r})")
    if write_csv:
        tag = cfg.get("tag", "")
        out_dir = Path(results_dir) if results_dir is not None else paths.simulations_dir()
        main = write_results_csv(out_dir / f"exp_{tag}.csv", lum, e_ratio)
        write_x_ax_csv(x_ax_path(main), x_ax)
    return x_ax, lum, e_ratio, configs


def main(argv: Optional[Sequence[str]] = None) -> int:
    ap = argparse.ArgumentParser(description="Run the Monte Carlo sweep of a config and write "
                                             "results/simulations/exp_{tag}.csv")
    ap.add_argument("--config-dir", type=Path, default=None,
                    help="directory holding config_fp.yaml (default: the package's conf/)")
    ap.add_argument("--config-name", default="config_fp")
    ap.add_argument("--seed", type=lambda s: int(s, 0), default=DEFAULT_SEED)
    ap.add_argument("overrides", nargs="*", help="Hydra-style overrides: key=value, +key=value")
    args = ap.parse_args(argv)
    cfg = compose(args.overrides, config_dir=args.config_dir, config_name=args.config_name)
    x_ax, lum, e_ratio, configs = simulate(cfg, seed=args.seed)
    tag = cfg.get("tag", "")
    print(f"{len(configs)} runs x {lum.shape[1]} sims: {int(lum.sum())} recombinations; wrote "
          f"{paths.simulations_dir() / f'exp_{tag}.csv'}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
