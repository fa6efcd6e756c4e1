"""Global optimiser for the trap-model parameters: the drop-in for the reference's
src/class/optimizer.py (DE over ten parameters against a lab objective).

    python -m mcluminescence_amd.optimizer task=train exp=tl_clbr gens=10 \\
        exp_type_fp=TLlab physics_fp=lab_TL
    python -m mcluminescence_amd.optimizer task=replay exp=tl_clbr exp_type_fp=TLlab physics_fp=lab_TL

Same bounds, parameter order, DE settings (randtobest1bin, sobol init, mutation
0.5, recombination 0.3, tol 1e-7, polish) and result file
(results/lab_sims/result_{exp}.csv, one row param_0..param_9, mse appended per
run) as optimizer.py:32-136.  The objective is TL_lab or ISO_lab on the lab
kernel; the whole DE population is evaluated in one kernel launch (scipy's
vectorized objective: candidates x lab rows as one batch of clusters) instead of
one process per candidate.  ``sims`` clusters per lab row average the simulated
fraction (the reference uses one).
"""
from __future__ import annotations

import argparse
import sys
from pathlib import Path
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import lab, paths
from ._abi import DEFAULT_SEED, Context
from .config import Config, compose
from .engine import initialize_runs

# optimizer.py:32-43 (order and ranges kept: result files index these columns)
DEFAULT_BOUNDS: List[Tuple[float, float]] = [
    (1e-8, 1e-3),     # rho_prime
    (1.9, 2.4),       # E_cb
    (1.2, 1.7),       # E_loc_1
    (1.0, 1.5),       # E_loc_2
    (1e2, 4e2),       # D0
    (1e12, 1e14),     # s
    (1e10, 1e13),     # b
    (1e9, 5e10),      # alpha
    (1e2, 7.5e2),     # holes
    (0.0, 1.0),       # retrap
]

LAB_TABLES = {"iso": "CLBR_IR50_ISO", "tl_clbr": "CLBR_IRSL50_0.25KperGy",
              "tl_fsm-13": "FSM-13_IRSL50_0.25KperGy"}             # optimizer.py:69-78


def cfg_with_params(base: Config, params) -> Config:
    """Copy of base with the ten parameters written in (optimizer.py:49-64)."""
    cfg = base.copy()
    (rho_prime, E_cb, E_loc_1, E_loc_2, D0, s, b, alpha, holes, retrap) = (float(v) for v in params)
    cfg.exp_type_fp.rho_prime = rho_prime
    cfg.exp_type_fp.holes = holes
    cfg.physics_fp.E_cb = E_cb
    cfg.physics_fp.E_loc_1 = E_loc_1
    cfg.physics_fp.E_loc_2 = E_loc_2
    cfg.physics_fp.D0 = D0
    cfg.physics_fp.s = s
    cfg.physics_fp.b = b
    cfg.physics_fp.alpha = alpha
    cfg.physics_fp.Retrap = retrap
    return cfg


def _table(exp: str, table=None):
    if exp not in LAB_TABLES:
        raise ValueError(f"Unknown exp '{exp}'")
    return lab.read_table(LAB_TABLES[exp]) if table is None else lab._as_frame(table)


def batched_objective(cfg: Config, exp: str, population: np.ndarray, *, ctx: Context,
                      sims: int = 1, table=None, seed: int = DEFAULT_SEED,
                      exact_fill: bool = False) -> np.ndarray:
    """MSE of every candidate (rows of population, [S, 10]) in one lab-kernel launch."""
    population = np.atleast_2d(population)
    tab = _table(exp, table)
    runs, obs, targets, spans = [], [], [], []
    for p in population:
        run = initialize_runs(cfg_with_params(cfg, p))[0]
        if exp == "iso":
            r, o, tg = lab.iso_lab_runs(run, tab, sims=sims)
        else:
            r, o = lab.tl_lab_runs(run, tab, sims=sims)
            tg = [[float(f)] for f in tab["Fill"]]
        spans.append((len(runs), len(r)))
        runs += r
        obs += o
        targets += tg
    ratio, _, _ = ctx.run_lab(runs, obs, exact_fill=exact_fill, seed=seed)
    means = lab._mean_over_sims(ratio, runs)
    out = np.empty(len(population))
    for k, (first, n) in enumerate(spans):
        errs = np.concatenate([means[first + j][:len(targets[first + j])] -
                               np.asarray(targets[first + j]) for j in range(n)])
        out[k] = float(np.mean(errs ** 2))
    return out


def run_one_sim(cfg: Config, exp: str, plot: bool = False, *, ctx: Optional[Context] = None,
                sims: int = 1, table=None) -> float:
    """MSE of one parameter set and one experiment label (optimizer.py:66-79)."""
    from .tl_trap_lab import TLTrapSim
    runs = initialize_runs(cfg)
    sim = TLTrapSim(runs[0], ctx=ctx)
    if exp == "iso":
        return sim.ISO_lab(LAB_TABLES[exp], plot=plot, sims=sims, table=table)
    if exp in ("tl_clbr", "tl_fsm-13"):
        return sim.TL_lab(LAB_TABLES[exp], plot=plot, sims=sims, table=table)
    raise ValueError(f"Unknown exp '{exp}'")


def train(cfg: Config, exp: str, *, gens: int = 7, pop: int = 15, sims: int = 1,
          ctx: Optional[Context] = None, table=None, seed: Optional[int] = None,
          polish: bool = True, csv_path: Optional[Path] = None, disp: bool = True,
          replot: bool = True):
    """DE as optimizer.py:101-121, with the population batched on the GPU: appends
    the best row to results/lab_sims/result_{exp}.csv, then runs the best
    parameters once more with the lab plot written (replot=False skips the plot,
    not the run).  The result gains ``best_rerun_mse``, that run's objective."""
    from scipy.optimize import differential_evolution
    import pandas as pd
    if ctx is None:
        from .simulate import default_context
        ctx = default_context()
    tab = _table(exp, table)
    calls = [0]

    def f(x):
        calls[0] += 1
        pop_ = np.asarray(x).T if np.ndim(x) == 2 else np.asarray(x)[None]
        vals = batched_objective(cfg, exp, pop_, ctx=ctx, sims=sims, table=tab,
                                 seed=DEFAULT_SEED + 7919 * calls[0])
        return vals if np.ndim(x) == 2 else float(vals[0])

    result = differential_evolution(
        f, DEFAULT_BOUNDS, strategy="randtobest1bin", init="sobol", mutation=0.5,
        recombination=0.3, maxiter=gens, popsize=pop, tol=1e-7, disp=disp, polish=polish,
        vectorized=True, updating="deferred", seed=seed)
    csv_path = Path(csv_path) if csv_path else (paths.project_root() / "results" / "lab_sims" /
                                                f"result_{exp}.csv")
    csv_path.parent.mkdir(parents=True, exist_ok=True)
    cols = [f"param_{i}" for i in range(len(result.x))] + ["mse"]
    pd.DataFrame([list(result.x) + [float(result.fun)]], columns=cols).to_csv(
        csv_path, mode="a", index=False, header=not csv_path.exists())
    print(f"Saved run to {csv_path}")
    # quick check of the best parameters, with the lab plot (optimizer.py:119-121)
    result.best_rerun_mse = run_one_sim(cfg_with_params(cfg, result.x), exp, plot=replot, ctx=ctx,
                                        sims=sims, table=tab)
    return result


def replay(cfg: Config, exp: str, *, ctx: Optional[Context] = None, table=None,
           csv_path: Optional[Path] = None, plot: bool = True) -> float:
    """Re-run the best row of the result file (optimizer.py:125-134)."""
    import pandas as pd
    csv_path = Path(csv_path) if csv_path else (paths.project_root() / "results" / "lab_sims" /
                                                f"result_{exp}.csv")
    if not csv_path.exists():
        raise FileNotFoundError(csv_path)
    df = pd.read_csv(csv_path)
    best = df.loc[df.mse.idxmin()]
    params = best.filter(like="param_").values.astype(float)
    mse = run_one_sim(cfg_with_params(cfg, params), exp, plot=plot, ctx=ctx, table=table)
    print("Re-plotted best parameters for exp:", exp)
    return mse


def main(argv: Optional[Sequence[str]] = None) -> int:
    ap = argparse.ArgumentParser(description="DE fit of the trap parameters to lab data")
    ap.add_argument("--config-dir", type=Path, default=None)
    ap.add_argument("--config-name", default="config_fp")
    ap.add_argument("overrides", nargs="*",
                    help="task=train|replay exp=tl_clbr|tl_fsm-13|iso gens=N pop=P works=W sims=S, "
                         "plus config overrides")
    args = ap.parse_args(argv)
    opts = {"task": "train", "exp": "tl_clbr", "gens": "7", "pop": "15", "works": "4", "sims": "1"}
    rest = []
    for ov in args.overrides:
        key = ov.lstrip("+").split("=", 1)[0]
        if key in opts and "=" in ov:
            opts[key] = ov.split("=", 1)[1]
        else:
            rest.append(ov)
    cfg = compose(rest, config_dir=args.config_dir, config_name=args.config_name)
    if opts["task"] == "train":
        train(cfg, opts["exp"], gens=int(opts["gens"]), pop=int(opts["pop"]), sims=int(opts["sims"]))
    elif opts["task"] == "replay":
        replay(cfg, opts["exp"])
    else:
        raise ValueError("task must be 'train' or 'replay'")
    return 0


if __name__ == "__main__":
    sys.exit(main())
