"""The legacy lab objectives of src/est_params (functions.py, estimate.py) on the
lab kernel, for reproducing results fitted with that stack.

    python -m mcluminescence_amd.legacy task=train exp=tl_clbr runs=1 \\
        exp_type_fp=TLlab physics_fp=lab_TL
    python -m mcluminescence_amd.legacy task=opt exp=tl_clbr exp_type_fp=TLlab physics_fp=lab_TL

What differs from the class-based objectives (tl_trap_lab.py:65-179), each an
MCL_LAB_* option of mcl_run_lab:

  * one localisation-channel draw per step shared by all electrons
    (lifetime_fading, functions.py:115-135);
  * the recombining pair is always the oldest electron with its nearest hole
    (recomber as the lab loops call it, functions.py:241-261);
  * in the TL loop, after each recombination a new electron-hole pair with
    probability Retrap (functions.py:328-331);
  * fills recompute every nearest hole (functions.py:150-161);
  * the box holds int(holes) core holes and int(density * shell) boundary holes
    (initialize_box_bg, functions.py:51-80);
  * the MSE is taken over every (row, sim) simulation, not over sim means
    (functions.py:348-350, 422-427), and the isothermal objective runs the
    experiments 0 .. last exp_no - 1 (functions.py:370).

Every simulation of an objective (rows x sims, or experiments x sims) is one
kernel launch; the DE population of the fit is one launch per generation.
"""
from __future__ import annotations

import argparse
import sys
from pathlib import Path
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import lab, paths
from ._abi import DEFAULT_SEED, LAB_LEGACY, LAB_RETRAP, Context
from .config import Config, compose
from .engine import initialize_runs
from .optimizer import DEFAULT_BOUNDS, cfg_with_params

BOUNDS: List[Tuple[float, float]] = list(DEFAULT_BOUNDS)   # estimate.py:17-28 (same ranges)
TABLES = {"tl_clbr": "CLBR_IRSL50_0.25KperGy", "tl_fsm-13": "FSM-13_IRSL50_0.25KperGy",
          "iso": "CLBR_IR50_ISO"}                          # estimate.py:49-56

OPT_TL = LAB_LEGACY
OPT_ISO = LAB_LEGACY & ~LAB_RETRAP


def _context(ctx: Optional[Context]) -> Context:
    if ctx is not None:
        return ctx
    from .simulate import default_context
    return default_context()


def _sims(run_cfg: Config) -> int:
    return int(run_cfg.exp_type_fp.get("sims", 1))


def printer(run_cfg: Config, avg_err: float, mse: float) -> None:
    """The parameter line of functions.py:458-463."""
    p, m = run_cfg.physics_fp, run_cfg.exp_type_fp
    print(f"absError: {avg_err}, MSE: {mse} with params: rho_prime = {m.rho_prime}, "
          f"E_cb = {p.E_cb}, D0 = {p.D0}, E_loc_1 = {p.E_loc_1}, E_loc_2 = {p.E_loc_2}, "
          f"s = {p.s}, b = {p.b}, alpha = {p.alpha}, holes = {m.holes}, P_retrap = {p.Retrap}")


def _tl_setup(run_cfg: Config, table, sims: int):
    runs, obs = lab.tl_lab_runs(run_cfg, table, sims=sims)
    fill = table["Fill"].to_numpy(dtype=np.float64)
    return runs, obs, [np.full((sims, 1), f) for f in fill]


def _iso_table(table):
    """Experiments 0 .. last exp_no - 1 (functions.py:370)."""
    return table[table["exp_no"] < table["exp_no"].iloc[-1]]


def _iso_setup(run_cfg: Config, table, sims: int):
    runs, obs, targets = lab.iso_lab_runs(run_cfg, _iso_table(table), sims=sims)
    return runs, obs, [np.tile(np.asarray(t, np.float64), (sims, 1)) for t in targets]


def _errors(ratio: np.ndarray, runs, targets) -> np.ndarray:
    """Per-simulation errors of every run against its targets, concatenated."""
    out, c = [], 0
    for r, tg in zip(runs, targets):
        n = int(r.sims)
        out.append((ratio[c:c + n, :tg.shape[1]] - tg).ravel())
        c += n
    return np.concatenate(out)


def sim_lab_TL_residuals(run_cfg: Config, lab_data: str = "CLBR_IRSL50_0.25KperGy",
                         PLOT: bool = False, *, ctx: Optional[Context] = None,
                         seed: int = DEFAULT_SEED, table=None, verbose: bool = True) -> float:
    """MSE of the legacy TL loop over every row and sim (functions.py:270-359)."""
    table = lab.read_table(lab_data) if table is None else lab._as_frame(table)
    runs, obs, targets = _tl_setup(run_cfg, table, _sims(run_cfg))
    ratio, _, _ = _context(ctx).run_lab(runs, obs, seed=seed, options=OPT_TL)
    err = _errors(ratio, runs, targets)
    mse = float(np.mean(err ** 2))
    if verbose:
        printer(run_cfg, float(np.mean(np.abs(err))), mse)
    if PLOT:
        final = np.array([ratio[k * int(runs[0].sims):(k + 1) * int(runs[0].sims), 0].mean()
                          for k in range(len(runs))])
        lab._plot(table["Duration"].to_numpy(), table["Fill"].to_numpy(), final,
                  f"{lab_data}_legacy")
    return mse


def sim_lab_TL_residuals_iso(run_cfg: Config, lab_data: str = "CLBR_IR50_ISO",
                             PLOT: bool = False, *, ctx: Optional[Context] = None,
                             seed: int = DEFAULT_SEED, table=None, verbose: bool = True) -> float:
    """MSE of the legacy isothermal loop over every experiment, sim and observation
    (functions.py:361-434)."""
    table = lab.read_table(lab_data) if table is None else lab._as_frame(table)
    runs, obs, targets = _iso_setup(run_cfg, table, _sims(run_cfg))
    ratio, _, _ = _context(ctx).run_lab(runs, obs, seed=seed, options=OPT_ISO)
    err = _errors(ratio, runs, targets)
    mse = float(np.mean(err ** 2))
    if verbose:
        printer(run_cfg, float(np.mean(np.abs(err))), mse)
    if PLOT:
        t = np.concatenate(obs)
        sims = int(runs[0].sims)
        sim = np.concatenate([ratio[k * sims:(k + 1) * sims, :len(o)].mean(axis=0)
                              for k, o in enumerate(obs)])
        lab._plot(t, np.concatenate([tg[0] for tg in targets]), sim, f"{lab_data}_legacy")
    return mse


def sim_starter(cfg: Config, params, exp_type: str, PLOT: bool = False, *,
                ctx: Optional[Context] = None, table=None, seed: int = DEFAULT_SEED) -> float:
    """estimate.py:34-57: write the ten parameters in and run the chosen objective."""
    run_cfg = initialize_runs(cfg_with_params(cfg, params))[0]
    if exp_type == "iso":
        return sim_lab_TL_residuals_iso(run_cfg, PLOT=PLOT, ctx=ctx, table=table, seed=seed)
    if exp_type in ("tl_clbr", "tl_fsm-13"):
        return sim_lab_TL_residuals(run_cfg, lab_data=TABLES[exp_type], PLOT=PLOT, ctx=ctx,
                                    table=table, seed=seed)
    raise ValueError(f"Unknown experiment type: {exp_type}")


def batched_objective(cfg: Config, exp_type: str, population: np.ndarray, *, ctx: Context,
                      table=None, seed: int = DEFAULT_SEED) -> np.ndarray:
    """Legacy MSE of every candidate (rows of population, [P, 10]) in one launch."""
    if exp_type not in TABLES:
        raise ValueError(f"Unknown experiment type: {exp_type}")
    tab = lab.read_table(TABLES[exp_type]) if table is None else lab._as_frame(table)
    runs, obs, targets, spans = [], [], [], []
    for p in np.atleast_2d(population):
        run_cfg = initialize_runs(cfg_with_params(cfg, p))[0]
        setup = _iso_setup if exp_type == "iso" else _tl_setup
        r, o, tg = setup(run_cfg, tab, _sims(run_cfg))
        spans.append((len(runs), len(r)))
        runs += r
        obs += o
        targets += tg
    opts = OPT_ISO if exp_type == "iso" else OPT_TL
    ratio, _, _ = ctx.run_lab(runs, obs, seed=seed, options=opts)
    firsts = np.cumsum([0] + [int(r.sims) for r in runs])
    out = np.empty(len(spans))
    for k, (first, n) in enumerate(spans):
        lo, hi = firsts[first], firsts[first + n]
        err = _errors(ratio[lo:hi], runs[first:first + n], targets[first:first + n])
        out[k] = float(np.mean(err ** 2))
    return out


def result_path(exp_type: str) -> Path:
    return paths.project_root() / "results" / "lab_sims" / f"result_{exp_type}.csv"


def train(cfg: Config, exp_type: str, *, runs: int = 1, gens: int = 7, pop: int = 15,
          ctx: Optional[Context] = None, table=None, seed: Optional[int] = None,
          polish: bool = True, csv_path: Optional[Path] = None, disp: bool = True):
    """estimate.py:62-87: `runs` DE fits (randtobest1bin, sobol, mutation 0.5,
    recombination 0.3, maxiter 7, popsize 15, tol 1e-7, polish), each result row
    put in front of the result file's rows."""
    from scipy.optimize import differential_evolution
    import pandas as pd
    ctx = _context(ctx)
    calls = [0]

    def f(x):
        calls[0] += 1
        pop_ = np.asarray(x).T if np.ndim(x) == 2 else np.asarray(x)[None]
        vals = batched_objective(cfg, exp_type, pop_, ctx=ctx, table=table,
                                 seed=DEFAULT_SEED + 104729 * calls[0])
        return vals if np.ndim(x) == 2 else float(vals[0])

    path = Path(csv_path) if csv_path else result_path(exp_type)
    results = []
    for _ in range(runs):
        res = differential_evolution(
            f, BOUNDS, strategy="randtobest1bin", init="sobol", mutation=0.5, recombination=0.3,
            maxiter=gens, popsize=pop, tol=1e-7, disp=disp, polish=polish, vectorized=True,
            updating="deferred", seed=seed)
        cols = [f"param_{i}" for i in range(len(res.x))] + ["mse"]
        df = pd.DataFrame([list(res.x) + [float(res.fun)]], columns=cols)
        path.parent.mkdir(parents=True, exist_ok=True)
        if path.exists():
            df = pd.concat([df, pd.read_csv(path)])
        df.to_csv(path, index=False)
        print(f"Saved CSV to {path}")
        results.append(res)
    return results


def replay(cfg: Config, exp_type: str, *, ctx: Optional[Context] = None, table=None,
           csv_path: Optional[Path] = None, plot: bool = True) -> float:
    """estimate.py:88-94: re-run the lowest-MSE row of the result file."""
    import pandas as pd
    path = Path(csv_path) if csv_path else result_path(exp_type)
    df = pd.read_csv(path)
    best = df[df["mse"] == df.mse.min()].iloc[0]
    params = best.iloc[:-1].to_numpy(dtype=np.float64)
    return sim_starter(cfg, params, exp_type, PLOT=plot, ctx=ctx, table=table)


def main(argv: Optional[Sequence[str]] = None) -> int:
    ap = argparse.ArgumentParser(description="Legacy DE fit of the trap parameters (estimate.py)")
    ap.add_argument("--config-dir", type=Path, default=None)
    ap.add_argument("--config-name", default="config_fp")
    ap.add_argument("overrides", nargs="*",
                    help="task=train|opt exp=tl_clbr|tl_fsm-13|iso runs=R gens=N pop=P, plus "
                         "config overrides")
    args = ap.parse_args(argv)
    opts = {"task": "train", "exp": "tl_clbr", "runs": "15", "gens": "7", "pop": "15"}
    rest = []
    for ov in args.overrides:
        key = ov.lstrip("+").split("=", 1)[0]
        if key in opts and "=" in ov:
            opts[key] = ov.split("=", 1)[1]
        else:
            rest.append(ov)
    cfg = compose(rest, config_dir=args.config_dir, config_name=args.config_name)
    if opts["task"] == "train":
        train(cfg, opts["exp"], runs=int(opts["runs"]), gens=int(opts["gens"]),
              pop=int(opts["pop"]))
    elif opts["task"] == "opt":
        replay(cfg, opts["exp"])
    else:
        raise ValueError("task must be 'train' or 'opt'")
    return 0


if __name__ == "__main__":
    sys.exit(main())
