"""Lab-comparison objectives: the reference's TLTrapSim.TL_lab and ISO_lab
(src/class/tl_trap_lab.py:65-179) on the lab kernel (mcl_run_lab).

TL_lab: one experiment per row of data/processed/{csv}.csv (Duration, T_start,
T_end [degC], Fill, e_ratio): start at e_ratio, ramp at (T_end - T_start) /
Duration while dosing at the physics' D, and compare n_e / N_e at the end of the
ramp with Fill.  ISO_lab: one experiment per exp_no (temp [degC], dose [Gy/s],
the first row's e_ratio as the start) compared with e_ratio at every listed time.
Both return the mean squared error, as the reference does.

``sims`` clusters per experiment average the simulated fraction (the reference
runs one); ``exact_fill`` switches from the reference's nearest-hole update on a
fill (the new electron ignores the new hole, no other electron looks at it,
engine.py:141-152) to an exact one.
"""
from __future__ import annotations

from pathlib import Path
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import paths
from ._abi import DEFAULT_SEED, RunParams
from .engine import run_params


def read_table(csv_name: str, root: Optional[Path] = None):
    """data/processed/{csv_name}.csv under the project root (tl_trap_lab.py:67-68)."""
    import pandas as pd
    base = Path(root) if root is not None else paths.data_dir()
    return pd.read_csv(base / "processed" / f"{csv_name}.csv")


def _as_frame(table):
    import pandas as pd
    return table if isinstance(table, pd.DataFrame) else pd.DataFrame(table)


def tl_lab_runs(cfg, table, *, sims: int = 1) -> (List[RunParams], List[List[float]]):
    """One run per TL row: e_ratio start, linear T from T_start to T_end over Duration,
    the physics' dose rate D (tl_trap_lab.py:77-85)."""
    table = _as_frame(table)
    dose = float(cfg.physics_fp.get("D", 0.0))
    runs, obs = [], []
    for _, row in table.iterrows():
        p = run_params(cfg, sims=sims, e_ratio_start=float(row["e_ratio"]), dose_rate=dose,
                       n_steps=1)
        p.T_start = float(row["T_start"])
        p.T_rate = (float(row["T_end"]) - float(row["T_start"])) / float(row["Duration"])
        runs.append(p)
        obs.append([float(row["Duration"])])
    return runs, obs


def iso_lab_runs(cfg, table, *, sims: int = 1):
    """One run per exp_no: isothermal at temp, dose rate dose, starting from the first
    row's e_ratio; observations at the listed times (tl_trap_lab.py:132-141).
    Returns (runs, observation lists, target fractions per run)."""
    table = _as_frame(table)
    runs, obs, targets = [], [], []
    for exp_no in sorted(table["exp_no"].unique()):
        sub = table[table["exp_no"] == exp_no]
        p = run_params(cfg, sims=sims, e_ratio_start=float(sub["e_ratio"].iloc[0]),
                       dose_rate=float(sub["dose"].iloc[0]), n_steps=1)
        p.T_start = float(sub["temp"].iloc[0])
        p.T_rate = 0.0
        times = sub["time"].to_numpy(dtype=np.float64)
        runs.append(p)
        obs.append(list(times))
        # the target at each time: the first row with that time (tl_trap_lab.py:167)
        targets.append([float(sub[sub["time"] == t]["e_ratio"].iloc[0]) for t in times])
    return runs, obs, targets


def _mean_over_sims(ratio: np.ndarray, runs: Sequence[RunParams]) -> List[np.ndarray]:
    out, c = [], 0
    for r in runs:
        out.append(ratio[c:c + r.sims].mean(axis=0))
        c += r.sims
    return out


def _printer(mean_abs_err: float, mse: float) -> None:
    print(f"mean |error| = {mean_abs_err:.4f}, MSE = {mse:.6f}")


def tl_lab(sim, csv_name: str, *, plot: bool = False, sims: int = 1, exact_fill: bool = False,
           table=None, seed: Optional[int] = None, verbose: bool = True) -> float:
    """MSE of the final trapped fraction against Fill over the rows of a TL table."""
    table = read_table(csv_name) if table is None else _as_frame(table)
    runs, obs = tl_lab_runs(sim.cfg, table, sims=sims)
    ratio, _, _ = sim._context().run_lab(runs, obs, exact_fill=exact_fill,
                                         seed=sim.seed if seed is None else seed)
    final = np.array([m[0] for m in _mean_over_sims(ratio, runs)])
    err = final - table["Fill"].to_numpy(dtype=np.float64)
    mse = float(np.mean(err ** 2))
    if verbose:
        _printer(float(np.mean(np.abs(err))), mse)
    if plot:
        _plot(table["Duration"].to_numpy(), table["Fill"].to_numpy(), final, csv_name)
    return mse


def iso_lab(sim, csv_name: str, *, plot: bool = False, sims: int = 1, exact_fill: bool = False,
            table=None, seed: Optional[int] = None, verbose: bool = True) -> float:
    """MSE of the trapped fraction against e_ratio at every observation of every
    isothermal experiment."""
    table = read_table(csv_name) if table is None else _as_frame(table)
    runs, obs, targets = iso_lab_runs(sim.cfg, table, sims=sims)
    ratio, _, _ = sim._context().run_lab(runs, obs, exact_fill=exact_fill,
                                         seed=sim.seed if seed is None else seed)
    means = _mean_over_sims(ratio, runs)
    errs = np.concatenate([m[:len(tg)] - np.asarray(tg) for m, tg in zip(means, targets)])
    mse = float(np.mean(errs ** 2))
    if verbose:
        _printer(float(np.mean(np.abs(errs))), mse)
    if plot:
        t = np.concatenate(obs)
        _plot(t, np.concatenate(targets), np.concatenate([m[:len(o)] for m, o in zip(means, obs)]),
              csv_name)
    return mse


def _plot(x, lab, sim, name: str) -> None:
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    fig, ax = plt.subplots()
    ax.plot(x, lab, "o", label="lab")
    ax.plot(x, sim, "x", label="simulation")
    ax.set_xlabel("time [s]")
    ax.set_ylabel("n_e / N_e")
    ax.legend()
    out = paths.project_root() / "results" / "lab_sims" / f"{name}.png"
    out.parent.mkdir(parents=True, exist_ok=True)
    fig.savefig(out)
    plt.close(fig)
