"""Plots of a sweep: the drop-in for the reference's src/class/plots.py.

    python -m mcluminescence_amd.plots +plots=TL,iso +tag=X [config overrides]
    python -m mcluminescence_amd.plots +plots=TL_norm +tag=X +from_csv=true

Same registry keys (TL, TL_norm, iso, bg), plotter signature
``(cfg, ax, x_ax, lum, e_ratio, configs)`` and output files
results/plots/{key}_run_{tag}.png (plots.py:28-37,138-171).  The sweep comes
from ``simulate(cfg)``, or with ``from_csv`` from the results the driver wrote:
exp_{tag}.csv plus its time-axis side-car exp_{tag}_x_ax.csv.

The glow curves of TL / TL_norm (plots.py:39-66: the first sim's time axis, the
mean over sims of Lum, 1-degree bins, a 50-degree boxcar) are computed on the
GPU by mcl_glow_curve for every run in one launch; matplotlib only draws.
plot_iso takes the configs dict as its last argument (the reference's registry
passes it, plots.py:167, while its plot_iso signature omits it and indexes the
root config, plots.py:103-108).
"""
from __future__ import annotations

import argparse
import sys
from pathlib import Path
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np

from . import paths
from ._abi import Context
from .config import Config, compose
from .engine import initialize_runs, run_params

BIN_WIDTH = 1.0      # degC per histogram bin (plots.py:39)
WIN_DEG = 50.0       # boxcar width in degC (plots.py:39)


def running_mean(a: np.ndarray, k: int = 5) -> np.ndarray:
    """Boxcar mean, 'valid' mode (plots.py:20-22); smooths the iso curves."""
    return np.convolve(a, np.ones(k) / k, "valid")


def colour_cycle(n: int):
    import matplotlib.pyplot as plt
    return iter(plt.cm.rainbow(np.linspace(0, 1, n)))


def out(name: str, subfolder: str, tag: Optional[str] = None) -> Path:
    """results/<subfolder>/<name>[_<tag>].png (plots.py:27-37)."""
    base = paths.project_root() / "results" / subfolder
    base.mkdir(parents=True, exist_ok=True)
    return base / (f"{name}_{tag}.png" if tag else f"{name}.png")


def _context(ctx: Optional[Context]) -> Context:
    if ctx is not None:
        return ctx
    from .simulate import default_context
    return default_context()


def glow_curves(x_ax: np.ndarray, lum: np.ndarray, configs: Dict[int, Config], *,
                bin_width: float = BIN_WIDTH, win_deg: float = WIN_DEG,
                ctx: Optional[Context] = None) -> List[np.ndarray]:
    """TL glow curve of every run (plots.py:55-64) from the sweep's arrays
    ([steps, sims, runs]): one mcl_glow_curve launch.  Lum must hold counts
    (whole numbers, as both Monte Carlo modes write)."""
    steps, sims, n_runs = lum.shape
    if x_ax.shape != lum.shape:
        raise ValueError("x_ax and lum must have the same shape")
    if len(configs) != n_runs:
        raise ValueError(f"{len(configs)} configs for {n_runs} runs")
    counts = lum.transpose(2, 1, 0).reshape(n_runs * sims, steps)
    if counts.size and (counts.min() < 0 or counts.max() > np.iinfo(np.uint32).max
                        or np.any(counts != np.floor(counts))):
        raise ValueError("lum must hold recombination counts (whole numbers >= 0)")
    counts = counts.astype(np.uint32)
    t_axis = x_ax.transpose(2, 1, 0).reshape(n_runs * sims, steps)
    runs = [run_params(configs[r], sims=sims, n_steps=max(steps, 1)) for r in range(n_runs)]
    curves, _ = _context(ctx).glow_curve(runs, counts, t_axis, bin_width=bin_width, win=win_deg)
    return curves


def _ramp(configs: Dict[int, Config], r: int):
    mc = configs[r].exp_type_fp
    return float(mc.T_rate), float(mc.T_start)


def plot_TL(cfg: Config, ax, x_ax, lum, e_ratio, configs, *, ctx: Optional[Context] = None):
    """Glow curve per run against temperature (plots.py:50-73)."""
    curves = glow_curves(x_ax, lum, configs, ctx=ctx)
    colours = colour_cycle(len(curves))
    for r, smooth in enumerate(curves):
        T_rate, _ = _ramp(configs, r)
        ax.plot(np.arange(len(smooth)), smooth, color=next(colours), label=f"{T_rate} °C/s")
    ax.set_xlim(0, _max_T(cfg))
    ax.set_xlabel("Temperature (°C)")
    ax.set_ylabel("TL intensity (per °C)")
    ax.legend()


def plot_TL_normalized(cfg: Config, ax, x_ax, lum, e_ratio, configs, *,
                       ctx: Optional[Context] = None):
    """plot_TL with every curve divided by its sum (plots.py:76-100)."""
    curves = glow_curves(x_ax, lum, configs, ctx=ctx)
    colours = colour_cycle(len(curves))
    for r, smooth in enumerate(curves):
        T_rate, _ = _ramp(configs, r)
        total = smooth.sum()
        ax.plot(np.arange(len(smooth)), smooth / total if total else smooth,
                color=next(colours), label=f"{T_rate} °C/s")
    ax.set_xlim(0, _max_T(cfg))
    ax.set_xlabel("Temperature (°C)")
    ax.set_ylabel("TL intensity (per °C)")
    ax.legend()


def plot_iso(cfg: Config, ax, x_ax, lum, e_ratio, configs=None):
    """Filled fraction against log time, 10-step running means (plots.py:102-114)."""
    configs = configs if configs is not None else initialize_runs(cfg)
    colours = colour_cycle(e_ratio.shape[-1])
    for r in range(e_ratio.shape[-1]):
        t_start = configs[r].exp_type_fp.T_start
        f = running_mean(e_ratio[:, :, r].mean(axis=1), k=10)
        t = running_mean(x_ax[:, 0, r], k=10)
        ax.plot(t, f, color=next(colours), lw=0.8, label=f"Temp = {t_start} °C")
    ax.set_xscale("log")
    ax.set_xlabel("Time (s)")
    ax.set_ylabel("Filled fraction")
    ax.legend()


def plot_bg(cfg: Config, ax, x_ax, lum, e_ratio, configs):
    """Filled fraction against time per dose rate (plots.py:117-131)."""
    colours = colour_cycle(e_ratio.shape[-1])
    for r in range(e_ratio.shape[-1]):
        D_rate = configs[r].physics_fp.get("D", 0.0)
        ax.plot(x_ax[:, 0, r], e_ratio[:, :, r].mean(axis=1), color=next(colours), lw=0.8,
                label=f"D={D_rate} Gy/s")
    ax.set_xlabel("Time (s)")
    ax.set_ylabel("Filled fraction")
    ax.set_ylim(0, 1)
    ax.legend()


def _max_T(cfg: Config) -> float:
    """Upper x limit of the TL plots: the first T_end of the config (plots.py:55)."""
    T_end = cfg.exp_type_fp.get("T_end", None)
    if T_end is None:
        return 800.0
    return float(T_end[0] if isinstance(T_end, (list, tuple)) else T_end)


PLOTTERS: Dict[str, Callable] = {
    "TL": plot_TL,
    "TL_norm": plot_TL_normalized,
    "iso": plot_iso,
    "bg": plot_bg,
}


def plot_all(cfg: Config, *, from_csv: bool = False, ctx: Optional[Context] = None,
             results_dir: Optional[Path] = None) -> List[Path]:
    """Simulate (or read the saved sweep) and draw every requested plotter
    (plots.py:149-171); returns the files written."""
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    from .simulate import read_results, simulate
    tag = cfg.get("tag", "")
    if from_csv:
        x_ax, lum, e_ratio = read_results(tag, results_dir=results_dir)
        configs = initialize_runs(cfg)
    else:
        x_ax, lum, e_ratio, configs = simulate(cfg, ctx=ctx, results_dir=results_dir)
    written = []
    for key in str(cfg.get("plots", "TL_norm")).split(","):
        if key not in PLOTTERS:
            print(f"[plot_runs] unknown key '{key}', skipping")
            continue
        fig, ax = plt.subplots(figsize=(6, 4))
        if key in ("TL", "TL_norm"):
            PLOTTERS[key](cfg, ax, x_ax, lum, e_ratio, configs, ctx=ctx)
        else:
            PLOTTERS[key](cfg, ax, x_ax, lum, e_ratio, configs)
        fig.tight_layout()
        path = out(f"{key}_run", "plots", tag)
        fig.savefig(path)
        plt.close(fig)
        written.append(path)
        print(f"[plot_runs] wrote {path.name}")
    return written


def main(argv: Optional[Sequence[str]] = None) -> int:
    ap = argparse.ArgumentParser(description="Simulate a sweep and plot it (TL, TL_norm, iso, bg)")
    ap.add_argument("--config-dir", type=Path, default=None)
    ap.add_argument("--config-name", default="config_fp")
    ap.add_argument("overrides", nargs="*",
                    help="+plots=TL,iso +tag=X +from_csv=true, plus config overrides")
    args = ap.parse_args(argv)
    cfg = compose(args.overrides, config_dir=args.config_dir, config_name=args.config_name)
    from_csv = str(cfg.get("from_csv", "false")).lower() in ("1", "true", "yes")
    plot_all(cfg, from_csv=from_csv)
    return 0


if __name__ == "__main__":
    sys.exit(main())
