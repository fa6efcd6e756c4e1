"""Numpy restatement of the reference's glow-curve binning (src/class/plots.py).

TEST INFRASTRUCTURE ONLY: the checker of the GPU glow-curve kernel
(mcl_glow_curve), never imported by the product.

  running_mean     plots.py:20-22   np.convolve(a, ones(k) / k, 'valid')
  hist_and_smooth  plots.py:39-48   weighted histogram over np.arange(0, max + w, w)
                                    edges, / w, boxcar of max(1, int(win / w)) bins
  tl_curves        plots.py:50-66   per run: the first sim's time axis, the mean
                                    over sims of Lum, the t > 0 mask, T = T_start +
                                    T_rate t, then hist_and_smooth

Pinned in tests/test_glow_oracle.py against a loop-by-loop brute force of the
histogram rule (edges j * w, half-open bins, the last one closed), the closed
forms of constant and single-step inputs, and conservation of the total weight.
"""
from __future__ import annotations

from typing import List, Sequence, Tuple

import numpy as np


def running_mean(a: np.ndarray, k: int = 5) -> np.ndarray:
    """plots.py:20-22."""
    kernel = np.ones(k) / k
    return np.convolve(a, kernel, "valid")


def hist_and_smooth(t_axis: np.ndarray, events: np.ndarray, bin_width: float = 1.0,
                    win_deg: float = 50.0) -> np.ndarray:
    """plots.py:39-48 (t_axis is the temperature axis there)."""
    bins = np.arange(0, t_axis.max() + bin_width, bin_width)
    hist, _ = np.histogram(t_axis, bins=bins, weights=events)
    hist = hist / bin_width
    k = max(1, int(win_deg / bin_width))
    return running_mean(hist, k=k)


def tl_curves(x_ax: np.ndarray, lum: np.ndarray, ramps: Sequence[Tuple[float, float]],
              bin_width: float = 1.0, win_deg: float = 50.0) -> List[np.ndarray]:
    """The per-run data path of plot_TL (plots.py:55-64): x_ax and lum are
    [steps, sims, runs]; ramps[r] = (T_start, T_rate)."""
    out = []
    for r, (T_start, T_rate) in enumerate(ramps):
        t = x_ax[:, 0, r]
        L = lum[:, :, r].mean(axis=1)
        m = t > 0
        T_axis = T_start + T_rate * t[m]
        out.append(hist_and_smooth(T_axis, L[m], bin_width, win_deg))
    return out
