"""Numpy restatement of the legacy lab loops of src/est_params/functions.py.

TEST INFRASTRUCTURE ONLY: the statistical checker of the lab kernel's legacy
options (MCL_LAB_LEGACY), never imported by the product.

  box ................ initialize_box_bg, functions.py:51-80 (int(holes) core holes,
                       boundary holes int(density * shell volume))
  lifetimes .......... lifetime_fading, functions.py:115-135: ONE uniform per call
                       picks E_loc_1 (U > Retrap) or E_loc_2 for every electron
  filling ............ filling_time, functions.py:138-148; add_electron and
                       recalc_distances, functions.py:150-161,211-226 (every nearest
                       hole recomputed over the full matrix afterwards)
  recombination ...... recomber and helpers, functions.py:176-261, called without
                       e_timer: the mask time >= waits + zeros((1, 1)) is 2-D, so
                       np.where(...)[0] holds row indices (all 0) and the pair
                       removed is electron 0 with its nearest hole; electrons that
                       shared that hole search again
  TL loop ............ sim_lab_TL_residuals, functions.py:270-359 (retrapping: after
                       a recombination, with probability Retrap, one more pair)
  ISO loop ........... sim_lab_TL_residuals_iso, functions.py:361-434 (no retrapping;
                       an observation is taken once t > its time)

The functions return trapped fractions, one simulation at a time, from a seeded
numpy Generator in place of the reference's global numpy RNG.
"""
from __future__ import annotations

from typing import Dict, List, Tuple

import numpy as np


def box_geometry(exp: Dict, phys: Dict, e_ratio_start: float) -> Tuple[int, int, float, float]:
    """(electrons, holes, core side d, hole side bf * d) of initialize_box_bg."""
    e = int(exp["N_e"] * e_ratio_start)
    rho = exp["rho_prime"] * (3 / (4 * np.pi) * phys["alpha"] ** 3)
    h = int(exp["holes"])
    d = (h / rho) ** (1 / 3)
    db = d * exp["boundary_factor"]
    core = d * d * d
    shell = db * db * db - core
    return e, h + int(h / (d * d * d) * shell), d, db


class LegacyBox:
    """Electrons, holes, the dense distance matrix and the nearest-hole arrays."""

    def __init__(self, exp: Dict, phys: Dict, e_ratio_start: float, rng: np.random.Generator):
        self.exp, self.phys, self.rng = exp, phys, rng
        n_e, n_h, self.d, self.db = box_geometry(exp, phys, e_ratio_start)
        self.e = rng.random((n_e, 3)) * self.d
        self.h = rng.random((n_h, 3)) * self.db
        self.dist = np.linalg.norm(self.e[:, None, :] - self.h[None, :, :], axis=2)
        self.waits = np.array([])
        self._nearest_all()

    def _nearest_all(self) -> None:
        if self.dist.shape[0] and self.dist.shape[1]:
            self.min_d = self.dist.min(axis=1)
            self.nn = self.dist.argmin(axis=1)
        else:
            self.min_d = np.zeros(self.dist.shape[0])
            self.nn = np.zeros(self.dist.shape[0], dtype=np.int64)

    def draw_waits(self, T: float) -> None:
        """lifetime_fading plus the exponential draws (one shared channel draw)."""
        if self.e.shape[0] == 0:
            self.waits = np.array([])
            return
        p = self.phys
        k_cb = p["s"] * np.exp(-p["E_cb"] / (p["k_b"] * T))
        E_loc = p["E_loc_1"] if self.rng.random() > p["Retrap"] else p["E_loc_2"]
        with np.errstate(over="ignore", divide="ignore"):
            k_tun = p["b"] * np.exp(-E_loc / (p["k_b"] * T) - p["alpha"] * self.min_d)
            self.waits = self.rng.exponential(1.0 / (k_cb + k_tun))

    def filling_time(self, D: float) -> float:
        N = int(self.exp["N_e"])
        n = self.e.shape[0]
        lifetime = 1e20 if (n == N or D == 0) else self.phys["D0"] / D * 1 / (N - n)
        return self.rng.exponential(lifetime)

    def add_pair(self) -> None:
        """add_electron + recalc_distances + min over the whole matrix."""
        ne = self.rng.random((1, 3)) * self.d
        nh = self.rng.random((1, 3)) * self.db
        self.e = np.vstack([self.e, ne])
        self.h = np.vstack([self.h, nh])
        row = np.linalg.norm(ne[0] - self.h[:-1], axis=1)
        col = np.linalg.norm(self.e - nh[0], axis=1)
        self.dist = np.hstack([np.vstack([self.dist, row]), col[:, None]])
        self._nearest_all()

    def recombine(self, t: float) -> int:
        """recomber with its default e_timer; returns the electron index removed."""
        fired = np.where(t >= self.waits + np.zeros([1, 1]))[0]
        if fired.size == 0:
            return -1
        k = int(fired[0])                      # every entry is 0 (row indices)
        gone = int(self.nn[k])
        again = np.where((self.nn == gone) & (np.arange(self.nn.size) != k))[0]
        again = again - (again > k)
        nn = self.nn.copy()
        nn[nn > gone] -= 1
        self.e = np.delete(self.e, k, 0)
        self.h = np.delete(self.h, gone, 0)
        self.dist = np.delete(np.delete(self.dist, k, 0), gone, 1)
        self.min_d = np.delete(self.min_d, k)
        self.nn = np.delete(nn, k)
        self.waits = np.delete(self.waits, k)
        if again.size and self.dist.shape[1]:
            self.min_d[again] = self.dist[again].min(axis=1)
            self.nn[again] = self.dist[again].argmin(axis=1)
        return k


def tl_fraction(exp: Dict, phys: Dict, row: Dict, rng: np.random.Generator,
                retrap: bool = True) -> float:
    """Trapped fraction at the end of one TL row (functions.py:288-343)."""
    box = LegacyBox(exp, phys, float(row["e_ratio"]), rng)
    D = float(phys["D"])
    dt_fill = box.filling_time(D)
    T = row["T_start"] + 273.15
    box.draw_waits(T)
    T_rate = (row["T_end"] - row["T_start"]) / row["Duration"]
    t = 0.0
    while t < row["Duration"]:
        dt_rec = np.min(box.waits) if box.waits.size else dt_fill
        dt = min(dt_rec, dt_fill)
        dt = max(dt, 0.0)
        T = T + dt * T_rate
        t = t + dt
        if dt == dt_fill:
            box.add_pair()
        else:
            box.recombine(t)
            if rng.random() < phys["Retrap"] and retrap:
                box.add_pair()
        box.draw_waits(T)
        dt_fill = box.filling_time(D)
    return box.e.shape[0] / exp["N_e"]


def iso_fractions(exp: Dict, phys: Dict, ex: Dict, rng: np.random.Generator) -> np.ndarray:
    """Trapped fraction at each observation time of one isothermal experiment
    (functions.py:376-421); ex = {e_ratio, temp, dose, times}."""
    box = LegacyBox(exp, phys, float(ex["e_ratio"]), rng)
    D = float(ex["dose"])
    dt_fill = box.filling_time(D)
    T = ex["temp"] + 273.15
    box.draw_waits(T)
    times = list(np.asarray(ex["times"], dtype=np.float64))
    out: List[float] = []
    duration = max(times)
    t = 0.0
    while t < duration and times:
        dt_rec = np.min(box.waits) if box.waits.size else dt_fill
        dt = min(dt_rec, dt_fill)
        if dt == dt_fill:
            t = t + dt
            box.add_pair()
            dt_fill = box.filling_time(D)
            box.draw_waits(T)
        else:
            t = t + max(dt, 0.0)
            box.recombine(t)
            dt_fill = box.filling_time(D)
            box.draw_waits(T)
        while times and t > times[0]:
            out.append(box.e.shape[0] / exp["N_e"])
            times.pop(0)
    return np.array(out)
