"""Event-mode restatement of the reference's Monte Carlo loop, in numpy.

TEST INFRASTRUCTURE ONLY: the ensemble oracle for the HIP kernels and the
CPU-baseline leg of bench.py.  Never imported by the product.

It follows the reference step for step and keeps its cost structure (a dense
float64 electron x hole distance matrix, np.delete per recombination, every
electron's wait redrawn every step):

  Physics.rate_cb / rate_tunnel / lifetime ....... src/class/engine.py:65-77
  Box._rebuild / seed / add_electron / remove_pair  src/class/engine.py:113-175
  TLTrapSim.__init__ / _update_lifetimes /
      _filling_time .............................. src/class/tl_trap_lab.py:27-60
  simulate() run x sim x step loop ............... src/class/simulate.py:36-92

including its quirks: dose filling reads ``phys["dose_rate"]`` (never set by the
shipped YAML, so lambda = 1e-20 and one Exp(1e20) draw per step); an empty
trap population takes dt = dt_fill and fires one spurious add_electron before
the loop ends; add_electron leaves existing nearest holes stale.  The only
change is the random source: a seeded numpy Generator instead of the
reference's unseeded global numpy RNG.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Dict, Tuple

import numpy as np


@dataclass(frozen=True)
class Physics:
    """engine.py:44-77."""
    alpha: float
    b: float
    s: float
    E_cb: float
    E_loc_1: float
    E_loc_2: float
    D0: float
    Retrap: float
    k_b: float = 8.617333262145e-5
    D: float = 0.0

    def rate_cb(self, T: float) -> float:
        return self.s * np.exp(-self.E_cb / (self.k_b * T))

    def rate_tunnel(self, d: np.ndarray, T: float, rng: np.random.Generator) -> np.ndarray:
        E_loc = np.where(rng.random(d.shape) < self.Retrap, self.E_loc_2, self.E_loc_1)
        return self.b * np.exp(-E_loc / (self.k_b * T) - self.alpha * d)

    def lifetime(self, d: np.ndarray, T: float, rng: np.random.Generator) -> np.ndarray:
        return 1.0 / (self.rate_cb(T) + self.rate_tunnel(d, T, rng))


class Box:
    """engine.py:83-188: coordinates plus a lazy dense distance matrix."""

    def __init__(self, side: float, bf: float, rng: np.random.Generator):
        self.core = np.array([side, side, side])
        self.boundary = self.core * bf
        self.bf = bf
        self.rng = rng
        self.electrons = np.zeros((0, 3))
        self.holes = np.zeros((0, 3))
        self._d = None
        self._min_d = None
        self._nearest = None

    def _rebuild(self) -> None:
        self._d = np.linalg.norm(self.electrons[:, None, :] - self.holes[None, :, :], axis=2)
        self._min_d = np.min(self._d, axis=1) if self._d.shape[1] else np.full(len(self._d), np.inf)
        self._nearest = (np.argmin(self._d, axis=1) if self._d.shape[1]
                         else np.zeros(len(self._d), np.int64))

    def seed(self, n_e: int, n_h: int) -> None:
        self.electrons = self.rng.random((n_e, 3)) * self.core
        total_h = int(n_h * self.bf ** 3)
        self.holes = self.rng.random((total_h, 3)) * self.boundary
        self._d = None
        self._min_d = None

    def add_electron(self) -> None:
        new_e = self.rng.random((1, 3)) * self.core
        new_h = self.rng.random((1, 3)) * self.boundary
        old_holes = self.holes.copy()
        self.electrons = np.vstack([self.electrons, new_e])
        self.holes = np.vstack([self.holes, new_h])
        if self._d is None:
            self._rebuild()
            return
        new_row = np.linalg.norm(new_e - old_holes, axis=1)
        new_col = np.linalg.norm(self.electrons - new_h, axis=1)[:, None]
        self._d = np.vstack([self._d, new_row])
        self._d = np.hstack([self._d, new_col])
        self._min_d = np.append(self._min_d, np.min(new_row) if new_row.size else np.inf)
        self._nearest = np.append(self._nearest, np.argmin(new_row) if new_row.size else 0)

    def remove_pair(self, e_idx: int, h_idx: int) -> None:
        self.electrons = np.delete(self.electrons, e_idx, axis=0)
        self.holes = np.delete(self.holes, h_idx, axis=0)
        if self._d is None:
            return
        self._d = np.delete(self._d, e_idx, axis=0)
        self._d = np.delete(self._d, h_idx, axis=1)
        self._min_d = np.delete(self._min_d, e_idx)
        self._nearest = np.delete(self._nearest, e_idx)
        self._nearest[self._nearest > h_idx] -= 1
        mask = self._nearest == h_idx
        if mask.any() and self._d.shape[1]:
            sub = self._d[mask]
            self._min_d[mask] = np.min(sub, axis=1)
            self._nearest[mask] = np.argmin(sub, axis=1)

    def nearest(self) -> Tuple[np.ndarray, np.ndarray]:
        if self._min_d is None:
            self._rebuild()
        return self._min_d, self._nearest


class Sim:
    """TLTrapSim core, tl_trap_lab.py:27-60."""

    def __init__(self, exp: Dict, phys: Dict, e_ratio_start: float, rng: np.random.Generator):
        self.mc = exp
        self.phys = Physics(**{k: phys[k] for k in ("alpha", "b", "s", "E_cb", "E_loc_1",
                                                       "E_loc_2", "D0", "Retrap")},
                            k_b=phys.get("k_b", 8.617333262145e-5), D=phys.get("D", 0.0))
        self.rng = rng
        rho = exp["rho_prime"] * (3 / (4 * np.pi) * self.phys.alpha ** 3)
        d = (exp["holes"] / rho) ** (1 / 3)
        self.side = d
        self.box = Box(d, exp["boundary_factor"], rng)
        e0 = int(exp["N_e"] * e_ratio_start)
        self.box.seed(e0, exp["holes"])
        self._recomb_wait = None

    def _update_lifetimes(self, T: float) -> None:
        d_min, _ = self.box.nearest()
        lifetime = self.phys.lifetime(d_min, T, self.rng)
        self._recomb_wait = self.rng.exponential(lifetime)

    def _filling_time(self, D: float) -> float:
        e = self.box.electrons.shape[0]
        N = int(self.mc["N_e"])
        lam = 1e-20 if (e == N or D == 0) else (D / self.phys.D0) * (N - e)
        return self.rng.exponential(1 / lam) if lam > 0 else 1e20


def simulate_run(exp: Dict, phys: Dict, sims: int, steps: int, seed: int,
                 first_sim: int = 0):
    """simulate.py:36-92 for one run: returns x_ax, lum, e_ratio of shape (steps, sims).

    Raises IndexError, like the reference, if the loop needs more than ``steps``
    steps before ``duration`` is reached (simulate.py:64)."""
    x_ax = np.zeros((steps, sims))
    lum = np.zeros((steps, sims))
    e_ratio = np.zeros((steps, sims))
    duration = float(exp.get("duration", 0.0))
    e_ratio_start = float(exp.get("e_ratio_start", 0.0))
    dose_rate = float(phys.get("dose_rate", 0.0))
    T_rate = float(exp.get("T_rate", 0.0))
    max_dt = float(exp.get("max_dt", 0.5))
    dt_cap = max_dt / T_rate if T_rate else 1e20
    T0 = exp.get("T_start", 0)
    for j in range(sims):
        rng = np.random.default_rng([seed, first_sim + j])
        sim = Sim(exp, phys, e_ratio_start, rng)
        t_cur = 0.0
        i = 0
        while t_cur <= duration:
            T_now = T0 + T_rate * t_cur + 273.15
            sim._update_lifetimes(T_now)
            dt_fill = sim._filling_time(dose_rate)
            dt_recomb = np.min(sim._recomb_wait) if sim._recomb_wait.any() else dt_fill
            dt = min(dt_fill, dt_recomb, dt_cap)
            t_cur += dt
            x_ax[i, j] = t_cur
            if dt == dt_fill:
                sim.box.add_electron()
                event = 0
            elif dt == dt_recomb:
                e_idx = int(np.argmin(sim._recomb_wait))
                _, nearest = sim.box.nearest()
                sim.box.remove_pair(e_idx, int(nearest[e_idx]))
                event = 1
            else:
                event = 0
            lum[i, j] = event
            e_ratio[i, j] = sim.box.electrons.shape[0] / sim.mc["N_e"]
            i += 1
            if duration and t_cur >= duration:
                break
    return x_ax, lum, e_ratio


def bin_events(x_ax: np.ndarray, lum: np.ndarray, dt: float, n_bins: int) -> np.ndarray:
    """Recombinations per fixed-dt step: an event at time t (x_ax after the step)
    falls in step floor(t / dt); returns counts (n_bins, sims)."""
    sims = x_ax.shape[1]
    out = np.zeros((n_bins, sims))
    for j in range(sims):
        m = lum[:, j] > 0
        idx = np.floor(x_ax[m, j] / dt).astype(np.int64)
        idx = idx[(idx >= 0) & (idx < n_bins)]
        np.add.at(out[:, j], idx, 1.0)
    return out


def live_electron_steps(x_ax: np.ndarray, lum: np.ndarray, n0: int, dt: float,
                        n_steps: int) -> np.ndarray:
    """Fixed-dt-equivalent work of an event-mode run: for each sim, the number of
    electrons alive at the start of each dt step, summed over the steps."""
    per_step = bin_events(x_ax, lum, dt, n_steps)        # recombinations per step
    gone_before = np.cumsum(per_step, axis=0) - per_step
    live = np.clip(n0 - gone_before, 0, None)
    return live.sum(axis=0)


def run_params_to_cfg(run) -> Tuple[Dict, Dict]:
    """Map a RunParams-like object onto the reference's exp_type_fp / physics_fp keys."""
    exp = dict(rho_prime=run.rho_prime, holes=run.holes, boundary_factor=run.boundary_factor,
               N_e=run.N_e, e_ratio_start=run.e_ratio_start, T_start=run.T_start,
               T_rate=run.T_rate, duration=run.n_steps * run.dt,
               max_dt=run.dt * (run.T_rate if run.T_rate else 1.0))
    phys = dict(alpha=run.alpha, b=run.b, s=run.s, E_cb=run.E_cb, E_loc_1=run.E_loc_1,
                E_loc_2=run.E_loc_2, D0=run.D0, Retrap=run.Retrap, k_b=run.k_b, D=run.D)
    return exp, phys


def randall_wilkins_peak(E: float, s: float, beta: float, k_b: float = 8.617343e-5) -> float:
    """First-order TL peak: beta E / (k T_m^2) = s exp(-E / (k T_m)), solved by Newton."""
    T = 400.0
    for _ in range(100):
        f = beta * E / (k_b * T * T) - s * math.exp(-E / (k_b * T))
        df = -2 * beta * E / (k_b * T ** 3) - s * math.exp(-E / (k_b * T)) * E / (k_b * T * T)
        T_new = T - f / df
        if abs(T_new - T) < 1e-12:
            break
        T = T_new
    return T
