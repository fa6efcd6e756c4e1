"""Numpy restatement of the reference's lab objectives TLTrapSim.TL_lab and
ISO_lab (src/class/tl_trap_lab.py:65-179), on the event-mode classes of
oracle/event_mc.py (Box with the reference's incremental add_electron,
engine.py:133-152).

TEST INFRASTRUCTURE ONLY: the statistical checker of the GPU lab kernel, never
imported by the product.  Each function returns the simulated trapped fractions
(not the MSE), one simulation per row or experiment as in the reference; the
random source is a seeded numpy Generator.
"""
from __future__ import annotations

from typing import Dict, List

import numpy as np

from .event_mc import Sim


def tl_lab_fractions(exp: Dict, phys: Dict, rows: List[Dict], rng: np.random.Generator) -> np.ndarray:
    """Final n_e / N_e of every TL row (tl_trap_lab.py:75-107)."""
    out = []
    for row in rows:
        sim = Sim(exp, phys, float(row["e_ratio"]), rng)
        T_start = row["T_start"] + 273.15
        T_rate = (row["T_end"] - row["T_start"]) / row["Duration"]
        sim._update_lifetimes(T_start)
        D = float(phys.get("D", 0.0))
        tf = sim._filling_time(D)
        t = 0.0
        ratio = np.nan
        while t < row["Duration"]:
            dt_recomb = np.min(sim._recomb_wait) if sim._recomb_wait.size else tf
            dt = min(dt_recomb, tf)
            T_now = T_start + T_rate * (t + dt)
            t += dt
            if dt == tf:
                sim.box.add_electron()
            else:
                e_idx = int(np.argmin(sim._recomb_wait))
                _, nearest = sim.box.nearest()
                sim.box.remove_pair(e_idx, int(nearest[e_idx]))
            tf = sim._filling_time(D)
            sim._update_lifetimes(T_now)
            ratio = sim.box.electrons.shape[0] / exp["N_e"]
        out.append(ratio)
    return np.array(out)


def iso_lab_fractions(exp: Dict, phys: Dict, experiments: List[Dict],
                      rng: np.random.Generator) -> List[np.ndarray]:
    """n_e / N_e at every observation time of every isothermal experiment
    (tl_trap_lab.py:130-172).  An experiment is {e_ratio, temp, dose, times}."""
    out = []
    for ex in experiments:
        sim = Sim(exp, phys, float(ex["e_ratio"]), rng)
        T_iso = ex["temp"] + 273.15
        D = float(ex["dose"])
        sim._update_lifetimes(T_iso)
        tf = sim._filling_time(D)
        t = 0.0
        times = np.asarray(ex["times"], dtype=np.float64)
        k = 0
        got = np.zeros(len(times))
        while k < len(times):
            dt_recomb = np.min(sim._recomb_wait) if sim._recomb_wait.size else tf
            dt = min(dt_recomb, tf)
            t += dt
            if dt == tf:
                sim.box.add_electron()
            else:
                e_idx = int(np.argmin(sim._recomb_wait))
                _, nearest = sim.box.nearest()
                sim.box.remove_pair(e_idx, int(nearest[e_idx]))
            tf = sim._filling_time(D)
            sim._update_lifetimes(T_iso)
            while k < len(times) and t >= times[k]:
                got[k] = sim.box.electrons.shape[0] / exp["N_e"]
                k += 1
        out.append(got)
    return out
