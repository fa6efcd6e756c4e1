"""Sweep expansion, physics parameters and the mapping of a run config onto the
kernel's parameter block.

Mirrors the public names of the reference's src/class/engine.py:
``initialize_runs`` (engine.py:13-38) and ``Physics`` (engine.py:44-77).  The
reference's ``Box`` (engine.py:83-188) has no host counterpart: coordinates,
nearest holes and removals live in the GPU kernel's LDS; ``TLTrapSim.box``
(tl_trap_lab.py) exposes the kernel's view of a cluster for inspection.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Dict, Optional

import numpy as np

from ._abi import RunParams
from .config import Config


def initialize_runs(cfg: Config) -> Dict[int, Config]:
    """Expand list-valued keys of exp_type_fp / physics_fp into runs (engine.py:13-38).

    Despite its name in the reference this is zip-with-repetition, not a
    cartesian product: run i takes element i of every list repeated to the
    longest length, and every list length must divide the longest."""
    groups = {"exp_type_fp": cfg.exp_type_fp, "physics_fp": cfg.physics_fp}
    lengths = [len(v) for sub in groups.values() for v in sub.values() if isinstance(v, list)]
    max_length = max(lengths) if lengths else 1
    runs: Dict[int, Config] = {}
    for i in range(max_length):
        run = {}
        for name, sub in groups.items():
            out = {}
            for k, v in sub.items():
                if isinstance(v, list):
                    if max_length % len(v) != 0:
                        raise ValueError(f"max_length {max_length} is not divisible by "
                                         f"len({name}.{k})={len(v)}")
                    out[k] = (list(v) * (max_length // len(v)))[i]
                else:
                    out[k] = v
            run[name] = Config(out)
        runs[i] = Config(run)
    return runs


@dataclass(frozen=True)
class Physics:
    """Physics parameters of one run (engine.py:44-77): unknown keys raise TypeError
    exactly like ``Physics(**cfg.physics_fp)`` in the reference."""

    alpha: float
    b: float
    s: float
    E_cb: float
    E_loc_1: float
    E_loc_2: float
    D0: float
    Retrap: float
    name: Optional[str] = None
    D: Optional[float] = None
    rho_trap: Optional[float] = None
    k_b: float = 8.617333262145e-5

    def rate_cb(self, T: float) -> float:
        """Band escape rate s exp(-E_cb / kT) (engine.py:65-67)."""
        return self.s * math.exp(-self.E_cb / (self.k_b * T))

    def rate_tunnel(self, d, T: float, channel2=False):
        """Tunnelling rate b exp(-E_loc / kT - alpha d) for the given channel
        (engine.py:69-73 draws the channel per electron; here it is an argument)."""
        E_loc = np.where(channel2, self.E_loc_2, self.E_loc_1)
        return self.b * np.exp(-E_loc / (self.k_b * T) - self.alpha * np.asarray(d))

    def lifetime(self, d, T: float, channel2=False):
        return 1.0 / (self.rate_cb(T) + self.rate_tunnel(d, T, channel2))


def physics_of(run_cfg: Config) -> Physics:
    return Physics(**run_cfg.physics_fp)


def schedule(run_cfg: Config) -> Dict[str, float]:
    """Fixed-dt schedule of one run (simulate.py:40-45).

    TL: dt = max_dt / T_rate, the reference's temperature-step cap; isothermal:
    dt = exp_type_fp.dt if given, else max_dt.  n_steps covers the duration."""
    mc = run_cfg.exp_type_fp
    duration = float(mc.get("duration", 0.0))
    T_rate = float(mc.get("T_rate", 0.0))
    max_dt = float(mc.get("max_dt", 0.5))
    if T_rate:
        dt = max_dt / T_rate
    else:
        dt = float(mc.get("dt", max_dt))
    if not dt > 0:
        raise ValueError(f"time step must be > 0 (max_dt={max_dt}, T_rate={T_rate})")
    n_steps = max(1, int(math.ceil(duration / dt - 1e-9))) if duration else 1
    return {"dt": dt, "n_steps": n_steps, "duration": duration, "T_rate": T_rate,
            "T_start": float(mc.get("T_start", 0.0))}


def run_params(run_cfg: Config, *, sims: Optional[int] = None,
               e_ratio_start: Optional[float] = None, n_steps: Optional[int] = None,
               dt: Optional[float] = None, dose_rate: float = 0.0) -> RunParams:
    """Pack one run config into the kernel's parameter block (mcl_run_params)."""
    phys = physics_of(run_cfg)
    mc = run_cfg.exp_type_fp
    sch = schedule(run_cfg)
    if e_ratio_start is None:
        e_ratio_start = float(mc.get("e_ratio_start", 0.0))
    return RunParams.make(
        alpha=float(phys.alpha), b=float(phys.b), s=float(phys.s), E_cb=float(phys.E_cb),
        E_loc_1=float(phys.E_loc_1), E_loc_2=float(phys.E_loc_2), D0=float(phys.D0),
        Retrap=float(phys.Retrap), k_b=float(phys.k_b), D=float(dose_rate),
        rho_prime=float(mc.rho_prime), holes=float(mc.holes),
        boundary_factor=float(mc.boundary_factor), N_e=float(mc.N_e),
        e_ratio_start=float(e_ratio_start), T_start=sch["T_start"], T_rate=sch["T_rate"],
        dt=float(dt if dt is not None else sch["dt"]),
        n_steps=int(n_steps if n_steps is not None else sch["n_steps"]),
        sims=int(sims if sims is not None else mc.get("sims", 1)))
