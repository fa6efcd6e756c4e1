"""Named workloads: parameter sets for the glow-curve, decay and sweep runs.

Each builder returns a list of RunParams (one per parameter set).  The physics
keys follow the reference YAML (conf/physics_fp/*.yaml); a single trap depth E
and frequency factor s map onto the reference's two channels as E_cb = E_loc_1
= E_loc_2 = E and s = b, so k = s e^{-E/kT} (1 + e^{-alpha r}).  The hole
density rho fixes rho_prime = rho / (3/(4 pi) alpha^3) (tl_trap_lab.py:33).
"""
from __future__ import annotations

import math
from typing import List

import numpy as np

from ._abi import RunParams

K_B = 8.617343e-5          # eV/K, conf/physics_fp/basicTL12.yaml:4
D0 = 400.35                # conf/physics_fp/basicTL12.yaml:8


def rho_prime_for(rho: float, alpha: float) -> float:
    return rho / (3.0 / (4.0 * math.pi) * alpha ** 3)


def single_trap(E: float, s: float, alpha: float, rho: float, n: int, *, T_start_K: float,
                T_rate: float, dt: float, n_steps: int, sims: int, bf: float = 1.2,
                fill: float = 1.0, retrap: float = 0.5) -> RunParams:
    """One trap depth E [eV] / frequency s [1/s] with tunnelling alpha [1/m] at hole
    density rho [1/m^3]; n electrons and n core holes per cluster."""
    return RunParams.make(
        alpha=alpha, b=s, s=s, E_cb=E, E_loc_1=E, E_loc_2=E, D0=D0, Retrap=retrap, k_b=K_B,
        D=0.0, rho_prime=rho_prime_for(rho, alpha), holes=float(n), boundary_factor=bf,
        N_e=float(n), e_ratio_start=fill, T_start=T_start_K - 273.15, T_rate=T_rate, dt=dt,
        n_steps=n_steps, sims=sims)


def tl_glow_16x500(sims: int = 16) -> List[RunParams]:
    """TL glow curve: 16 clusters x 500 electrons / 500 core holes (863 with the 1.2
    boundary), rho = 1e23 m^-3, E = 1.3 eV, s = 1e12 1/s, alpha = 4e9 1/m,
    300 -> 700 K at 1 K/s, dt = 0.1 s, 4000 steps."""
    return [single_trap(1.3, 1e12, 4e9, 1e23, 500, T_start_K=300.0, T_rate=1.0, dt=0.1,
                        n_steps=4000, sims=sims)]


def iso_decay_4096x400(sims: int = 4096) -> List[RunParams]:
    """Isothermal decay at 500 K for 2000 s, dt = 0.5 s (4000 steps): 4096 clusters x
    400 electrons / 400 core holes (691 with boundary), rho = 1e23 m^-3, E = 1.4 eV,
    s = 1e13 1/s, alpha = 4e9 1/m."""
    return [single_trap(1.4, 1e13, 4e9, 1e23, 400, T_start_K=500.0, T_rate=0.0, dt=0.5,
                        n_steps=4000, sims=sims)]


def tunnelling_tl12(sims: int = 2, n: int = 2000, T_rate: float = 1.0) -> List[RunParams]:
    """The reference's shipped TL12 + basicTL12 physics (conf/exp_type_fp/TL12.yaml,
    conf/physics_fp/basicTL12.yaml): tunnelling only (E_cb = 1000 disables the band),
    alpha = 1.6e9, b = 1e12, E_loc = 1.2 eV, rho' = 3e-4, ramp from 0 degC to 800 degC."""
    duration = 800.0 / T_rate
    dt = 1.0 / T_rate if T_rate > 0 else 1.0      # 1 degC per step (max_dt = 1)
    n_steps = int(round(duration / dt))
    return [RunParams.make(alpha=1.6e9, b=1e12, s=1e12, E_cb=1000.0, E_loc_1=1.2, E_loc_2=1.2,
                           D0=D0, Retrap=0.5, k_b=K_B, D=0.0, rho_prime=3e-4, holes=float(n),
                           boundary_factor=1.2, N_e=float(n), e_ratio_start=1.0, T_start=0.0,
                           T_rate=T_rate, dt=dt, n_steps=n_steps, sims=sims)]


def tunnelling_iso(rho_prime: float, *, n: int = 2000, fill: float = 0.05, dt: float = 1.0,
                   n_steps: int = 4000, sims: int = 256, s_eff: float = 1.0, T_K: float = 500.0,
                   E_loc: float = 1.0, alpha: float = 4e9) -> List[RunParams]:
    """Tunnelling-only isothermal decay (the regime of Huntley's closed form,
    functions.py:264-268): the band is off (E_cb = 1000 eV, as in basicTL12), one
    localisation level E_loc at T_K with b chosen so that the effective tunnelling
    frequency b e^{-E_loc / kT} is s_eff, n core holes (N_e = n) filled to fill, so
    electrons are few against the holes."""
    b = s_eff / math.exp(-E_loc / (K_B * T_K))
    return [RunParams.make(alpha=alpha, b=b, s=1e12, E_cb=1000.0, E_loc_1=E_loc, E_loc_2=E_loc,
                           D0=D0, Retrap=0.5, k_b=K_B, D=0.0, rho_prime=rho_prime,
                           holes=float(n), boundary_factor=1.2, N_e=float(n), e_ratio_start=fill,
                           T_start=T_K - 273.15, T_rate=0.0, dt=dt, n_steps=n_steps, sims=sims)]


def uneven_sweep(n_sets: int = 10000, seed: int = 2024, n_bins: int = 256,
                 n_steps: int = 1024) -> List[RunParams]:
    """Batched parameter sweep with uneven clusters: half TL (0.5-5 K/s, 300-750 K),
    half ISO (400-600 K); E log-uniform in [1.0, 1.8] eV, s log-uniform in [1e10, 1e14],
    alpha uniform in [2e9, 6e9], rho log-uniform in [1e22, 1e24], fill fraction
    uniform in [0.1, 1.0], electrons per cluster uniform in [64, 2048]."""
    rng = np.random.default_rng(seed)
    runs = []
    for i in range(n_sets):
        E = float(np.exp(rng.uniform(np.log(1.0), np.log(1.8))))
        s = float(10 ** rng.uniform(10, 14))
        alpha = float(rng.uniform(2e9, 6e9))
        rho = float(10 ** rng.uniform(22, 24))
        fill = float(rng.uniform(0.1, 1.0))
        n = int(rng.integers(64, 2049))
        if i % 2 == 0:
            beta = float(rng.uniform(0.5, 5.0))
            dt = (750.0 - 300.0) / beta / n_steps
            r = single_trap(E, s, alpha, rho, n, T_start_K=300.0, T_rate=beta, dt=dt,
                            n_steps=n_steps, sims=1, fill=1.0)
        else:
            T_iso = float(rng.uniform(400.0, 600.0))
            r = single_trap(E, s, alpha, rho, n, T_start_K=T_iso, T_rate=0.0, dt=2000.0 / n_steps,
                            n_steps=n_steps, sims=1, fill=1.0)
        # dose fill fraction: electrons present at the start = fill * N_e
        r.N_e = float(n)
        r.e_ratio_start = fill
        runs.append(r)
    return runs
