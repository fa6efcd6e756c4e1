"""Trap-parameter estimation from a glow curve, in the form of the reference's
optimizer: differential evolution ("randtobest1bin", mutation 0.5, recombination
0.3 — src/class/optimizer.py:104-110) minimising the mean squared difference
between a simulated and a measured curve (optimizer.py:66-84 objective).

Two ways to evaluate a candidate population:

  fit_mc         every candidate of every generation is simulated by the Monte
                 Carlo (one batched launch per generation) — the plain method;
  fit_surrogate  candidates are scored by the trained network (one inference batch
                 per generation); the best one is re-simulated by the Monte Carlo
                 judge (64 clusters x 1000 electrons), and if the network's curve
                 misses the judge's by more than the gate (2 % NRMSE) the search
                 falls back to Monte Carlo generations seeded with the network's
                 final population.

The unknowns are E, s, alpha, rho and the fill fraction; the sixth input is the
experiment's own: the heating rate of a TL curve, the temperature of an
isothermal decay (the kind of the curve, surrogate.TL or surrogate.ISO, the
reference's TL_lab and ISO_lab objectives, optimizer.py:73-78).  A TL curve
depends on alpha and rho only through rho' = rho / (3/(4 pi) alpha^3)
(tl_trap_lab.py:33), so TL recovery is reported for E, s, rho' and fill.  An
isothermal decay at one temperature depends on E and s only through the rate
k = s e^{-E/kT} (the tunnelling term is below 1e-4 of it over the box), so its
recovery is reported for log10 k and fill.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np

from . import surrogate as S
from ._abi import DEFAULT_SEED, Context
from .workloads import rho_prime_for

N_FIT = 5          # E, s, alpha, rho, fill
JUDGE_SIMS, JUDGE_N = 64, 1000
K_B = 8.617343e-5


def full_theta(u5: np.ndarray, fixed: float, kind: S.CurveKind = S.TL) -> np.ndarray:
    """Encoded fit vector(s) [.., 5] -> physical parameters [.., 6] with the fixed
    sixth input (heating rate for TL, temperature for ISO)."""
    u5 = np.atleast_2d(u5)
    th = kind.decode(np.concatenate([u5, np.zeros((u5.shape[0], 1))], axis=1))
    th[:, 5] = fixed
    return th


def iso_rate(theta: np.ndarray) -> np.ndarray:
    """k = s e^{-E / kT_iso} [1/s] of isothermal parameter rows."""
    th = np.atleast_2d(theta)
    return th[:, 1] * np.exp(-th[:, 0] / (K_B * th[:, 5]))


def lab_curve(ctx: Context, theta: np.ndarray, peak_counts: float = 1e4,
              rng: Optional[np.random.Generator] = None, seed: int = DEFAULT_SEED + 101,
              kind: S.CurveKind = S.TL):
    """Synthetic measured curve: Monte Carlo at theta (64 clusters x 1000 electrons)
    scaled to a peak of peak_counts, with Poisson counting noise.  Returns (counts,
    scale) where scale converts a fraction-of-N_e curve into counts."""
    rng = rng or np.random.default_rng(0)
    c = kind.mc_curves(ctx, np.atleast_2d(theta), n=JUDGE_N, sims=JUDGE_SIMS, seed=seed)[0]
    scale = peak_counts / max(c.max(), 1e-30)
    return rng.poisson(c * scale).astype(np.float64), scale


def objective(curves: np.ndarray, y: np.ndarray, scale: float) -> np.ndarray:
    """Mean squared difference of scaled candidate curves and the measured curve,
    divided by the measured peak squared (so sqrt(objective) is an NRMSE)."""
    d = np.atleast_2d(curves) * scale - y
    return np.mean(d * d, axis=1) / max(y.max(), 1.0) ** 2


class DifferentialEvolution:
    """Vectorised DE/rand-to-best/1/bin over the box [-1, 1]^d (scipy's
    'randtobest1bin': mutant = x_r0 + F (best - x_r0) + F (x_r1 - x_r2), binomial
    crossover with at least one mutant coordinate; coordinates leaving the box are
    redrawn uniformly)."""

    def __init__(self, d: int, pop: int, rng: np.random.Generator, F: float = 0.5,
                 CR: float = 0.3, init: Optional[np.ndarray] = None):
        self.d, self.P, self.F, self.CR, self.rng = d, pop, F, CR, rng
        self.x = init.copy() if init is not None else rng.uniform(-1, 1, (pop, d))
        self.f = np.full(pop, np.inf)

    def seed_fitness(self, f: np.ndarray) -> None:
        self.f = np.asarray(f, np.float64).copy()

    def propose(self) -> np.ndarray:
        P, d, rng = self.P, self.d, self.rng
        best = self.x[np.argmin(self.f)]
        # three distinct partners per target, all different from the target
        me = np.arange(P)[:, None]
        r = rng.integers(0, P, (P, 3))
        for _ in range(64):
            bad = ((r == me).any(axis=1) | (r[:, 0] == r[:, 1]) | (r[:, 0] == r[:, 2]) |
                   (r[:, 1] == r[:, 2]))
            if not bad.any():
                break
            r[bad] = rng.integers(0, P, (int(bad.sum()), 3))
        x0, x1, x2 = self.x[r[:, 0]], self.x[r[:, 1]], self.x[r[:, 2]]
        mutant = x0 + self.F * (best - x0) + self.F * (x1 - x2)
        cross = rng.uniform(size=(P, d)) < self.CR
        cross[np.arange(P), rng.integers(0, d, P)] = True
        trial = np.where(cross, mutant, self.x)
        bad = (trial < -1) | (trial > 1)
        trial[bad] = rng.uniform(-1, 1, bad.sum())
        return trial

    def accept(self, trial: np.ndarray, f: np.ndarray) -> None:
        better = f <= self.f
        self.x[better] = trial[better]
        self.f[better] = f[better]

    @property
    def best(self):
        i = int(np.argmin(self.f))
        return self.x[i].copy(), float(self.f[i])


class NetworkDE:
    """The network-scored search of the surrogate-assisted fit, in device memory
    (mcl_de_*, include/mcl_surrogate.h): the DE/rand-to-best/1/bin step of
    DifferentialEvolution, with the population, the trials and the network's
    curves on the GPU.  A generation is one kernel (trials, the network forward in
    LDS, objectives and selection), or with layered=True a propose kernel, the
    network's per-layer kernels and a score-and-accept kernel, with identical
    results; run(k) queues k generations with no host round trip.  The objective
    is objective() of the network's de-normalised curve: mean over bins of
    (curve - y)^2 / peak(y)^2."""

    def __init__(self, net: S.Network, y: np.ndarray, scale: float, fixed: Sequence[float],
                 dim: int, pop: int, seed: int, F: float = 0.5, CR: float = 0.3,
                 layered: bool = False, peak: Optional[float] = None):
        import torch
        from ._abi import DE_LAYERED, DeParams
        self.net, self.dim, self.pop = net, dim, pop
        self.target = torch.from_numpy(np.asarray(y, np.float32)).cuda()
        p = DeParams()
        p.pop, p.dim, p.F, p.CR, p.seed = pop, dim, F, CR, int(seed) & (2 ** 64 - 1)
        for k, v in enumerate(fixed):
            p.fixed[k] = float(v)
        p.log_span = math.log1p(1.0 / net.c0)
        p.out_scale = net.c0 * scale
        p.inv_peak2 = 1.0 / max(float(np.max(y)) if peak is None else float(peak), 1.0) ** 2
        p.flags = DE_LAYERED if layered else 0
        h = C.c_void_p()
        with net._on_stream(self.target):
            net.ctx.check(net.lib.mcl_de_create(net.handle, C.byref(p), self.target.data_ptr(),
                                                C.byref(h)), "mcl_de_create")
        self.handle = h

    def run(self, gens: int) -> None:
        with self.net._on_stream(self.target):
            self.net.ctx.check(self.net.lib.mcl_de_run(self.handle, int(gens)), "mcl_de_run")

    def best(self):
        """(coordinates [dim] float64, objective) of the best member."""
        x = np.zeros(self.dim, np.float32)
        f = C.c_float()
        with self.net._on_stream(self.target):
            self.net.ctx.check(self.net.lib.mcl_de_best(self.handle, x.ctypes.data, C.byref(f)),
                               "mcl_de_best")
        return x.astype(np.float64), float(f.value)

    def population(self):
        """(x [pop][dim], f [pop]) as CUDA float32 tensors."""
        import torch
        x = torch.empty((self.pop, self.dim), dtype=torch.float32, device="cuda")
        f = torch.empty(self.pop, dtype=torch.float32, device="cuda")
        with self.net._on_stream(self.target, x, f):
            self.net.ctx.check(self.net.lib.mcl_de_population(self.handle, x.data_ptr(),
                                                              f.data_ptr()), "mcl_de_population")
        return x, f

    def close(self) -> None:
        """Free the search (after the work queued on the network's stream)."""
        if getattr(self, "handle", None):
            with self.net._on_stream():
                self.net.lib.mcl_de_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class FitResult:
    theta: np.ndarray                 # physical parameters [6] (beta / T_iso included)
    objective_mc: float               # objective of theta on the Monte Carlo judge curve
    judge_nrmse: float                # network vs judge curve at theta (nan for fit_mc)
    seconds: float
    mc_candidates: int                # candidate curves simulated by the Monte Carlo
    net_candidates: int               # candidate curves evaluated by the network
    fallback: bool = False
    history: List[float] = field(default_factory=list)
    kind: str = "tl"

    def recovery(self, truth: np.ndarray) -> dict:
        """Errors against the true parameters.  TL: |dE| [eV], |d log10 s|,
        |d log10 rho'|, |d fill|.  ISO: |d log10 k| (k = s e^{-E/kT}), |d fill|."""
        t, h = np.asarray(truth, float), self.theta
        if self.kind == "iso":
            return {"log10_k": abs(math.log10(iso_rate(h)[0] / iso_rate(t)[0])),
                    "fill": abs(h[4] - t[4])}
        rp = lambda th: rho_prime_for(th[3], th[2])
        return {"E": abs(h[0] - t[0]), "log10_s": abs(math.log10(h[1] / t[1])),
                "log10_rho_prime": abs(math.log10(rp(h) / rp(t))), "fill": abs(h[4] - t[4])}


def judge_objective(ctx: Context, theta: np.ndarray, y: np.ndarray, scale: float,
                    seed: int = DEFAULT_SEED + 999, kind: S.CurveKind = S.TL) -> float:
    """Objective of one parameter set on a fresh Monte Carlo judge curve (64 clusters x
    1000 electrons, a seed no fit uses): the common yardstick for comparing fits."""
    c = kind.mc_curves(ctx, np.atleast_2d(theta), n=JUDGE_N, sims=JUDGE_SIMS, seed=seed)
    return float(objective(c, y, scale)[0])


def _mc_eval(ctx: Context, th: np.ndarray, sims: int, n: int, seed: int,
             kind: S.CurveKind = S.TL) -> np.ndarray:
    return kind.mc_curves(ctx, th, n=n, sims=sims, seed=seed)


def fit_mc(ctx: Context, y: np.ndarray, scale: float, fixed: float, pop: int = 64,
           gens: int = 30, sims: int = JUDGE_SIMS, n: int = JUDGE_N, seed: int = 0,
           init: Optional[np.ndarray] = None, log: Optional[Callable[[str], None]] = None,
           kind: S.CurveKind = S.TL) -> FitResult:
    """Plain Monte Carlo DE: every candidate simulated with sims clusters of n electrons.
    fixed: the heating rate (TL) or the temperature (ISO)."""
    t0 = time.time()
    rng = np.random.default_rng(seed)
    de = DifferentialEvolution(N_FIT, pop, rng, init=init)
    key = DEFAULT_SEED + 1000 * seed
    de.seed_fitness(objective(_mc_eval(ctx, full_theta(de.x, fixed, kind), sims, n, key, kind),
                              y, scale))
    hist = [de.best[1]]
    for g in range(gens):
        trial = de.propose()
        f = objective(_mc_eval(ctx, full_theta(trial, fixed, kind), sims, n, key + g + 1, kind),
                      y, scale)
        de.accept(trial, f)
        hist.append(de.best[1])
        if log:
            log(f"mc gen {g + 1}: best {hist[-1]:.3e}")
    u, fbest = de.best
    return FitResult(full_theta(u, fixed, kind)[0], fbest, float("nan"), time.time() - t0,
                     pop * (gens + 1), 0, history=hist, kind=kind.name)


def _network_search(net: S.Network, y: np.ndarray, scale: float, u_fixed: float, pop: int,
                    gens: int, seed: int, patience: int, peak: float):
    """One network-scored DE; returns (search, best-objective history, generations)."""
    de = NetworkDE(net, y, scale, [u_fixed], N_FIT, pop, seed, peak=peak)
    hist = [de.best()[1]]
    g = 0
    try:
        while g < gens:
            k = min(patience, gens - g)
            de.run(k)
            g += k
            hist.append(de.best()[1])
            if hist[-1] > hist[-2] * (1.0 - 1e-4):
                break
    except Exception:
        de.close()
        raise
    return de, hist, g


def fit_surrogate(ctx: Context, net: S.Network, y: np.ndarray, scale: float, fixed: float,
                  pop: int = 1024, gens: int = 150, gate: float = 0.02, seed: int = 0,
                  fallback_pop: int = 64, fallback_gens: int = 20, patience: int = 25,
                  log: Optional[Callable[[str], None]] = None) -> FitResult:
    """Network-scored DE (population on the device), Monte Carlo judge of the best
    candidate, Monte Carlo fallback when the network misses the judge by more than
    gate (NRMSE).  The curve kind is the network's (TL: fixed is the heating rate;
    ISO: the temperature).  Every `patience` generations the best objective is read
    back; the search stops once it has improved by less than 1e-4 of itself since."""
    t0 = time.time()
    kind = net.kind
    peak = max(float(np.max(y)), 1.0)
    # the encoded fit coordinates are the network's inputs; the fixed one is the sixth
    u_fixed = float(kind.encode(full_theta(np.zeros(N_FIT), fixed, kind))[0, 5])
    de, hist, g = _network_search(net, y, scale, u_fixed, pop, gens, seed, patience, peak)
    n_net = pop * (g + 1)
    try:
        u, _ = de.best()
        th = full_theta(u, fixed, kind)
        judge = _mc_eval(ctx, th, JUDGE_SIMS, JUDGE_N, DEFAULT_SEED + 77 + seed, kind)[0]
        pred = S.predict_curves(net, th)[0]
        err = float(S.nrmse(pred, judge)[0])
        if log:
            log(f"surrogate fit: {g} generations, best {hist[-1]:.3e}, judge NRMSE {err:.4f}")
        if err <= gate:
            return FitResult(th[0], float(objective(judge, y, scale)[0]), err, time.time() - t0,
                             1, n_net, False, hist, kind.name)
        xs, fs = de.population()        # the fallback's starting members
    finally:
        de.close()
    # fallback: Monte Carlo generations from the network's best candidates
    import torch
    order = torch.argsort(fs)[:fallback_pop]
    res = fit_mc(ctx, y, scale, fixed, pop=fallback_pop, gens=fallback_gens, seed=seed + 1,
                 init=xs[order].double().cpu().numpy(), log=log, kind=kind)
    return FitResult(res.theta, res.objective_mc, err, time.time() - t0,
                     1 + res.mc_candidates, n_net, True, hist + res.history, kind.name)
