"""Surrogate-assisted against plain Monte Carlo parameter fits on synthetic lab curves.

    python -m mcluminescence_amd.fit_bench --targets 4 --model mcluminescence_amd/surrogate_tl.npz
    python -m mcluminescence_amd.fit_bench --kind iso --targets 4

For each target: hidden parameters drawn inside the surrogate's box (TL: heating
rate 1 K/s; ISO: a temperature in [420, 580] K and a decay the 2000 s window
resolves), a measured curve from the Monte Carlo with Poisson noise (peak 1e4 counts),
then fit_surrogate and fit_mc on the same curve.  Prints one JSON line with the
wall time of each fit, fits/hour, the judge NRMSE, how often the fallback ran,
each fit's objective on a fresh Monte Carlo judge curve (beside the true
parameters' own) and the parameter recovery errors of both methods.
"""
from __future__ import annotations

import argparse
import json
from pathlib import Path

import numpy as np

from . import fit as F
from . import surrogate as S
from ._abi import Context

MODEL = Path(__file__).resolve().parent / "surrogate_tl.npz"
MODELS = {"tl": MODEL, "iso": Path(__file__).resolve().parent / "surrogate_iso.npz"}
ISO_TIME_CONSTANT = (5.0, 500.0)          # s: 1/k of the hidden isothermal decays


def draw_truth(rng: np.random.Generator, beta: float = 1.0) -> np.ndarray:
    """Hidden TL parameters away from the box edges (encoded coordinates in
    [-0.7, 0.7]); fill fraction in [0.3, 1]."""
    u = rng.uniform(-0.7, 0.7, S.N_IN)
    th = S.decode(u[None])[0]
    th[4] = rng.uniform(0.3, 1.0)
    th[5] = beta
    return th


def draw_truth_iso(rng: np.random.Generator) -> np.ndarray:
    """Hidden isothermal parameters away from the box edges (encoded coordinates in
    [-0.7, 0.7]), fill in [0.3, 1], T_iso in [420, 580] K, redrawn until the decay
    time 1/k lies in ISO_TIME_CONSTANT (the 2000 s window shows the decay)."""
    while True:
        u = rng.uniform(-0.7, 0.7, S.N_IN)
        th = S.ISO.decode(u[None])[0]
        th[4] = rng.uniform(0.3, 1.0)
        th[5] = rng.uniform(420.0, 580.0)
        tau = 1.0 / F.iso_rate(th)[0]
        if ISO_TIME_CONSTANT[0] <= tau <= ISO_TIME_CONSTANT[1]:
            return th


def truth_for(kind: S.CurveKind, rng: np.random.Generator) -> np.ndarray:
    return draw_truth_iso(rng) if kind.name == "iso" else draw_truth(rng)


def run(targets: int = 4, model: Path = MODEL, mc_pop: int = 128, mc_gens: int = 40,
        sur_pop: int = 1024, sur_gens: int = 150, seed: int = 0, log=print) -> dict:
    ctx = Context(0)
    net = S.Network.load(ctx, model, max_batch=max(sur_pop, 4096))
    kind = net.kind
    rng = np.random.default_rng(seed)
    rows = []
    # one untimed pass of each method loads the kernels both use
    warm = truth_for(kind, np.random.default_rng(seed + 99))
    yw, sw = F.lab_curve(ctx, warm, kind=kind)
    F.fit_surrogate(ctx, net, yw, sw, warm[5], pop=sur_pop, gens=2)
    F.fit_mc(ctx, yw, sw, warm[5], pop=8, gens=1, kind=kind)
    keys = ("log10_k", "fill") if kind.name == "iso" else ("E", "log10_s", "log10_rho_prime", "fill")
    for k in range(targets):
        truth = truth_for(kind, rng)
        y, scale = F.lab_curve(ctx, truth, rng=rng, seed=S.DEFAULT_SEED + 500 + k, kind=kind)
        rs = F.fit_surrogate(ctx, net, y, scale, truth[5], pop=sur_pop, gens=sur_gens, seed=k)
        rm = F.fit_mc(ctx, y, scale, truth[5], pop=mc_pop, gens=mc_gens, seed=k, kind=kind)
        jo = lambda th: F.judge_objective(ctx, th, y, scale, kind=kind)
        row = {"truth": truth.tolist(),
               "truth_objective": jo(truth),
               "surrogate": {"seconds": rs.seconds,
                             "objective": jo(rs.theta),
                             "judge_nrmse": rs.judge_nrmse, "fallback": rs.fallback,
                             "mc_candidates": rs.mc_candidates, "net_candidates": rs.net_candidates,
                             "recovery": rs.recovery(truth)},
               "mc": {"seconds": rm.seconds, "objective": jo(rm.theta),
                      "mc_candidates": rm.mc_candidates, "recovery": rm.recovery(truth)}}
        if log:
            log(json.dumps(row))
        rows.append(row)
    ts = np.array([r["surrogate"]["seconds"] for r in rows])
    tm = np.array([r["mc"]["seconds"] for r in rows])

    def med(method, key):
        return float(np.median([r[method]["recovery"][key] for r in rows]))

    out = {
        "kind": kind.name, "targets": targets, "model": str(model),
        "surrogate_fit_s": float(ts.mean()), "mc_fit_s": float(tm.mean()),
        "speedup": float(tm.mean() / ts.mean()),
        "surrogate_fits_per_hour": float(3600 / ts.mean()), "mc_fits_per_hour": float(3600 / tm.mean()),
        "judge_nrmse_median": float(np.median([r["surrogate"]["judge_nrmse"] for r in rows])),
        "fallback_rate": float(np.mean([r["surrogate"]["fallback"] for r in rows])),
        "objective_surrogate_median": float(np.median([r["surrogate"]["objective"] for r in rows])),
        "objective_mc_median": float(np.median([r["mc"]["objective"] for r in rows])),
        "objective_truth_median": float(np.median([r["truth_objective"] for r in rows])),
        "recovery_median": {m: {k: med(m, k) for k in keys} for m in ("surrogate", "mc")},
        "mc_config": {"pop": mc_pop, "gens": mc_gens},
        "surrogate_config": {"pop": sur_pop, "gens": sur_gens},
    }
    net.close()
    ctx.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--kind", choices=sorted(MODELS), default="tl")
    ap.add_argument("--targets", type=int, default=4)
    ap.add_argument("--model", type=Path, default=None, help="default: the kind's committed network")
    ap.add_argument("--mc-pop", type=int, default=128)
    ap.add_argument("--mc-gens", type=int, default=40)
    ap.add_argument("--sur-pop", type=int, default=1024)
    ap.add_argument("--sur-gens", type=int, default=150)
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)
    out = run(a.targets, a.model or MODELS[a.kind], a.mc_pop, a.mc_gens, a.sur_pop, a.sur_gens,
              a.seed)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
