"""Train a curve surrogate on Monte Carlo curves, on the GPU.

    python -m mcluminescence_amd.train_surrogate --curves 200000 --epochs 200 \\
        --out mcluminescence_amd/surrogate_tl.npz
    python -m mcluminescence_amd.train_surrogate --kind iso --curves 200000 --sims 32 \\
        --epochs 400 --out mcluminescence_amd/surrogate_iso.npz

1. samples parameter sets in the encoded box of the kind (surrogate.RANGES for
   TL glow curves, surrogate.ISO_RANGES for isothermal decays): uniformly, and for
   ISO half of them uniformly among the sets whose decay rate k = s e^{-E/kT}
   lies in [10^-4.5, 10] 1/s (the decays the 2000 s window resolves; the rest of
   the box gives empty or first-step curves),
2. simulates each with the fixed-dt Monte Carlo (sims clusters of 400 electrons;
   TL 300 -> 750 K, ISO 2000 s at dt 0.5 s; 256 bins) — the training targets,
3. trains the BF16 network with the library's forward / backward / AdamW kernels,
4. scores it on held-out parameter sets against low-noise Monte Carlo curves
   (eval_sims clusters each): NRMSE = rms(curve difference) / peak of the MC curve.

Prints one JSON summary line; the network is saved as an .npz of float32 arrays.
"""
from __future__ import annotations

import argparse
import json
import time
from pathlib import Path

import numpy as np

from . import surrogate as S
from ._abi import Context


LOG10_K_RESOLVED = (-4.5, 1.0)


def sample(kind: S.CurveKind, n: int, rng: np.random.Generator,
           resolved_only: bool = False) -> np.ndarray:
    """Training parameter sets of a kind (see the module docstring); resolved_only:
    every ISO set among those the window resolves (held-out scoring: an empty or
    first-step curve has no peak to normalise by)."""
    if kind.name != "iso":
        return kind.sample_params(n, rng)
    n_uniform = 0 if resolved_only else n - n // 2
    out = [kind.sample_params(n_uniform, rng)]
    need = n - n_uniform
    while need > 0:
        th = kind.sample_params(4 * need + 64, rng)
        lk = np.log10(th[:, 1]) - th[:, 0] / (8.617343e-5 * th[:, 5]) / np.log(10.0)
        ok = th[(lk >= LOG10_K_RESOLVED[0]) & (lk <= LOG10_K_RESOLVED[1])][:need]
        out.append(ok)
        need -= len(ok)
    theta = np.concatenate(out)
    return theta[rng.permutation(len(theta))]


def make_dataset(ctx: Context, n: int, sims: int, seed: int, kind: S.CurveKind = S.TL,
                 part: int = 32768, resolved_only: bool = False):
    """theta [n, 6] and their Monte Carlo curves, simulated part sets at a time (a
    progress line per part; the same curves as one call, chunking being by 8192)."""
    rng = np.random.default_rng(seed)
    theta = sample(kind, n, rng, resolved_only)
    t0 = time.time()
    curves = np.empty((n, S.N_BINS))
    for p0 in range(0, n, part):
        # mc_curves seeds chunk c0 with seed + c0: offset by p0 to keep one stream
        curves[p0:p0 + part] = kind.mc_curves(ctx, theta[p0:p0 + part], n=400, sims=sims,
                                              seed=S.DEFAULT_SEED + 7919 * seed + p0)
        if n > part:
            print(f"  {min(n, p0 + part)}/{n} curves, {time.time() - t0:.1f} s", flush=True)
    return theta, curves, time.time() - t0


def evaluate(net: S.Network, theta: np.ndarray, curves: np.ndarray) -> dict:
    pred = S.predict_curves(net, theta)
    e = S.nrmse(pred, curves)
    return {"n": int(len(e)), "nrmse_median": float(np.median(e)),
            "nrmse_p90": float(np.percentile(e, 90)), "nrmse_mean": float(e.mean()),
            "frac_below_2pct": float(np.mean(e <= 0.02))}


def main(argv=None) -> dict:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--kind", choices=sorted(S.KINDS), default="tl")
    ap.add_argument("--curves", type=int, default=200_000)
    ap.add_argument("--sims", type=int, default=8, help="clusters per training curve")
    ap.add_argument("--epochs", type=int, default=200)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--eval", type=int, default=2000, help="held-out parameter sets")
    ap.add_argument("--eval-sims", type=int, default=64)
    ap.add_argument("--c0", type=float, default=S.C0, help="log-normalisation floor")
    ap.add_argument("--eval-fine", type=int, default=256,
                    help="held-out sets also scored against 256-cluster Monte Carlo")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", type=Path, default=None)
    args = ap.parse_args(argv)

    import torch
    kind = S.KINDS[args.kind]
    ctx = Context(0)
    theta, curves, t_mc = make_dataset(ctx, args.curves, args.sims, args.seed, kind)
    print(f"MC training curves ({kind.name}): {args.curves} x {args.sims} clusters in {t_mc:.1f} s",
          flush=True)
    x = torch.from_numpy(kind.encode(theta)).cuda()
    t = torch.from_numpy(S.normalise(curves, args.c0)).cuda()
    net = S.Network(ctx, max_batch=max(args.batch, 8192), seed=args.seed, c0=args.c0, kind=kind)
    cfg = S.TrainConfig(epochs=args.epochs, batch=args.batch, lr=args.lr, seed=args.seed)
    hist = S.train(net, x, t, cfg)
    ev_theta, ev_curves, t_ev = make_dataset(ctx, args.eval, args.eval_sims, args.seed + 1000, kind,
                                             resolved_only=True)
    ev = evaluate(net, ev_theta, ev_curves)
    # noise floor: the training-grade MC (sims clusters) against the low-noise MC
    tr_grade = kind.mc_curves(ctx, ev_theta, n=400, sims=args.sims, seed=S.DEFAULT_SEED + 31)
    floor = S.nrmse(tr_grade, ev_curves)
    fine = {}
    if args.eval_fine:
        k = min(args.eval_fine, args.eval)
        fine_curves = kind.mc_curves(ctx, ev_theta[:k], n=400, sims=256, seed=S.DEFAULT_SEED + 57)
        fine = evaluate(net, ev_theta[:k], fine_curves)
        fine["eval_sims"] = 256
    out = {
        "kind": kind.name,
        "curves": args.curves, "sims": args.sims, "mc_seconds": round(t_mc, 2),
        "mc_curves_per_s": round(args.curves / t_mc, 1),
        "epochs": args.epochs, "steps": hist["steps"], "train_seconds": round(hist["wall_s"], 2),
        "train_step_ms": round(float(np.median(hist["step_ms_last"])), 4),
        "final_loss": float(hist["loss"][-50:].mean()),
        "eval": ev, "eval_sims": args.eval_sims,
        "mc_noise_nrmse_median": float(np.median(floor)),
        "eval_fine": fine, "c0": args.c0,
    }
    if args.out:
        net.save(args.out)
        out["saved"] = str(args.out)
    print(json.dumps(out), flush=True)
    net.close()
    ctx.close()
    return out


if __name__ == "__main__":
    main()
