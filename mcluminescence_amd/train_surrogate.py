"""Train the glow-curve surrogate on Monte Carlo curves, on the GPU.

    python -m mcluminescence_amd.train_surrogate --curves 200000 --epochs 200 \\
        --out mcluminescence_amd/surrogate_tl.npz

1. samples parameter sets uniformly in the encoded box (surrogate.RANGES),
2. simulates each with the fixed-dt Monte Carlo (sims clusters of 400 electrons,
   300 -> 750 K, 256 bins) — the training targets,
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


def make_dataset(ctx: Context, n: int, sims: int, seed: int):
    rng = np.random.default_rng(seed)
    theta = S.sample_params(n, rng)
    t0 = time.time()
    curves = S.mc_curves(ctx, theta, n=400, sims=sims, seed=S.DEFAULT_SEED + 7919 * seed)
    return theta, curves, time.time() - t0


def evaluate(net: S.Network, theta: np.ndarray, curves: np.ndarray) -> dict:
    pred = S.predict_curves(net, theta)
    e = S.nrmse(pred, curves)
    return {"n": int(len(e)), "nrmse_median": float(np.median(e)),
            "nrmse_p90": float(np.percentile(e, 90)), "nrmse_mean": float(e.mean()),
            "frac_below_2pct": float(np.mean(e <= 0.02))}


def main(argv=None) -> dict:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
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
    ctx = Context(0)
    theta, curves, t_mc = make_dataset(ctx, args.curves, args.sims, args.seed)
    print(f"MC training curves: {args.curves} x {args.sims} clusters in {t_mc:.1f} s", flush=True)
    x = torch.from_numpy(S.encode(theta)).cuda()
    t = torch.from_numpy(S.normalise(curves, args.c0)).cuda()
    net = S.Network(ctx, max_batch=max(args.batch, 8192), seed=args.seed, c0=args.c0)
    cfg = S.TrainConfig(epochs=args.epochs, batch=args.batch, lr=args.lr, seed=args.seed)
    hist = S.train(net, x, t, cfg)
    ev_theta, ev_curves, t_ev = make_dataset(ctx, args.eval, args.eval_sims, args.seed + 1000)
    ev = evaluate(net, ev_theta, ev_curves)
    # noise floor: the training-grade MC (sims clusters) against the low-noise MC
    tr_grade = S.mc_curves(ctx, ev_theta, n=400, sims=args.sims, seed=S.DEFAULT_SEED + 31)
    floor = S.nrmse(tr_grade, ev_curves)
    fine = {}
    if args.eval_fine:
        k = min(args.eval_fine, args.eval)
        fine_curves = S.mc_curves(ctx, ev_theta[:k], n=400, sims=256, seed=S.DEFAULT_SEED + 57)
        fine = evaluate(net, ev_theta[:k], fine_curves)
        fine["eval_sims"] = 256
    out = {
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
