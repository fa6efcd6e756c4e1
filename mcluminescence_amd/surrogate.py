"""Learned curve surrogates: BF16 matrix-core MLPs trained on Monte Carlo curves.

A network (include/mcl_surrogate.h, kernels in csrc/mlp.hip) maps six trap
parameters to a 256-bin curve.  Two kinds of curve, each its own network:

  TL   thermoluminescence glow curve: inputs E [eV], s [1/s], alpha [1/m], rho
       [1/m^3] (hole density), fill fraction (e_ratio_start), heating rate beta
       [K/s]; recombinations per temperature bin between 300 K and 750 K (2048
       fixed steps, 8 per bin).
       The curve is the fraction of N_e recombining in each bin.
  ISO  isothermal decay curve: the same five trap inputs and the temperature
       T_iso [K] in [400, 600]; the trapped fraction n_e / N_e at 256 log-spaced
       times of the decay at T_iso for 2000 s with dt = 0.5 s (4000 steps, the
       schedule of the reference's isothermal runs and of the 4096 x 400 decay
       workload): the ends of 47 single-step bins, then of 209 bins growing
       geometrically to step 4000.  This is what ISO_lab compares with the lab's
       trapped fractions at its observation times (tl_trap_lab.py:163-172).

Both outputs are fractions of N_e, log-normalised:
y = ln(1 + c / c0) / ln(1 + 1 / c0) (c0 = 1e-3 by default, stored with the network),
with the single-trap physics of ``workloads.single_trap`` (E_cb = E_loc = E,
b = s).  Training curves come from the fixed-dt Monte Carlo (mcl_run_fixed_dt);
the networks are trained on the GPU by the library's own forward, backward and
AdamW kernels.  The reference has no counterpart: its fit
(src/class/optimizer.py:68-111) simulates every candidate, for TL_lab and
ISO_lab alike (tl_trap_lab.py:65-179, optimizer.py:73-78).
"""
from __future__ import annotations

import contextlib
import ctypes as C
import math
import time
from dataclasses import dataclass
from pathlib import Path
from typing import Callable, Optional, Sequence, Tuple

import numpy as np

from . import workloads as W
from ._abi import DEFAULT_SEED, F_PERSISTENT, AdamW, Context, MclError, MlpDesc, RunParams, run_table

N_IN = 6
N_BINS = 256
T_LO, T_HI = 300.0, 750.0          # K: glow-curve window
STEPS_PER_BIN = 8
N_STEPS = N_BINS * STEPS_PER_BIN
C0 = 1e-3                          # default log-normalisation floor
ISO_DT, ISO_STEPS = 0.5, 4000      # s, steps: the isothermal schedule (2000 s)
ISO_LINEAR_BINS = 47               # single-step bins before the geometric ones


def iso_edges(n_steps: int = ISO_STEPS, n_bins: int = N_BINS,
              linear: int = ISO_LINEAR_BINS) -> np.ndarray:
    """Step edges of the log-time bins: 0, 1, .., linear, then
    round(linear r^k) for k = 1 .. n_bins - linear with r = (n_steps / linear)^(1 /
    (n_bins - linear)), so the last edge is n_steps.  linear (r - 1) >= 1 keeps the
    rounded edges strictly increasing."""
    r = (n_steps / linear) ** (1.0 / (n_bins - linear))
    if linear * (r - 1.0) < 1.0:
        raise ValueError("geometric bins narrower than one step")
    geo = np.rint(linear * r ** np.arange(1, n_bins - linear + 1)).astype(np.int64)
    edges = np.concatenate([np.arange(linear + 1), geo]).astype(np.int32)
    edges[-1] = n_steps
    assert np.all(np.diff(edges) > 0) and edges.size == n_bins + 1
    return edges


@dataclass(frozen=True)
class CurveKind:
    """One kind of surrogate curve: its parameter box, Monte Carlo schedule and bins.
    A curve point is the fraction of N_e recombining in the bin, or with
    trapped=True the fraction still trapped at the end of the bin."""
    name: str
    ranges: tuple                       # (name, low, high, log-scaled) per input
    n_steps: int
    edges: Optional[tuple] = None       # step edges; None: n_steps / N_BINS steps a bin
    trapped: bool = False

    @property
    def low(self) -> np.ndarray:
        return np.array([r[1] for r in self.ranges])

    @property
    def high(self) -> np.ndarray:
        return np.array([r[2] for r in self.ranges])

    @property
    def logs(self) -> np.ndarray:
        return np.array([r[3] for r in self.ranges])

    def encode(self, theta: np.ndarray) -> np.ndarray:
        """Physical parameters [N, 6] -> network inputs in [-1, 1] (log axes where set)."""
        th = np.asarray(theta, np.float64)
        lo = np.where(self.logs, np.log(self.low), self.low)
        hi = np.where(self.logs, np.log(self.high), self.high)
        v = np.where(self.logs, np.log(np.maximum(th, 1e-300)), th)
        return (2.0 * (v - lo) / (hi - lo) - 1.0).astype(np.float32)

    def decode(self, u: np.ndarray) -> np.ndarray:
        """Inverse of encode."""
        u = np.asarray(u, np.float64)
        lo = np.where(self.logs, np.log(self.low), self.low)
        hi = np.where(self.logs, np.log(self.high), self.high)
        v = lo + (u + 1.0) * 0.5 * (hi - lo)
        out = v.copy()
        out[..., self.logs] = np.exp(v[..., self.logs])
        return out

    def sample_params(self, n: int, rng: np.random.Generator) -> np.ndarray:
        """Uniform in the encoded box."""
        return self.decode(rng.uniform(-1.0, 1.0, size=(n, N_IN)))

    def run(self, theta: Sequence[float], n: int, sims: int) -> RunParams:
        """Monte Carlo run of one parameter set: n core holes and N_e = n electron
        traps filled to the fill fraction, on this kind's schedule."""
        E, s, alpha, rho, fill, x = (float(v) for v in theta)
        if self.name == "tl":
            dt = (T_HI - T_LO) / x / self.n_steps
            return W.single_trap(E, s, alpha, rho, n, T_start_K=T_LO, T_rate=x, dt=dt,
                                 n_steps=self.n_steps, sims=sims, fill=fill)
        return W.single_trap(E, s, alpha, rho, n, T_start_K=x, T_rate=0.0, dt=ISO_DT,
                             n_steps=self.n_steps, sims=sims, fill=fill)

    def table(self, theta: np.ndarray, n: int, sims: int) -> np.ndarray:
        """run() for every row of theta [N, 6], as a run table (RUN_DTYPE), built
        column-wise; the same numbers as the per-set runs (a test compares the bytes)."""
        theta = np.atleast_2d(np.asarray(theta, np.float64))
        E, s, alpha, rho, fill, x = (theta[:, j] for j in range(6))
        t = run_table(theta.shape[0])
        t["alpha"], t["b"], t["s"] = alpha, s, s
        t["E_cb"], t["E_loc_1"], t["E_loc_2"] = E, E, E
        t["D0"], t["Retrap"], t["k_b"], t["D"] = W.D0, 0.5, W.K_B, 0.0
        # the scalar formula per set (its pow and division, as single_trap computes them)
        t["rho_prime"] = [W.rho_prime_for(float(r), float(a)) for r, a in zip(rho, alpha)]
        t["holes"], t["boundary_factor"], t["N_e"], t["e_ratio_start"] = float(n), 1.2, float(n), fill
        if self.name == "tl":
            t["T_start"], t["T_rate"], t["dt"] = T_LO - 273.15, x, (T_HI - T_LO) / x / self.n_steps
        else:
            t["T_start"], t["T_rate"], t["dt"] = x - 273.15, 0.0, ISO_DT
        t["n_steps"], t["sims"] = self.n_steps, sims
        return t

    def mc_curves(self, ctx: Context, theta: np.ndarray, n: int = 400, sims: int = 8,
                  seed: int = DEFAULT_SEED, chunk: int = 8192) -> np.ndarray:
        """Monte Carlo curves [N, 256]: recombinations per bin summed over the sims,
        divided by sims * N_e (trapped kinds: the initial electrons less the
        recombinations up to the end of each bin, over sims * N_e).  A chunk of
        several runs is one persistent launch (clusters taken largest expected work
        first), summed over the sims on the device: by the kernel's own binning for
        equal bins, else from per-step counts by mcl_bin_counts.  A single TL set
        runs one workgroup per cluster.  The counts are the same either way."""
        import torch
        theta = np.atleast_2d(theta)
        out = np.empty((theta.shape[0], N_BINS), np.float64)
        for c0 in range(0, theta.shape[0], chunk):
            part = theta[c0:c0 + chunk]
            runs = self.table(part, n, sims)
            if self.edges is None and len(part) == 1:
                counts, _, _ = ctx.run_fixed_dt(runs, seed=seed + c0, n_bins=N_BINS, flags=0)
                out[c0] = counts.sum(axis=0) / (sims * n)
                continue
            n_bins = N_BINS if self.edges is None else 0
            plan = ctx.plan(runs, seed=seed + c0, n_bins=n_bins, flags=F_PERSISTENT)
            try:
                row = plan.info.row_len
                counts = torch.empty((plan.info.total_clusters, row), dtype=torch.int32,
                                     device="cuda")
                torch.cuda.current_stream().synchronize()
                plan.launch(counts.data_ptr())
                if self.edges is None:
                    plan.sync()
                    sums = counts.view(len(part), sims, N_BINS).sum(dim=1, dtype=torch.int64)
                else:
                    sums = torch.empty((len(part), N_BINS), dtype=torch.int64, device="cuda")
                    ctx.bin_counts(counts.data_ptr(), len(part), sims, row, self.edges,
                                   sums.data_ptr())
                    plan.sync()
                sums = sums.cpu().numpy()
            finally:
                plan.close()
            if self.trapped:
                # electrons at the start: int(N_e * fill), the geometry's truncation
                n0 = (float(n) * runs["e_ratio_start"]).astype(np.int64)[:, None] * sims
                sums = n0 - np.cumsum(sums, axis=1)
            out[c0:c0 + len(part)] = sums / (sims * n)
        return out

    def axis(self) -> np.ndarray:
        """Where each curve point sits: bin-centre temperature [K] for TL, the time
        at the end of each bin [s] for ISO."""
        if self.name == "tl":
            w = (T_HI - T_LO) / N_BINS
            return T_LO + w * (np.arange(N_BINS) + 0.5)
        return np.asarray(self.edges[1:], np.float64) * ISO_DT


# parameter boxes: (name, low, high, log-scaled)
RANGES = (
    ("E", 1.0, 1.8, True),
    ("s", 1e10, 1e14, True),
    ("alpha", 2e9, 6e9, False),
    ("rho", 1e22, 1e24, True),
    ("fill", 0.1, 1.0, False),
    ("beta", 0.5, 5.0, False),
)
ISO_RANGES = RANGES[:5] + (("T_iso", 400.0, 600.0, False),)

TL = CurveKind("tl", RANGES, N_STEPS)
ISO = CurveKind("iso", ISO_RANGES, ISO_STEPS, tuple(int(v) for v in iso_edges()), trapped=True)
KINDS = {"tl": TL, "iso": ISO}

# the TL kind's functions under their module names
LOW, HIGH, LOGS = TL.low, TL.high, TL.logs
encode, decode, sample_params = TL.encode, TL.decode, TL.sample_params
tl_run, tl_table, mc_curves = TL.run, TL.table, TL.mc_curves


def normalise(c: np.ndarray, c0: float = C0) -> np.ndarray:
    return (np.log1p(np.asarray(c, np.float64) / c0) / math.log1p(1.0 / c0)).astype(np.float32)


def denormalise(y: np.ndarray, c0: float = C0) -> np.ndarray:
    return c0 * np.expm1(np.asarray(y, np.float64) * math.log1p(1.0 / c0))


def nrmse(pred: np.ndarray, ref: np.ndarray) -> np.ndarray:
    """Root-mean-square difference of two curves over the bins, divided by the
    reference curve's peak (per row)."""
    pred = np.atleast_2d(pred)
    ref = np.atleast_2d(ref)
    rms = np.sqrt(np.mean((pred - ref) ** 2, axis=1))
    return rms / np.maximum(ref.max(axis=1), 1e-30)


def default_params(desc: MlpDesc, seed: int = 0) -> np.ndarray:
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights and biases, flat layout."""
    rng = np.random.default_rng(seed)
    dims = [desc.n_in] + [desc.n_hidden] * desc.n_hidden_layers + [desc.n_out]
    parts = []
    for i in range(len(dims) - 1):
        lim = 1.0 / math.sqrt(dims[i])
        parts.append(rng.uniform(-lim, lim, dims[i + 1] * dims[i]))
        parts.append(rng.uniform(-lim, lim, dims[i + 1]))
    return np.concatenate(parts).astype(np.float32)


class Network:
    """The GPU network (BF16 weights and activations, FP32 accumulation and master
    weights).  Batch tensors are torch CUDA float32 tensors; the library reads and
    writes them in place on the context's stream."""

    def __init__(self, ctx: Context, n_in: int = N_IN, n_hidden: int = 512,
                 n_hidden_layers: int = 4, n_out: int = N_BINS, max_batch: int = 4096,
                 params: Optional[np.ndarray] = None, seed: int = 0, c0: float = C0,
                 kind: CurveKind = TL):
        self.ctx = ctx
        self.c0 = float(c0)
        self.kind = kind
        self.lib = ctx.lib
        self.desc = MlpDesc(n_in, n_hidden, n_hidden_layers, n_out, max_batch)
        h = C.c_void_p()
        ctx.check(self.lib.mcl_mlp_create(ctx.handle, C.byref(self.desc), C.byref(h)),
                  "mcl_mlp_create")
        self.handle = h
        self._torch_stream = None
        self.n_params = int(self.lib.mcl_mlp_param_count(h))
        self.set_params(default_params(self.desc, seed) if params is None else params)

    def close(self) -> None:
        """Free the network after the work queued on its own stream: the destroy
        synchronises the stream the context is set to, so it runs on that stream."""
        if getattr(self, "handle", None):
            if self._torch_stream is not None:
                with self._on_stream():
                    self.lib.mcl_mlp_destroy(self.handle)
            else:
                self.lib.mcl_mlp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self) -> np.ndarray:
        p = np.empty(self.n_params, np.float32)
        self.ctx.check(self.lib.mcl_mlp_get_params(self.handle, p.ctypes.data), "mcl_mlp_get_params")
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {p.size}")
        self.ctx.check(self.lib.mcl_mlp_set_params(self.handle, p.ctypes.data), "mcl_mlp_set_params")

    @contextlib.contextmanager
    def _on_stream(self, *tensors):
        """Run library calls on the network's own torch stream, ordered after the
        caller's current stream and before its later work."""
        import torch
        if self._torch_stream is None:
            self._torch_stream = torch.cuda.Stream()
        caller = torch.cuda.current_stream()
        s = self._torch_stream
        s.wait_stream(caller)
        for t in tensors:
            if t is not None:
                t.record_stream(s)
        self.ctx.set_stream(s.cuda_stream)
        try:
            with torch.cuda.stream(s):
                yield
        finally:
            self.ctx.set_stream(None)
            caller.wait_stream(s)

    def forward(self, x):
        """x [B, n_in] float32 CUDA tensor -> [B, n_out] float32 (B padded to 64 inside)."""
        import torch
        B = x.shape[0]
        Bp = -(-B // 64) * 64
        if Bp > self.desc.max_batch:
            raise ValueError(f"batch {B} exceeds max_batch {self.desc.max_batch}")
        xin = x.contiguous().float()
        if Bp != B:
            xin = torch.cat([xin, xin.new_zeros(Bp - B, x.shape[1])])
        y = torch.empty((Bp, self.desc.n_out), dtype=torch.float32, device=x.device)
        with self._on_stream(xin, y):
            self.ctx.check(self.lib.mcl_mlp_forward(self.handle, xin.data_ptr(), Bp, y.data_ptr()),
                           "mcl_mlp_forward")
        return y[:B]

    def grad(self, x, t, idx=None):
        """(loss, flat gradient) of the mean squared error on rows idx (or all)."""
        import torch
        B = x.shape[0] if idx is None else idx.shape[0]
        g = torch.empty(self.n_params, dtype=torch.float32, device=x.device)
        loss = torch.empty(1, dtype=torch.float32, device=x.device)
        with self._on_stream(x, t, idx, g, loss):
            self.ctx.check(self.lib.mcl_mlp_grad(self.handle, x.data_ptr(), t.data_ptr(),
                                                 None if idx is None else idx.data_ptr(), B,
                                                 g.data_ptr(), loss.data_ptr()), "mcl_mlp_grad")
        return float(loss.item()), g

    def train_step(self, x, t, idx, hp: AdamW, loss_out=None) -> None:
        """One AdamW step on rows idx (int32 CUDA tensor) of x / t; the batch loss
        goes to loss_out (a one-element CUDA tensor) without a host sync."""
        B = idx.shape[0]
        with self._on_stream(x, t, idx, loss_out):
            self.ctx.check(self.lib.mcl_mlp_train_step(
                self.handle, x.data_ptr(), t.data_ptr(), idx.data_ptr(), B, C.byref(hp),
                None if loss_out is None else loss_out.data_ptr()), "mcl_mlp_train_step")

    def last_ms(self) -> float:
        v = C.c_double()
        self.ctx.check(self.lib.mcl_mlp_last_ms(self.handle, C.byref(v)), "mcl_mlp_last_ms")
        return v.value

    def save(self, path: Path) -> None:
        np.savez_compressed(path, params=self.get_params(),
                            desc=np.array([self.desc.n_in, self.desc.n_hidden,
                                           self.desc.n_hidden_layers, self.desc.n_out]),
                            c0=np.array(self.c0), kind=np.array(self.kind.name))

    @staticmethod
    def load(ctx: Context, path: Path, max_batch: int = 65536) -> "Network":
        z = np.load(path)                      # plain arrays: no pickle
        n_in, n_hidden, n_layers, n_out = (int(v) for v in z["desc"])
        kind = KINDS[str(z["kind"])] if "kind" in z.files else TL
        return Network(ctx, n_in, n_hidden, n_layers, n_out, max_batch, params=z["params"],
                       c0=float(z["c0"]), kind=kind)


@dataclass
class TrainConfig:
    epochs: int = 60
    batch: int = 4096
    lr: float = 2e-3
    lr_min: float = 5e-5
    warmup_steps: int = 100
    beta1: float = 0.9
    beta2: float = 0.999
    eps: float = 1e-8
    weight_decay: float = 1e-4
    seed: int = 0


def train(net: Network, x, t, cfg: TrainConfig = TrainConfig(),
          log: Optional[Callable[[str], None]] = print) -> dict:
    """Train on device tensors x [N, n_in] (encoded inputs) and t [N, n_out]
    (normalised curves): shuffled batches, linear warm-up then cosine decay of the
    learning rate.  Returns the loss history and device time per step."""
    import torch
    N = x.shape[0]
    steps_per_epoch = N // cfg.batch
    if steps_per_epoch < 1:
        raise ValueError(f"{N} samples are fewer than one batch of {cfg.batch}")
    total = steps_per_epoch * cfg.epochs
    gen = torch.Generator(device="cpu").manual_seed(cfg.seed)
    loss_buf = torch.zeros(total, dtype=torch.float32, device=x.device)
    hp = AdamW(cfg.lr, cfg.beta1, cfg.beta2, cfg.eps, cfg.weight_decay, 0)
    step = 0
    t0 = time.time()
    ms = []
    for ep in range(cfg.epochs):
        perm = torch.randperm(N, generator=gen).to(torch.int32).to(x.device)
        for b in range(steps_per_epoch):
            step += 1
            if step <= cfg.warmup_steps:
                lr = cfg.lr * step / cfg.warmup_steps
            else:
                f = (step - cfg.warmup_steps) / max(1, total - cfg.warmup_steps)
                lr = cfg.lr_min + 0.5 * (cfg.lr - cfg.lr_min) * (1.0 + math.cos(math.pi * f))
            hp.lr = lr
            hp.step = step
            idx = perm[b * cfg.batch:(b + 1) * cfg.batch]
            net.train_step(x, t, idx, hp, loss_buf[step - 1:step])
        ms.append(net.last_ms())
        if log and (ep % max(1, cfg.epochs // 10) == 0 or ep == cfg.epochs - 1):
            torch.cuda.synchronize()
            recent = loss_buf[max(0, step - steps_per_epoch):step].mean().item()
            log(f"epoch {ep + 1}/{cfg.epochs}: loss {recent:.3e} ({time.time() - t0:.1f} s)")
    torch.cuda.synchronize()
    return {"loss": loss_buf.cpu().numpy(), "steps": total, "step_ms_last": ms,
            "wall_s": time.time() - t0}


def predict_curves(net: Network, theta: np.ndarray) -> np.ndarray:
    """Surrogate curves [N, 256] (fractions of N_e per bin) for physical params, in
    the network's own kind of curve."""
    import torch
    x = torch.from_numpy(net.kind.encode(theta)).cuda()
    out = []
    for c0 in range(0, x.shape[0], net.desc.max_batch):
        out.append(net.forward(x[c0:c0 + net.desc.max_batch]).cpu().numpy())
    return denormalise(np.concatenate(out), net.c0)


__all__ = ["Network", "TrainConfig", "train", "encode", "decode", "sample_params", "normalise",
           "denormalise", "nrmse", "mc_curves", "tl_run", "predict_curves", "RANGES", "MclError",
           "CurveKind", "TL", "ISO", "KINDS", "iso_edges"]
