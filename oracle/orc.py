"""ctypes binding of liborc.so (oracle/mcl_oracle.c): the CPU replay harness.

TEST INFRASTRUCTURE ONLY — the checker for the HIP kernels, never the product.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path
from typing import Sequence

import numpy as np

from mcluminescence_amd._abi import DEFAULT_SEED, Geometry, RunParams, run_array

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liborc.so"

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        L.orc_philox4x32_10.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_philox4x32_10.restype = None
        L.orc_expm1f.argtypes = [C.c_float]
        L.orc_expm1f.restype = C.c_float
        L.orc_expf.argtypes = [C.c_float]
        L.orc_expf.restype = C.c_float
        L.orc_exp.argtypes = [C.c_double]
        L.orc_exp.restype = C.c_double
        L.orc_run_geometry.argtypes = [C.POINTER(RunParams), C.POINTER(Geometry)]
        L.orc_run_fixed_dt.argtypes = [C.POINTER(RunParams), C.c_int32, C.c_uint64, C.c_int32,
                                       C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]
        L.orc_init_state.argtypes = [C.POINTER(RunParams), C.c_int32, C.c_int32, C.c_uint64,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_final_state.argtypes = [C.POINTER(RunParams), C.c_int32, C.c_int32, C.c_uint64,
                                      C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_bruteforce_nn.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                        C.c_void_p]
        L.orc_bruteforce_nn.restype = None
        _lib = L
    return _lib


def philox(ctr, key) -> np.ndarray:
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(c.ctypes.data, k.ctypes.data, out.ctypes.data)
    return out


def expm1f(x: float) -> np.float32:
    return np.float32(lib().orc_expm1f(float(x)))


def expf(x: float) -> np.float32:
    return np.float32(lib().orc_expf(float(x)))


def run_event(runs: Sequence[RunParams], duration, max_steps: int, seed: int = DEFAULT_SEED,
              threads: int = 0):
    """Replay of the event kernel: (lum, t, e_ratio, steps) like Context.run_event."""
    L = lib()
    L.orc_run_event.argtypes = [C.POINTER(RunParams), C.c_int32, C.c_void_p, C.c_uint64,
                                C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_void_p]
    L.orc_run_event.restype = C.c_int
    arr = run_array(runs)
    total = sum(int(r.sims) for r in runs)
    dur = np.ascontiguousarray(np.asarray(duration, dtype=np.float64))
    lum = np.zeros((total, max_steps), dtype=np.uint8)
    t = np.zeros((total, max_steps), dtype=np.float64)
    er = np.zeros((total, max_steps), dtype=np.float64)
    steps = np.zeros(total, dtype=np.int32)
    st = L.orc_run_event(arr, len(runs), dur.ctypes.data, C.c_uint64(seed), int(max_steps),
                         int(threads), lum.ctypes.data, t.ctypes.data, er.ctypes.data,
                         steps.ctypes.data)
    if st != 0:
        raise RuntimeError(f"orc_run_event failed: {st}")
    return lum, t, er, steps


def run_lab(runs: Sequence[RunParams], observations, exact_fill: bool = False,
            seed: int = DEFAULT_SEED, max_steps: int = 10 ** 7, threads: int = 0,
            options: int = 0):
    """Replay of the lab kernel: (ratio [clusters, max_obs], status).  options: the
    MCL_LAB_* bits (exact_fill sets MCL_LAB_EXACT_FILL)."""
    L = lib()
    L.orc_run_lab.argtypes = [C.POINTER(RunParams), C.c_int32, C.c_void_p, C.c_void_p, C.c_uint32,
                              C.c_uint64, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p]
    L.orc_run_lab.restype = C.c_int
    arr = run_array(runs)
    counts = np.array([len(o) for o in observations], dtype=np.int32)
    obs = np.ascontiguousarray(np.concatenate([np.asarray(o, np.float64) for o in observations]))
    total = sum(int(r.sims) for r in runs)
    ratio = np.zeros((total, int(counts.max())), dtype=np.float64)
    status = np.zeros(total, dtype=np.int32)
    opts = int(options) | (1 if exact_fill else 0)
    st = L.orc_run_lab(arr, len(runs), obs.ctypes.data, counts.ctypes.data, opts,
                       C.c_uint64(seed), int(max_steps), int(threads), ratio.ctypes.data,
                       status.ctypes.data)
    if st != 0:
        raise RuntimeError(f"orc_run_lab failed: {st}")
    return ratio, status


def log64(x: float) -> float:
    L = lib()
    L.orc_log.argtypes = [C.c_double]
    L.orc_log.restype = C.c_double
    return float(L.orc_log(float(x)))


def fire_threshold(x: float) -> int:
    """T(x) = (uint32)((1 - e^{-x}) 2^32), or 2^32 when p rounds to 1."""
    L = lib()
    L.orc_fire_threshold.argtypes = [C.c_float]
    L.orc_fire_threshold.restype = C.c_uint64
    return int(L.orc_fire_threshold(float(x)))


def thr_monotone_scan(lo: float, hi: float):
    """(violations, largest drop) of the float32 fire threshold over every float
    in [lo, hi] taken in increasing order."""
    L = lib()
    L.orc_thr_monotone_scan.argtypes = [C.c_float, C.c_float, C.POINTER(C.c_uint64)]
    L.orc_thr_monotone_scan.restype = C.c_uint64
    worst = C.c_uint64()
    bad = L.orc_thr_monotone_scan(float(lo), float(hi), C.byref(worst))
    return int(bad), int(worst.value)


def exp64(x: float) -> float:
    return float(lib().orc_exp(float(x)))


def geometry(run: RunParams) -> Geometry:
    g = Geometry()
    lib().orc_run_geometry(C.byref(run), C.byref(g))
    return g


def geometry_legacy(run: RunParams) -> Geometry:
    """The legacy stack's box (functions.py:51-80)."""
    L = lib()
    L.orc_run_geometry_legacy.argtypes = [C.POINTER(RunParams), C.POINTER(Geometry)]
    g = Geometry()
    L.orc_run_geometry_legacy(C.byref(run), C.byref(g))
    return g


def run_fixed_dt(runs: Sequence[RunParams], seed: int = DEFAULT_SEED, n_bins=None,
                 fp64: bool = False, threads: int = 0):
    """Replay harness: counts [clusters, n_bins] uint32 and live_steps [clusters] uint64."""
    arr = run_array(runs)
    if n_bins is None:
        n_bins = 0
    total = sum(int(r.sims) for r in runs)
    row_len = n_bins if n_bins > 0 else max(int(r.n_steps) for r in runs if r.sims > 0)
    counts = np.zeros((total, row_len), np.uint32)
    live = np.zeros(total, np.uint64)
    st = lib().orc_run_fixed_dt(arr, len(runs), C.c_uint64(seed), int(n_bins), int(bool(fp64)),
                                int(threads), counts.ctypes.data, live.ctypes.data)
    if st != 0:
        raise ValueError(f"oracle rejected the runs (status {st})")
    return counts, live


def init_state(run: RunParams, run_index: int, sim: int, seed: int = DEFAULT_SEED) -> dict:
    g = geometry(run)
    ne, nh = g.n_e0, g.n_h
    e_xyz = np.zeros((ne, 3), np.float32)
    h_xyz = np.zeros((nh, 3), np.float32)
    nn = np.zeros(ne, np.uint32)
    w = np.zeros(ne, np.float32)
    lib().orc_init_state(C.byref(run), run_index, sim, C.c_uint64(seed), e_xyz.ctypes.data,
                         h_xyz.ctypes.data, nn.ctypes.data, w.ctypes.data)
    return dict(e_xyz=e_xyz, h_xyz=h_xyz, nn=nn, w=w)


def final_state(run: RunParams, run_index: int, sim: int, max_steps: int,
                seed: int = DEFAULT_SEED) -> dict:
    g = geometry(run)
    ne, nh = g.n_e0, g.n_h
    counts = np.zeros(int(run.n_steps), np.uint32)
    e_live = np.zeros(ne, np.uint8)
    h_alive = np.zeros(nh, np.uint8)
    nn = np.zeros(ne, np.uint32)
    e_xyz = np.zeros((ne, 3), np.float32)
    h_xyz = np.zeros((nh, 3), np.float32)
    st = lib().orc_final_state(C.byref(run), run_index, sim, C.c_uint64(seed), int(max_steps),
                               int(run.n_steps), counts.ctypes.data, e_live.ctypes.data,
                               h_alive.ctypes.data, nn.ctypes.data, e_xyz.ctypes.data,
                               h_xyz.ctypes.data)
    if st != 0:
        raise ValueError(f"oracle rejected the run (status {st})")
    return dict(counts=counts, e_live=e_live, h_alive=h_alive, nn=nn, e_xyz=e_xyz, h_xyz=h_xyz)


def bruteforce_nn(e_xyz: np.ndarray, h_xyz: np.ndarray, h_alive: np.ndarray) -> np.ndarray:
    e = np.ascontiguousarray(e_xyz, np.float32)
    h = np.ascontiguousarray(h_xyz, np.float32)
    a = np.ascontiguousarray(h_alive, np.uint8)
    nn = np.zeros(len(e), np.uint32)
    lib().orc_bruteforce_nn(e.ctypes.data, len(e), h.ctypes.data, a.ctypes.data, len(h),
                            nn.ctypes.data)
    return nn
