"""ctypes binding of libmcl.so, the C ABI declared in include/mcl.h and
include/mcl_surrogate.h.

This is the only way the package reaches the GPU.  There is no CPU fallback:
if the library is missing or no HIP device is present, ``Library()`` and
``Context()`` raise, naming what is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path
from typing import List, Optional, Sequence, Tuple

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libmcl.so"

ABI_VERSION = 1

MCL_OK = 0
MCL_EDEVICE = -5
MCL_ENOMEM = -12
MCL_EINVAL = -22
MCL_ERANGE = -34
MCL_EOVERFLOW = -75

F_DEVICE_PTRS = 0x1
F_PERSISTENT = 0x2
F_GRAPH = 0x4

MAX_GLOW_BINS = 8192          # MCL_GLOW_MAX_BINS

# mcl_run_lab options (MCL_LAB_*)
LAB_EXACT_FILL = 0x1
LAB_SHARED_CHANNEL = 0x2
LAB_RETRAP = 0x4
LAB_FIRST_ELECTRON = 0x8
LAB_LEGACY_GEOMETRY = 0x10
LAB_LEGACY = 0x1F

TAG_GEOM_E = 0
TAG_GEOM_H = 1
TAG_FIRE = 2
TAG_CHAN = 3
NO_HOLE = 0xFFFF

DEFAULT_SEED = 0x4D434C554D000001


class RunParams(C.Structure):
    """mcl_run_params: one parameter set (run) and its number of clusters (sims)."""

    _fields_ = [
        ("alpha", C.c_double), ("b", C.c_double), ("s", C.c_double), ("E_cb", C.c_double),
        ("E_loc_1", C.c_double), ("E_loc_2", C.c_double), ("D0", C.c_double),
        ("Retrap", C.c_double), ("k_b", C.c_double), ("D", C.c_double),
        ("rho_prime", C.c_double), ("holes", C.c_double), ("boundary_factor", C.c_double),
        ("N_e", C.c_double), ("e_ratio_start", C.c_double),
        ("T_start", C.c_double), ("T_rate", C.c_double), ("dt", C.c_double),
        ("n_steps", C.c_int32), ("sims", C.c_int32),
    ]

    FLOAT_FIELDS = ("alpha", "b", "s", "E_cb", "E_loc_1", "E_loc_2", "D0", "Retrap", "k_b", "D",
                    "rho_prime", "holes", "boundary_factor", "N_e", "e_ratio_start",
                    "T_start", "T_rate", "dt")

    @classmethod
    def make(cls, **kw) -> "RunParams":
        p = cls()
        for k, v in kw.items():
            if not hasattr(p, k):
                raise TypeError(f"unknown run parameter {k!r}")
            setattr(p, k, v)
        return p

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class Geometry(C.Structure):
    _fields_ = [("rho", C.c_double), ("side_m", C.c_double), ("n_e0", C.c_int32),
                ("n_h", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("live_electron_steps", C.c_uint64),
                ("recombinations", C.c_uint64), ("clusters", C.c_uint64),
                ("cluster_steps", C.c_uint64)]


class PlanInfo(C.Structure):
    _fields_ = [("total_clusters", C.c_int64), ("n_bins", C.c_int32), ("row_len", C.c_int32),
                ("lds_bytes", C.c_int32),
                ("max_electrons", C.c_int32), ("max_holes", C.c_int32), ("grid", C.c_int32),
                ("block", C.c_int32)]


class Work(C.Structure):
    _fields_ = [("live_electron_steps", C.c_uint64), ("distance_evals", C.c_uint64),
                ("recombinations", C.c_uint64), ("cluster_steps", C.c_uint64),
                ("kernel_ms", C.c_double)]


class MlpDesc(C.Structure):
    """mcl_mlp_desc (include/mcl_surrogate.h)."""
    _fields_ = [("n_in", C.c_int32), ("n_hidden", C.c_int32), ("n_hidden_layers", C.c_int32),
                ("n_out", C.c_int32), ("max_batch", C.c_int32)]


class AdamW(C.Structure):
    """mcl_adamw (include/mcl_surrogate.h)."""
    _fields_ = [("lr", C.c_double), ("beta1", C.c_double), ("beta2", C.c_double),
                ("eps", C.c_double), ("weight_decay", C.c_double), ("step", C.c_int64)]


class DeParams(C.Structure):
    """mcl_de_params (include/mcl_surrogate.h)."""
    _fields_ = [("pop", C.c_int32), ("dim", C.c_int32), ("F", C.c_float), ("CR", C.c_float),
                ("seed", C.c_uint64), ("fixed", C.c_float * 8), ("log_span", C.c_float),
                ("out_scale", C.c_float), ("inv_peak2", C.c_float), ("flags", C.c_int32)]


DE_LAYERED = 1   # mcl_de_params.flags: a generation as separate kernels


class MclError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"mcl error {code}: {msg}")
        self.code = code


_P = C.c_void_p
_SIGNATURES = {
    "mcl_abi_version": (C.c_int, []),
    "mcl_create": (C.c_int, [C.POINTER(_P), C.c_int]),
    "mcl_destroy": (None, [_P]),
    "mcl_last_error": (C.c_char_p, [_P]),
    "mcl_set_stream": (C.c_int, [_P, _P]),
    "mcl_cluster_limits": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "mcl_tune_set": (C.c_int, [_P, C.c_char_p, C.c_double]),
    "mcl_tune_get": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_double)]),
    "mcl_run_geometry": (C.c_int, [C.POINTER(RunParams), C.POINTER(Geometry)]),
    "mcl_run_fixed_dt": (C.c_int, [_P, C.POINTER(RunParams), C.c_int32, C.c_uint64, C.c_int32,
                                   C.c_uint32, _P, _P, C.POINTER(Stats)]),
    "mcl_plan_fixed_dt": (C.c_int, [_P, C.POINTER(RunParams), C.c_int32, C.c_uint64, C.c_int32,
                                    C.c_uint32, C.POINTER(_P)]),
    "mcl_plan_info_get": (C.c_int, [_P, C.POINTER(PlanInfo)]),
    "mcl_plan_launch": (C.c_int, [_P, _P, _P]),
    "mcl_plan_last_ms": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "mcl_plan_work": (C.c_int, [_P, C.POINTER(Work)]),
    "mcl_plan_destroy": (None, [_P]),
    "mcl_sync": (C.c_int, [_P]),
    "mcl_cluster_state": (C.c_int, [_P, C.POINTER(RunParams), C.c_int32, C.c_int32, C.c_uint64,
                                    C.c_int32, _P, _P, _P, _P, _P, _P]),
    "mcl_run_cluster": (C.c_int, [_P, C.POINTER(RunParams), C.c_int32, C.c_int32, C.c_uint64,
                                  C.c_int32, _P, _P]),
    "mcl_bin_counts": (C.c_int, [_P, _P, C.c_int64, C.c_int32, C.c_int64, _P, C.c_int32, _P]),
    "mcl_run_fixed_dt_noalloc": (C.c_int, [_P, C.POINTER(RunParams), C.c_int32, C.c_uint64,
                                           C.c_int32, _P, _P, C.POINTER(Stats)]),
    "mcl_run_event": (C.c_int, [_P, C.POINTER(RunParams), C.c_int32, _P, C.c_uint64, C.c_int32,
                                C.c_uint32, _P, _P, _P, _P, C.POINTER(Stats)]),
    "mcl_run_lab": (C.c_int, [_P, C.POINTER(RunParams), C.c_int32, _P, _P, C.c_uint32, C.c_uint64,
                              C.c_int64, _P, _P, C.POINTER(Stats)]),
    "mcl_glow_curve": (C.c_int, [_P, C.POINTER(RunParams), C.c_int32, _P, C.c_int32, C.c_int64,
                                 _P, C.c_double, C.c_double, C.c_uint32, _P, C.c_int64, _P,
                                 C.POINTER(Stats)]),
    # surrogate network (include/mcl_surrogate.h)
    "mcl_mlp_create": (C.c_int, [_P, C.POINTER(MlpDesc), C.POINTER(_P)]),
    "mcl_mlp_destroy": (None, [_P]),
    "mcl_mlp_param_count": (C.c_int64, [_P]),
    "mcl_mlp_get_params": (C.c_int, [_P, _P]),
    "mcl_mlp_set_params": (C.c_int, [_P, _P]),
    "mcl_mlp_forward": (C.c_int, [_P, _P, C.c_int32, _P]),
    "mcl_mlp_grad": (C.c_int, [_P, _P, _P, _P, C.c_int32, _P, _P]),
    "mcl_mlp_train_step": (C.c_int, [_P, _P, _P, _P, C.c_int32, C.POINTER(AdamW), _P]),
    "mcl_mlp_last_ms": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "mcl_de_create": (C.c_int, [_P, C.POINTER(DeParams), _P, C.POINTER(_P)]),
    "mcl_de_destroy": (None, [_P]),
    "mcl_de_run": (C.c_int, [_P, C.c_int32]),
    "mcl_de_best": (C.c_int, [_P, _P, _P]),
    "mcl_de_population": (C.c_int, [_P, _P, _P]),
}

EXPORTED = tuple(_SIGNATURES)

_lib: Optional[C.CDLL] = None


def library() -> C.CDLL:
    """Load libmcl.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build the HIP library first "
            "(python -m mcluminescence_amd.csrc.build); there is no CPU fallback")
    # PyTorch-ROCm ships its own libamdhip64 and links it by the unversioned
    # name.  Loading torch first makes libmcl.so bind to that same runtime (same
    # SONAME); loading libmcl.so first would let a later torch import start a
    # second HIP runtime in the process, which then sees no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.mcl_abi_version() != ABI_VERSION:
        raise RuntimeError(f"libmcl ABI {lib.mcl_abi_version()} != expected {ABI_VERSION}")
    _lib = lib
    return lib


# numpy view of mcl_run_params: a structured table of runs, built column-wise, is
# handed to the library as a ctypes array without a per-run Python object
RUN_DTYPE = np.dtype([(name, np.float64 if ct is C.c_double else np.int32)
                      for name, ct in RunParams._fields_])
assert RUN_DTYPE.itemsize == C.sizeof(RunParams)


def run_table(n: int) -> np.ndarray:
    """A zeroed table of n runs (structured array, RUN_DTYPE)."""
    return np.zeros(n, dtype=RUN_DTYPE)


def run_array(runs):
    """ctypes array of mcl_run_params from a list of RunParams, a RUN_DTYPE table
    (copied) or a ctypes array (passed through)."""
    if isinstance(runs, C.Array):
        return runs
    if isinstance(runs, np.ndarray):
        if runs.dtype != RUN_DTYPE:
            raise TypeError("run table must have dtype RUN_DTYPE")
        return (RunParams * len(runs)).from_buffer_copy(np.ascontiguousarray(runs))
    arr = (RunParams * len(runs))()
    for i, r in enumerate(runs):
        arr[i] = r
    return arr


def geometry(run: RunParams) -> Geometry:
    g = Geometry()
    st = library().mcl_run_geometry(C.byref(run), C.byref(g))
    if st != MCL_OK:
        raise MclError(st, "mcl_run_geometry")
    return g


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class Context:
    """A libmcl context: one HIP device and stream."""

    def __init__(self, device: int = 0):
        self.lib = library()
        h = C.c_void_p()
        st = self.lib.mcl_create(C.byref(h), int(device))
        if st != MCL_OK:
            raise MclError(st, f"mcl_create(device={device}) failed: no usable HIP device")
        self.handle = h

    def sync(self) -> None:
        """Wait for the work queued on the context's stream."""
        self.check(self.lib.mcl_sync(self.handle), "mcl_sync")

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.mcl_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, st: int, what: str) -> None:
        if st != MCL_OK:
            msg = self.lib.mcl_last_error(self.handle)
            raise MclError(st, f"{what}: {msg.decode() if msg else ''}")

    def set_stream(self, stream_handle: Optional[int]) -> None:
        self.check(self.lib.mcl_set_stream(self.handle, stream_handle or None), "mcl_set_stream")

    def tune_set(self, key: str, value: float) -> None:
        self.check(self.lib.mcl_tune_set(self.handle, key.encode(), float(value)), "mcl_tune_set")

    def tune_get(self, key: str) -> float:
        v = C.c_double()
        self.check(self.lib.mcl_tune_get(self.handle, key.encode(), C.byref(v)), "mcl_tune_get")
        return v.value

    def cluster_limits(self):
        e, h = C.c_int32(), C.c_int32()
        self.check(self.lib.mcl_cluster_limits(self.handle, C.byref(e), C.byref(h)),
                   "mcl_cluster_limits")
        return e.value, h.value

    def run_fixed_dt(self, runs: Sequence[RunParams], seed: int = DEFAULT_SEED,
                     n_bins: Optional[int] = None, flags: int = 0):
        """Run every cluster of every run; returns (counts [clusters, row_len] uint32,
        live_steps [clusters] uint64, Stats).  n_bins=None or 0: per-step counts,
        rows as long as the longest run."""
        arr = run_array(runs)
        if n_bins is None:
            n_bins = 0
        if isinstance(runs, np.ndarray):
            total = int(runs["sims"].sum())
            row_len = n_bins if n_bins > 0 else int(runs["n_steps"][runs["sims"] > 0].max())
        else:
            total = sum(int(r.sims) for r in runs)
            row_len = n_bins if n_bins > 0 else max(int(r.n_steps) for r in runs if r.sims > 0)
        counts = np.zeros((total, row_len), dtype=np.uint32)
        live = np.zeros(total, dtype=np.uint64)
        stats = Stats()
        flags = int(flags) & ~F_DEVICE_PTRS   # host arrays here
        st = self.lib.mcl_run_fixed_dt(self.handle, arr, len(runs), C.c_uint64(seed), int(n_bins),
                                       flags, _ptr(counts), _ptr(live), C.byref(stats))
        self.check(st, "mcl_run_fixed_dt")
        return counts, live, stats

    def run_event(self, runs: Sequence[RunParams], duration: Sequence[float], max_steps: int,
                  seed: int = DEFAULT_SEED):
        """Event-mode Monte Carlo (one event per step; each run's dt field is max_dt).
        Returns lum [clusters, max_steps] uint8, t [clusters, max_steps] float64,
        e_ratio [clusters, max_steps] float64, steps [clusters] int32, Stats.  Raises
        MclError with status MCL_EOVERFLOW when a cluster needs more than max_steps."""
        arr = run_array(runs)
        total = sum(int(r.sims) for r in runs)
        dur = np.ascontiguousarray(np.asarray(duration, dtype=np.float64))
        if dur.size != len(runs):
            raise ValueError("one duration per run")
        lum = np.zeros((total, max_steps), dtype=np.uint8)
        t = np.zeros((total, max_steps), dtype=np.float64)
        er = np.zeros((total, max_steps), dtype=np.float64)
        steps = np.zeros(total, dtype=np.int32)
        stats = Stats()
        st = self.lib.mcl_run_event(self.handle, arr, len(runs), _ptr(dur), C.c_uint64(seed),
                                    int(max_steps), 0, _ptr(lum), _ptr(t), _ptr(er), _ptr(steps),
                                    C.byref(stats))
        self.check(st, "mcl_run_event")
        return lum, t, er, steps, stats

    def run_lab(self, runs: Sequence[RunParams], observations: Sequence[Sequence[float]],
                exact_fill: bool = False, seed: int = DEFAULT_SEED, max_steps: int = 10 ** 7,
                options: int = 0):
        """Lab experiments with dose filling (one run per experiment, its T_start,
        T_rate, D and e_ratio_start).  options: LAB_* bits (exact_fill sets
        LAB_EXACT_FILL).  Returns (ratio [clusters, max_obs] float64: n_e / N_e at
        each observation time, status [clusters] int32, Stats)."""
        arr = run_array(runs)
        if len(observations) != len(runs):
            raise ValueError("one observation list per run")
        counts = np.array([len(o) for o in observations], dtype=np.int32)
        obs = np.ascontiguousarray(np.concatenate([np.asarray(o, np.float64) for o in observations]))
        total = sum(int(r.sims) for r in runs)
        ratio = np.zeros((total, int(counts.max())), dtype=np.float64)
        status = np.zeros(total, dtype=np.int32)
        stats = Stats()
        opts = int(options) | (LAB_EXACT_FILL if exact_fill else 0)
        st = self.lib.mcl_run_lab(self.handle, arr, len(runs), _ptr(obs), _ptr(counts),
                                  opts, C.c_uint64(seed), int(max_steps),
                                  _ptr(ratio), _ptr(status), C.byref(stats))
        self.check(st, "mcl_run_lab")
        return ratio, status, stats

    def glow_curve(self, runs: Sequence[RunParams], counts: np.ndarray,
                   t_axis: Optional[np.ndarray] = None, bin_width: float = 1.0,
                   win: float = 50.0) -> Tuple[List[np.ndarray], Stats]:
        """Glow curves (mcl_glow_curve) from host arrays: counts [clusters, steps]
        uint32 (fixed-dt counts) or uint8 (event lum), t_axis [clusters, steps]
        float64 or None (fixed-dt axis (i + 1) dt of each run).  Returns one float64
        curve per run and the Stats of the last launch."""
        arr = run_array(runs)
        counts = np.ascontiguousarray(counts)
        if counts.dtype not in (np.uint32, np.uint8) or counts.ndim != 2:
            raise ValueError("counts must be a 2-D uint32 or uint8 array")
        total = sum(int(r.sims) for r in runs)
        if counts.shape[0] != total:
            raise ValueError(f"counts has {counts.shape[0]} rows for {total} clusters")
        stride = counts.shape[1]
        if t_axis is not None:
            t_axis = np.ascontiguousarray(t_axis, dtype=np.float64)
            if t_axis.shape != counts.shape:
                raise ValueError("t_axis must have the shape of counts")
        n = len(runs)
        # first guess from the temperature span; a longer curve is reported and retried
        span = max(abs(float(r.T_start)) + abs(float(r.T_rate)) * float(r.n_steps) * float(r.dt)
                   for r in runs)
        k = max(1, int(win / bin_width)) if bin_width > 0 else 1
        guess = int(min(span / bin_width, 4 * MAX_GLOW_BINS)) + k + 4 if bin_width > 0 else 16
        for _ in range(2):
            curve = np.zeros((n, max(guess, 1)), dtype=np.float64)
            lens = np.zeros(n, dtype=np.int32)
            stats = Stats()
            st = self.lib.mcl_glow_curve(self.handle, arr, n, _ptr(counts), counts.itemsize, stride,
                                         _ptr(t_axis) if t_axis is not None else None,
                                         float(bin_width), float(win), 0, _ptr(curve),
                                         curve.shape[1], _ptr(lens), C.byref(stats))
            if st == MCL_EOVERFLOW:
                guess = int(lens.max())
                continue
            self.check(st, "mcl_glow_curve")
            return [curve[r, :lens[r]].copy() for r in range(n)], stats
        self.check(st, "mcl_glow_curve")
        raise AssertionError("unreachable")

    def glow_curve_device(self, runs: Sequence[RunParams], counts_ptr: int, count_bytes: int,
                          row_stride: int, t_axis_ptr: Optional[int], bin_width: float,
                          win: float, curve_ptr: int, curve_stride: int, len_ptr: int) -> Stats:
        """mcl_glow_curve on device buffers (MCL_F_DEVICE_PTRS); waits for the kernel."""
        arr = run_array(runs)
        stats = Stats()
        st = self.lib.mcl_glow_curve(self.handle, arr, len(runs), C.c_void_p(counts_ptr),
                                     int(count_bytes), int(row_stride),
                                     C.c_void_p(t_axis_ptr) if t_axis_ptr else None,
                                     float(bin_width), float(win), F_DEVICE_PTRS,
                                     C.c_void_p(curve_ptr), int(curve_stride),
                                     C.c_void_p(len_ptr), C.byref(stats))
        self.check(st, "mcl_glow_curve")
        return stats

    def cluster_state(self, run: RunParams, run_index: int, sim: int, seed: int = DEFAULT_SEED,
                      max_steps: int = 0) -> dict:
        g = geometry(run)
        ne, nh = max(g.n_e0, 0), max(g.n_h, 0)
        e_xyz = np.zeros((ne, 3), np.float32)
        h_xyz = np.zeros((nh, 3), np.float32)
        e_live = np.zeros(ne, np.uint8)
        h_alive = np.zeros(nh, np.uint8)
        nn = np.zeros(ne, np.uint32)
        w = np.zeros(ne, np.float32)
        st = self.lib.mcl_cluster_state(self.handle, C.byref(run), int(run_index), int(sim),
                                        C.c_uint64(seed), int(max_steps), _ptr(e_xyz), _ptr(h_xyz),
                                        _ptr(e_live), _ptr(h_alive), _ptr(nn), _ptr(w))
        self.check(st, "mcl_cluster_state")
        return dict(e_xyz=e_xyz, h_xyz=h_xyz, e_live=e_live, h_alive=h_alive, nn=nn, w=w)

    def bin_counts(self, counts_ptr: int, n_sets: int, sims: int, row_stride: int,
                   edges: Sequence[int], out_ptr: int) -> None:
        """mcl_bin_counts on device buffers (asynchronous on the context's stream):
        out[p][j] = sum of counts[p*sims+k][edges[j]:edges[j+1]] over the set's sims."""
        e = np.ascontiguousarray(np.asarray(edges, np.int32))
        st = self.lib.mcl_bin_counts(self.handle, C.c_void_p(counts_ptr), int(n_sets), int(sims),
                                     int(row_stride), _ptr(e), len(e) - 1, C.c_void_p(out_ptr))
        self.check(st, "mcl_bin_counts")

    def run_fixed_dt_noalloc(self, runs: Sequence[RunParams], counts_ptr: int,
                             live_ptr: Optional[int] = None, seed: int = DEFAULT_SEED,
                             n_bins: int = 0) -> "Stats":
        """mcl_run_fixed_dt_noalloc: counts and live steps into caller-owned DEVICE
        buffers (addresses), scratch from the library's static device array."""
        arr = run_array(runs)
        stats = Stats()
        st = self.lib.mcl_run_fixed_dt_noalloc(self.handle, arr, len(runs), C.c_uint64(seed),
                                               int(n_bins), C.c_void_p(counts_ptr),
                                               C.c_void_p(live_ptr) if live_ptr else None,
                                               C.byref(stats))
        self.check(st, "mcl_run_fixed_dt_noalloc")
        return stats

    def run_cluster(self, run: RunParams, run_index: int, sim: int, seed: int = DEFAULT_SEED,
                    n_bins: int = 0):
        """Counts [row_len] uint32 and live electron-steps of cluster (run_index, sim)
        of a sweep holding `run` at run_index, simulating that cluster alone."""
        row = int(n_bins) if n_bins > 0 else int(run.n_steps)
        counts = np.zeros(row, np.uint32)
        live = np.zeros(1, np.uint64)
        st = self.lib.mcl_run_cluster(self.handle, C.byref(run), int(run_index), int(sim),
                                      C.c_uint64(seed), int(n_bins), _ptr(counts), _ptr(live))
        self.check(st, "mcl_run_cluster")
        return counts, int(live[0])

    def plan(self, runs: Sequence[RunParams], seed: int = DEFAULT_SEED,
             n_bins: Optional[int] = None, flags: int = 0) -> "Plan":
        return Plan(self, runs, seed, n_bins, flags)


class Plan:
    """Device-resident parameter sets, launched many times (mcl_plan_*)."""

    def __init__(self, ctx: Context, runs: Sequence[RunParams], seed: int,
                 n_bins: Optional[int], flags: int):
        self.ctx = ctx
        arr = run_array(runs)
        if n_bins is None:
            n_bins = 0
        h = C.c_void_p()
        ctx.check(ctx.lib.mcl_plan_fixed_dt(ctx.handle, arr, len(runs), C.c_uint64(seed),
                                            int(n_bins), int(flags), C.byref(h)),
                  "mcl_plan_fixed_dt")
        self.handle = h
        info = PlanInfo()
        ctx.check(ctx.lib.mcl_plan_info_get(h, C.byref(info)), "mcl_plan_info_get")
        self.info = info

    def launch(self, counts_dev_ptr: int, live_dev_ptr: Optional[int] = None) -> None:
        self.ctx.check(self.ctx.lib.mcl_plan_launch(self.handle, counts_dev_ptr, live_dev_ptr),
                       "mcl_plan_launch")

    def last_ms(self) -> float:
        ms = C.c_double()
        self.ctx.check(self.ctx.lib.mcl_plan_last_ms(self.handle, C.byref(ms)), "mcl_plan_last_ms")
        return ms.value

    def work(self) -> Work:
        w = Work()
        self.ctx.check(self.ctx.lib.mcl_plan_work(self.handle, C.byref(w)), "mcl_plan_work")
        return w

    def sync(self) -> None:
        self.ctx.check(self.ctx.lib.mcl_sync(self.ctx.handle), "mcl_sync")

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.ctx.lib.mcl_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
