"""``TLTrapSim``: the drop-in for the reference's simulation object
(src/class/tl_trap_lab.py:24-60).

The reference's object is driven one event at a time from Python
(``_update_lifetimes`` / ``_filling_time`` / ``box.remove_pair``, called by
simulate.py:51-92).  Here a cluster's whole step loop is one kernel launch, so
the object keeps the constructor and the attributes callers read
(``cfg``, ``mc``, ``phys``, ``box.electrons``, ``box.holes``, ``box.nearest()``)
and adds ``run()``; the per-event private methods have no equivalent.
"""
from __future__ import annotations

from typing import Optional, Tuple

import numpy as np

from ._abi import DEFAULT_SEED, Context, geometry
from .config import Config
from .engine import Physics, run_params


class ClusterView:
    """The kernel's view of one cluster (the reference's Box, engine.py:83-188):
    coordinates in metres and the nearest live hole of every electron, as the
    kernel holds them after ``step`` steps."""

    def __init__(self, sim: "TLTrapSim"):
        self._sim = sim
        self._state = None
        self._step = 0

    def _fetch(self, step: int = 0):
        if self._state is None or self._step != step:
            s = self._sim
            self._state = s._context().cluster_state(s._params, s.run_index, s.sim, seed=s.seed,
                                                     max_steps=step)
            self._step = step
        return self._state

    @property
    def L(self) -> float:
        return self._sim.d

    @property
    def boundary_factor(self) -> float:
        return float(self._sim.mc.boundary_factor)

    @property
    def electrons(self) -> np.ndarray:
        st = self._fetch(self._step)
        live = st["e_live"].astype(bool)
        return st["e_xyz"][live].astype(np.float64) * 1e-9

    @property
    def holes(self) -> np.ndarray:
        st = self._fetch(self._step)
        alive = st["h_alive"].astype(bool)
        return st["h_xyz"][alive].astype(np.float64) * 1e-9

    def nearest(self) -> Tuple[np.ndarray, np.ndarray]:
        """(distance [m], index into ``holes``) of each live electron's nearest hole."""
        st = self._fetch(self._step)
        live = st["e_live"].astype(bool)
        alive = st["h_alive"].astype(bool)
        remap = np.cumsum(alive) - 1                  # original hole index -> index in holes
        nn = st["nn"][live].astype(np.int64)
        e = st["e_xyz"][live].astype(np.float64)
        h = st["h_xyz"].astype(np.float64)
        d = np.linalg.norm(e - h[np.clip(nn, 0, len(h) - 1)], axis=1) * 1e-9
        return d, remap[np.clip(nn, 0, len(h) - 1)]

    def advance_to(self, step: int) -> "ClusterView":
        self._fetch(step)
        return self


class TLTrapSim:
    """One cluster of one run config: ``TLTrapSim(cfg, e_ratio_start=...)``.

    ``cfg`` holds ``exp_type_fp`` and ``physics_fp`` (one run of
    ``initialize_runs``).  Geometry follows tl_trap_lab.py:33-39:
    rho = rho' 3/(4 pi) alpha^3, d = (holes / rho)^(1/3), int(N_e * e_ratio_start)
    electrons; ``e_ratio_start=None`` means 0, as in the reference.
    """

    def __init__(self, cfg: Config, *, e_ratio_start: Optional[float] = None,
                 ctx: Optional[Context] = None, seed: int = DEFAULT_SEED, run_index: int = 0,
                 sim: int = 0):
        self.cfg = cfg
        self.mc = cfg.exp_type_fp
        self.phys = Physics(**cfg.physics_fp)
        self.e_ratio_start = float(e_ratio_start) if e_ratio_start is not None else 0.0
        self.seed = seed
        self.run_index = run_index
        self.sim = sim
        self._ctx = ctx
        self._params = run_params(cfg, sims=1, e_ratio_start=self.e_ratio_start)
        g = geometry(self._params)
        self.rho = g.rho
        self.d = g.side_m
        self.n_e0 = g.n_e0
        self.n_h = g.n_h
        self.box = ClusterView(self)

    def _context(self) -> Context:
        if self._ctx is None:
            from .simulate import default_context
            self._ctx = default_context()
        return self._ctx

    def run(self, n_steps: Optional[int] = None, dt: Optional[float] = None) -> np.ndarray:
        """Per-step recombination counts of this cluster over its schedule (the
        row (run_index, sim) of the sweep's counts; only this cluster is simulated)."""
        p = run_params(self.cfg, sims=1, e_ratio_start=self.e_ratio_start, n_steps=n_steps, dt=dt)
        counts, _ = self._context().run_cluster(p, self.run_index, self.sim, seed=self.seed)
        return counts

    def TL_lab(self, csv_name: str, *, plot: bool = False, **kw) -> float:
        """MSE against a TL lab table (tl_trap_lab.py:65-118); keywords of lab.tl_lab."""
        from .lab import tl_lab
        return tl_lab(self, csv_name, plot=plot, **kw)

    def ISO_lab(self, csv_name: str, *, plot: bool = False, **kw) -> float:
        """MSE against an isothermal lab table (tl_trap_lab.py:120-179)."""
        from .lab import iso_lab
        return iso_lab(self, csv_name, plot=plot, **kw)
