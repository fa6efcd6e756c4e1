"""Lab-comparison objectives (the reference's TLTrapSim.TL_lab / ISO_lab,
src/class/tl_trap_lab.py:65-179): dose filling against data/processed tables."""
from __future__ import annotations


def tl_lab(sim, csv_name: str, *, plot: bool = False) -> float:
    raise NotImplementedError(
        "TL_lab needs dose filling (lambda = (D/D0)(N_e - n_e), tl_trap_lab.py:53-60), "
        "which the fixed-dt kernel does not run yet")


def iso_lab(sim, csv_name: str, *, plot: bool = False) -> float:
    raise NotImplementedError(
        "ISO_lab needs dose filling (lambda = (D/D0)(N_e - n_e), tl_trap_lab.py:53-60), "
        "which the fixed-dt kernel does not run yet")
