"""Ensemble fixtures of the legacy lab loops restated in numpy (oracle/legacy_mc.py:
sim_lab_TL_residuals / sim_lab_TL_residuals_iso of src/est_params/functions.py)
for the lab kernel's MCL_LAB_LEGACY options.

Physics and geometry are the shipped lab configs (conf/physics_fp/lab_TL.yaml,
conf/exp_type_fp/TLlab.yaml) with E_loc_2 = 0.9 eV, so that the shared channel
draw matters, and holes = 437.6, where the legacy box (437 + 318 holes) differs
from the class-based one (756).  Tables: the reference's processed lab data
(CLBR_IRSL50_0.25KperGy, CLBR_IR50_ISO experiments 0-6) as held in
reference_outputs.json.  For each row / observation: mean and standard deviation
of the trapped fraction over `sims` simulations.  Writes
tests/golden/legacy_ensembles.json.

Run: python tests/golden/make_legacy_fixtures.py   (about a minute on 8 cores)
"""
import json
import sys
from concurrent.futures import ProcessPoolExecutor
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import legacy_mc  # noqa: E402

OUT = Path(__file__).resolve().parent / "legacy_ensembles.json"
LAB = json.loads((Path(__file__).parent / "reference_outputs.json").read_text())["lab_data"]
SEED = 9173
EXP = dict(N_e=100.0, holes=437.6, rho_prime=1e-4, boundary_factor=1.2)
PHYS = dict(alpha=9e9, b=1e12, k_b=8.617343e-5, s=1e13, D=0.092, D0=400.35, E_cb=1.8,
            E_loc_1=1.4, E_loc_2=0.9, Retrap=0.5)


def tl_rows():
    t = LAB["CLBR_IRSL50_0.25KperGy"]
    return [dict(Duration=t["Duration"][k], T_start=t["T_start"][k], T_end=t["T_end"][k],
                 e_ratio=t["e_ratio"][k]) for k in range(len(t["Duration"]))]


def iso_experiments():
    t = LAB["CLBR_IR50_ISO"]
    last = t["exp_no"][-1]
    exps = []
    for n in sorted(set(t["exp_no"])):
        if n >= last:                       # functions.py:370 runs 0 .. last - 1
            continue
        idx = [i for i, e in enumerate(t["exp_no"]) if e == n]
        exps.append(dict(e_ratio=t["e_ratio"][idx[0]], temp=t["temp"][idx[0]],
                         dose=t["dose"][idx[0]], times=[t["time"][i] for i in idx]))
    return exps


def _tl(j):
    rng = np.random.default_rng([SEED, j])
    return np.array([legacy_mc.tl_fraction(EXP, PHYS, row, rng) for row in tl_rows()])


def _iso(j):
    rng = np.random.default_rng([SEED + 1, j])
    return np.concatenate([legacy_mc.iso_fractions(EXP, PHYS, ex, rng) for ex in iso_experiments()])


def main(sims_tl: int = 1024, sims_iso: int = 1024):
    with ProcessPoolExecutor(max_workers=8) as pool:
        tl = np.array(list(pool.map(_tl, range(sims_tl), chunksize=8)))
        iso = np.array(list(pool.map(_iso, range(sims_iso), chunksize=8)))
    fx = {
        "exp": EXP, "phys": PHYS,
        "tl": {"table": "CLBR_IRSL50_0.25KperGy", "sims": sims_tl,
               "mean": tl.mean(axis=0).tolist(), "std": tl.std(axis=0, ddof=1).tolist()},
        "iso": {"table": "CLBR_IR50_ISO", "experiments": len(iso_experiments()), "sims": sims_iso,
                "mean": iso.mean(axis=0).tolist(), "std": iso.std(axis=0, ddof=1).tolist()},
        "_about": ("numpy restatement of the legacy lab loops (oracle/legacy_mc.py) on the "
                   f"lab_TL + TLlab configs with E_loc_2 = 0.9, holes = 437.6, seed {SEED}; "
                   "made by tests/golden/make_legacy_fixtures.py"),
    }
    OUT.write_text(json.dumps(fx, indent=1) + "\n")
    print("wrote", OUT, "TL means", np.round(tl.mean(axis=0), 3))


if __name__ == "__main__":
    main()
