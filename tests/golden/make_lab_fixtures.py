"""Ensemble fixtures of the reference's lab objectives restated in numpy
(oracle/lab_mc.py: TL_lab and ISO_lab loops of tl_trap_lab.py:65-179) for the GPU
lab-kernel test.

Physics and geometry are the shipped lab configs (conf/physics_fp/lab_TL.yaml,
conf/exp_type_fp/TLlab.yaml); the tables are the reference's processed lab data
(CLBR_IRSL50_0.25KperGy, CLBR_IR50_ISO) as held in reference_outputs.json.  For
each row / observation: mean and standard deviation of the simulated trapped
fraction over `sims` independent simulations.  Writes tests/golden/lab_ensembles.json.

Run: python tests/golden/make_lab_fixtures.py   (about a minute on 8 cores)
"""
import json
import sys
from concurrent.futures import ProcessPoolExecutor
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from mcluminescence_amd import config, engine  # noqa: E402
from oracle import event_mc as E  # noqa: E402
from oracle import lab_mc  # noqa: E402

OUT = Path(__file__).resolve().parent / "lab_ensembles.json"
LAB = json.loads((Path(__file__).parent / "reference_outputs.json").read_text())["lab_data"]
SEED = 7051


def setup():
    cfg = config.compose(["exp_type_fp=TLlab", "physics_fp=lab_TL"])
    run = engine.initialize_runs(cfg)[0]
    p = engine.run_params(run, sims=1, n_steps=1)
    exp, phys = E.run_params_to_cfg(p)
    phys["D"] = float(run.physics_fp.D)
    return exp, phys


def tl_rows():
    t = LAB["CLBR_IRSL50_0.25KperGy"]
    return [dict(Duration=t["Duration"][k], T_start=t["T_start"][k], T_end=t["T_end"][k],
                 e_ratio=t["e_ratio"][k]) for k in range(len(t["Duration"]))]


def iso_experiments():
    t = LAB["CLBR_IR50_ISO"]
    exps = []
    for n in sorted(set(t["exp_no"])):
        idx = [i for i, e in enumerate(t["exp_no"]) if e == n]
        exps.append(dict(e_ratio=t["e_ratio"][idx[0]], temp=t["temp"][idx[0]],
                         dose=t["dose"][idx[0]], times=[t["time"][i] for i in idx]))
    return exps


def _tl(j):
    exp, phys = setup()
    return lab_mc.tl_lab_fractions(exp, phys, tl_rows(), np.random.default_rng([SEED, j]))


def _iso(j):
    exp, phys = setup()
    return np.concatenate(lab_mc.iso_lab_fractions(exp, phys, iso_experiments(),
                                                   np.random.default_rng([SEED + 1, j])))


def main(sims_tl: int = 512, sims_iso: int = 256):
    with ProcessPoolExecutor(max_workers=8) as pool:
        tl = np.array(list(pool.map(_tl, range(sims_tl), chunksize=8)))
        iso = np.array(list(pool.map(_iso, range(sims_iso), chunksize=8)))
    fx = {
        "tl": {"table": "CLBR_IRSL50_0.25KperGy", "sims": sims_tl,
               "mean": tl.mean(axis=0).tolist(), "std": tl.std(axis=0, ddof=1).tolist()},
        "iso": {"table": "CLBR_IR50_ISO", "sims": sims_iso,
                "mean": iso.mean(axis=0).tolist(), "std": iso.std(axis=0, ddof=1).tolist()},
        "_about": ("numpy restatement of TL_lab / ISO_lab (oracle/lab_mc.py) on the lab_TL + "
                   f"TLlab configs, seed {SEED}; made by tests/golden/make_lab_fixtures.py"),
    }
    OUT.write_text(json.dumps(fx, indent=1) + "\n")
    print("wrote", OUT, "TL means", np.round(tl.mean(axis=0), 3))


if __name__ == "__main__":
    main()
