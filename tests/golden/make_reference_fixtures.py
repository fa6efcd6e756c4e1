"""Extract small fixtures from the reference's committed outputs.

Reads (in this container only) /root/reference/results/simulations/exp_*.csv and
/root/reference/data/processed/*.csv, and writes tests/golden/reference_outputs.json:
  - per (run, sim) recombination totals and event-step counts of the three
    simulate.py result files (TL12 + basicTL12, 4 heating rates x 2 sims),
  - the CSV schema of the current writer (header, dtypes, first and last rows,
    row count) from exp_TLwithDose_emptystart.csv,
  - the processed lab tables the objectives read,
  - the parameter rows and recorded MSEs of the reference's lab fits
    (results/lab_sims/result_*.csv), as written (raw field strings).
Run: python tests/golden/make_reference_fixtures.py
"""
import json
from pathlib import Path

import numpy as np
import pandas as pd

REF = Path("/root/reference")
OUT = Path(__file__).resolve().parent / "reference_outputs.json"


def totals(path: Path, positional_names: bool):
    df = pd.read_csv(path)
    if positional_names:
        # older writer: header says step,sim,run but columns are run,sim,step by position
        df.columns = ["run", "sim", "step"] + list(df.columns[3:])
    out = {}
    for (run, sim), g in df.groupby(["run", "sim"]):
        rec = {"recombinations": int(g["lum"].sum())}
        if "electron_ratio" in g:
            nz = np.nonzero(g["electron_ratio"].to_numpy() > 0)[0]
            rec["last_nonzero_step"] = int(nz[-1]) if nz.size else -1
            rec["e_ratio_first"] = float(g["electron_ratio"].iloc[0])
        out[f"{run},{sim}"] = rec
    return out


def main():
    sims = REF / "results" / "simulations"
    fx = {
        "tl12_heating_rates_degC_per_s": [0.1, 1, 5, 20],
        "exp_": totals(sims / "exp_.csv", True),
        "exp_TLwithDose": totals(sims / "exp_TLwithDose.csv", True),
        "exp_TLwithDose_emptystart": totals(sims / "exp_TLwithDose_emptystart.csv", False),
    }
    df = pd.read_csv(sims / "exp_TLwithDose_emptystart.csv")
    raw_lines = (sims / "exp_TLwithDose_emptystart.csv").read_text().splitlines()
    fx["csv_schema"] = {
        "header": raw_lines[0],
        "n_rows": len(df),
        "dtypes": {c: str(t) for c, t in df.dtypes.items()},
        "first_lines": raw_lines[1:6],
        "last_lines": raw_lines[-3:],
        "sample_lines": [raw_lines[i] for i in (101, 845, 846, 20001, 159999)],
    }
    lab = {}
    for f in sorted((REF / "data" / "processed").glob("*.csv")):
        lab[f.stem] = pd.read_csv(f).to_dict(orient="list")
    fx["lab_data"] = lab
    fits = {}
    for f in sorted((REF / "results" / "lab_sims").glob("result_*.csv")):
        lines = f.read_text().splitlines()
        fits[f.stem] = {"header": lines[0].split(","),
                        "rows": {str(i + 1): ln.split(",") for i, ln in enumerate(lines) if i > 0}}
    fx["lab_fit_results"] = fits
    OUT.write_text(json.dumps(fx, indent=1))
    print("wrote", OUT)


if __name__ == "__main__":
    main()
