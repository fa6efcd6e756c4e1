"""Project paths (the reference's src/class/paths.py:1-7).

PROJECT_ROOT is $MCL_PROJECT_ROOT when set, else the current directory; results
go to PROJECT_ROOT/results/simulations and lab tables are read from
PROJECT_ROOT/data/processed, as in the reference.
"""
from __future__ import annotations

import os
from pathlib import Path


def project_root() -> Path:
    return Path(os.environ.get("MCL_PROJECT_ROOT", os.getcwd())).resolve()


def simulations_dir() -> Path:
    return project_root() / "results" / "simulations"


def data_dir() -> Path:
    return project_root() / "data"
