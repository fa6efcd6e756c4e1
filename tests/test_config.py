"""Config composition and sweep expansion (conf/ schema of the reference, Hydra grammar,
engine.initialize_runs zip expansion, engine.py:13-38)."""
import json
from pathlib import Path

import numpy as np
import pytest

from mcluminescence_amd import config, engine
from mcluminescence_amd.simulate import results_frame, write_results_csv

GOLDEN = json.loads((Path(__file__).parent / "golden" / "reference_outputs.json").read_text())


def test_yaml_scalars_typed_like_omegaconf():
    d = config.load_yaml("a: 3e-4\nb: 1e12\nc: 150e-9\nd: 1.5\ne: 7\nf: False\ng: [0.1,1,5]\nh: text\n")
    assert d == {"a": 3e-4, "b": 1e12, "c": 150e-9, "d": 1.5, "e": 7, "f": False,
                 "g": [0.1, 1, 5], "h": "text"}
    assert isinstance(d["a"], float) and isinstance(d["e"], int)


def test_default_composition_is_tl12_basicTL12():
    cfg = config.compose()
    assert cfg.exp_type_fp.name == "TL" and cfg.exp_type_fp.N_e == 2000
    assert cfg.physics_fp.E_cb == 1000 and cfg.physics_fp.alpha == 1.6e9
    assert cfg.exp_type_fp.rho_prime == 3e-4 and isinstance(cfg.exp_type_fp.rho_prime, float)
    assert "defaults" not in cfg


def test_overrides_grammar():
    cfg = config.compose(["+tag=TLwithDose", "exp_type_fp.e_ratio_start=0.05",
                          "physics_fp.D=0.092", "exp_type_fp.T_rate=[1,2]"])
    assert cfg.tag == "TLwithDose" and cfg.get("tag", "") == "TLwithDose"
    assert cfg.exp_type_fp.e_ratio_start == 0.05 and cfg.physics_fp.D == 0.092
    assert cfg.exp_type_fp.T_rate == [1, 2]
    with pytest.raises(KeyError):
        config.compose(["tag=x"])                      # new key needs '+'
    with pytest.raises(KeyError):
        config.compose(["+exp_type_fp.N_e=5"])         # existing key cannot be appended
    cfg = config.compose(["++exp_type_fp.N_e=5"])
    assert cfg.exp_type_fp.N_e == 5


def test_group_selection():
    cfg = config.compose(["exp_type_fp=iso_decay_4096x400", "physics_fp=trap_1p4eV"])
    assert cfg.exp_type_fp.sims == 4096 and cfg.physics_fp.E_cb == 1.4
    with pytest.raises(FileNotFoundError):
        config.compose(["exp_type_fp=nope"])


def test_initialize_runs_is_zip_with_repetition():
    runs = engine.initialize_runs(config.compose())
    assert len(runs) == 4
    got = [(r.exp_type_fp.T_rate, r.exp_type_fp.duration, r.exp_type_fp.boundary_factor,
            r.exp_type_fp.T_start) for r in runs.values()]
    assert got == [(0.1, 8000, 1.2, 0), (1, 800, 1.2, 0), (5, 160, 1.2, 0), (20, 40, 1.2, 0)]
    cfg = config.compose(["exp_type_fp.T_rate=[1,2,3]"])
    with pytest.raises(ValueError, match="not divisible"):
        engine.initialize_runs(cfg)
    cfg = config.compose(["exp_type_fp.T_rate=[1,2]", "exp_type_fp.duration=[10,20,30,40]"])
    runs = engine.initialize_runs(cfg)
    assert [r.exp_type_fp.T_rate for r in runs.values()] == [1, 2, 1, 2]


def test_physics_rejects_unknown_keys_like_the_reference():
    cfg = config.compose(["+physics_fp.E=1.2"])          # BG_basic-style key
    run = engine.initialize_runs(cfg)[0]
    with pytest.raises(TypeError):
        engine.physics_of(run)


def test_schedules():
    runs = engine.initialize_runs(config.compose())
    s = [engine.schedule(r) for r in runs.values()]
    assert [x["dt"] for x in s] == [10.0, 1.0, 0.2, 0.05]
    assert all(x["n_steps"] == 800 for x in s)
    iso = engine.initialize_runs(config.compose(["exp_type_fp=iso_decay_4096x400",
                                                 "physics_fp=trap_1p4eV"]))[0]
    assert engine.schedule(iso) == {"dt": 0.5, "n_steps": 4000, "duration": 2000.0,
                                    "T_rate": 0.0, "T_start": 226.85}
    p = engine.run_params(iso)
    assert (p.sims, p.n_steps, p.dt, p.E_cb, p.holes) == (4096, 4000, 0.5, 1.4, 400.0)


def test_csv_writer_matches_reference_file_format(tmp_path):
    # rebuild the reference file's values around its sampled rows: 4 runs x 2 sims
    # x 20000 steps, 100 electrons of N_e = 2000 recombining one at a time
    schema = GOLDEN["csv_schema"]
    steps, sims, runs = 20000, 2, 4
    lum = np.zeros((steps, sims, runs))
    er = np.zeros((steps, sims, runs))
    er[:845, 0, 0] = 0.05
    er[844, 0, 0] = 0.0055
    er[:3, 1, 0] = 0.05
    path = write_results_csv(tmp_path / "exp_x.csv", lum, er)
    lines = path.read_text().splitlines()
    assert lines[0] == schema["header"]
    assert len(lines) - 1 == schema["n_rows"] == steps * sims * runs
    assert lines[1:6] == schema["first_lines"]
    assert lines[-3:] == schema["last_lines"]
    for ln in schema["sample_lines"][:3]:
        run, sim, step = (int(v) for v in ln.split(",")[:3])
        assert lines[1 + (run * sims + sim) * steps + step] == ln
    df = results_frame(lum, er)
    assert {c: str(t) for c, t in df.dtypes.items()} == schema["dtypes"]
