"""Every struct of the C ABI has the same size and field offsets in the headers
(compiled by gcc, as a caller in C would see them) as in the ctypes binding, so
no argument is read at a shifted offset.  CPU only."""
import ctypes
import shutil
import subprocess
from pathlib import Path

import pytest

from mcluminescence_amd import _abi

INCLUDE = Path(__file__).resolve().parent.parent / "include"
STRUCTS = {
    "mcl_run_params": _abi.RunParams,
    "mcl_geometry": _abi.Geometry,
    "mcl_stats": _abi.Stats,
    "mcl_plan_info": _abi.PlanInfo,
    "mcl_work": _abi.Work,
    "mcl_mlp_desc": _abi.MlpDesc,
    "mcl_adamw": _abi.AdamW,
    "mcl_de_params": _abi.DeParams,
}


def _c_layout(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mcl.h"',
             '#include "mcl_surrogate.h"', "int main(void) {"]
    for name, cls in STRUCTS.items():
        lines.append(f'  printf("{name} size %zu\\n", sizeof({name}));')
        for field, _ in cls._fields_:
            lines.append(f'  printf("{name} {field} %zu\\n", offsetof({name}, {field}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", str(INCLUDE), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    table = {}
    for line in out.splitlines():
        name, field, value = line.split()
        table[(name, field)] = int(value)
    return table


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc is needed to read the C layout")
def test_ctypes_structs_match_the_headers(tmp_path):
    c = _c_layout(tmp_path)
    for name, cls in STRUCTS.items():
        assert c[(name, "size")] == ctypes.sizeof(cls), name
        for field, _ in cls._fields_:
            assert c[(name, field)] == getattr(cls, field).offset, (name, field)
