"""Build recipe for libmcl.so (HIP kernels + C ABI) for gfx950.

Run as ``python -m mcluminescence_amd.csrc.build`` or through
``__graft_entry__.build()``.  The library is built in-tree next to the package
so that it travels with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
PKG = HERE.parent
ROOT = PKG.parent
OUT = PKG / "libmcl.so"

SOURCES = ["mc_fixed_dt.hip", "mc_fixed_dt_window.hip", "mc_fixed_dt_global.hip",
           "mc_fixed_dt_global_window.hip", "mc_event.hip", "mc_lab.hip", "glow.hip",
           "curves.hip", "mcl_abi.hip", "mlp.hip", "mlp_fused.hip", "mlp_abi.hip", "de.hip"]

# -ffp-contract=off: every fused multiply-add in a decision path is an explicit
# __builtin_fma, so the CPU replay harness can reproduce each rounding.
# -packed-fp32-ops: no v_pk_*_f32 generated from scalar float code.
HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    "-Wno-unused-value",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(path).exists():
        raise RuntimeError("hipcc not found: the HIP toolchain is required to build libmcl.so")
    return path


def build(verbose: bool = False, jobs: int = 0) -> Path:
    """Compiles every source (jobs at a time; 0: one per CPU, at most 8) and links."""
    tmp = PKG / "_objs"
    tmp.mkdir(exist_ok=True)
    objs = [str(tmp / (Path(src).stem + ".o")) for src in SOURCES]

    def compile_one(k: int) -> None:
        cmd = [hipcc(), *HIPCC_FLAGS, "-c", str(HERE / SOURCES[k]), "-o", objs[k]]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    jobs = jobs or min(8, os.cpu_count() or 1, len(SOURCES))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        for f in [pool.submit(compile_one, k) for k in range(len(SOURCES))]:
            f.result()
    tmp_out = OUT.with_suffix(".so.tmp")
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", str(tmp_out)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp_out, OUT)
    shutil.rmtree(tmp, ignore_errors=True)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
