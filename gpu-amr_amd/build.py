"""Build recipe of libamr_hip.so (HIP, gfx950) -- used by __graft_entry__.build().

The library is compiled in-tree so that the shared object travels with the
source tree.  -ffp-contract=off keeps every multiply and add separately
rounded (the step is bit-reproducible against the CPU oracle); packed FP32
instructions are disabled for all device code.
"""
from __future__ import annotations

import os
import pathlib
import shutil
import subprocess

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parent
LIB = HERE / "libamr_hip.so"
SOURCES = [HERE / "csrc" / "amr_hip.hip"]
HEADERS = sorted((HERE / "csrc").glob("*.hpp")) + [ROOT / "include" / "amr_hip.h"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and pathlib.Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def flags():
    return [
        "--offload-arch=gfx950",
        "-O3",
        "-std=c++20",
        "-fPIC",
        "-shared",
        "-ffp-contract=off",
        "-Xclang",
        "-target-feature",
        "-Xclang",
        "-packed-fp32-ops",
        "-Wall",
        "-Wno-unused-function",
        f"-I{ROOT / 'include'}",
        f"-I{HERE / 'csrc'}",
    ]


CALIB_LIB = HERE / "libamr_calib.so"
CALIB_SOURCES = [HERE / "csrc" / "calibration.hip"]


def _compile(sources, out, verbose):
    tmp = out.with_suffix(".so.tmp")
    cmd = [hipcc(), *flags(), "-o", str(tmp), *map(str, sources)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if verbose or r.returncode != 0:
        print(" ".join(cmd))
        print(r.stdout[-4000:], r.stderr[-8000:])
    if r.returncode != 0:
        raise RuntimeError(f"{out.name} build failed")
    os.replace(tmp, out)


def build(verbose: bool = False) -> str:
    """libamr_hip.so (the solver) and libamr_calib.so (the copy kernel that
    calibrates the HBM byte counters)."""
    _compile(SOURCES, LIB, verbose)
    _compile(CALIB_SOURCES, CALIB_LIB, verbose)
    return str(LIB)


if __name__ == "__main__":
    print(build(verbose=True))
