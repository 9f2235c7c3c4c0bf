"""Builds the native libraries of newton_amd in-tree.

* ``newton_amd/lib/libvbd_hip.so``  -- gfx950 HIP kernels + C ABI (``include/vbd_hip.h``)
* ``newton_amd/lib/libvbd_ws.so``   -- gfx950 learned warm start (features, BF16 MFMA MLP, Adam;
  ``include/vbd_warmstart.h``)
* ``newton_amd/lib/libvbd_host.so`` -- host-side model construction: coloring, adjacency (C++)

Run ``python -m newton_amd.build`` (no GPU needed: hipcc cross-compiles).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
LIB_DIR = PKG / "lib"
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"

# FP32 arithmetic must follow the reference's operation order exactly: no FMA
# contraction, IEEE division/sqrt, denormals kept.  Packed-FP32 instructions
# are disabled for all device code.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build newton_amd")


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        sys.stderr.write(proc.stderr)
        raise RuntimeError(f"build command failed ({proc.returncode}): {' '.join(cmd)}")


def build_host() -> Path:
    LIB_DIR.mkdir(exist_ok=True)
    out = LIB_DIR / "libvbd_host.so"
    tmp = out.with_suffix(".so.tmp")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", str(tmp), str(CSRC / "vbd_host.cpp")])
    os.replace(tmp, out)
    return out


def build_hip() -> Path:
    LIB_DIR.mkdir(exist_ok=True)
    out = LIB_DIR / "libvbd_hip.so"
    tmp = out.with_suffix(".so.tmp")
    _run([_hipcc(), *HIP_FLAGS, "-I", str(INCLUDE), "-o", str(tmp), str(CSRC / "vbd_hip.hip")])
    os.replace(tmp, out)
    return out


def build_ws() -> Path:
    LIB_DIR.mkdir(exist_ok=True)
    out = LIB_DIR / "libvbd_ws.so"
    tmp = out.with_suffix(".so.tmp")
    _run([_hipcc(), *HIP_FLAGS, "-I", str(INCLUDE), "-o", str(tmp), str(CSRC / "vbd_warmstart.hip")])
    os.replace(tmp, out)
    return out


def build_all() -> list[Path]:
    return [build_host(), build_hip(), build_ws()]


if __name__ == "__main__":
    for p in build_all():
        print(p)
