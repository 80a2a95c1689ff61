"""Build libseevr_hip.so (gfx950 HIP kernels + C ABI) in-tree.

Usage: python -m seevr_amd.build        (or seevr_amd.build.build_library())

Every HIP translation unit is compiled with:
  -ffp-contract=off                         no fused multiply-add: each a*b+c
                                            rounds twice, like the oracle
  -fhip-fp32-correctly-rounded-divide-sqrt  IEEE division and square root
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                            no packed fp32 instructions
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "libseevr_hip.so"
OBJ = PKG / "_build"

HIP_SOURCES = ["sv_prims.hip", "sv_physics.hip", "sv_pgs.hip", "sv_bvh.hip", "sv_batch.hip", "sv_diff.hip", "sv_color.hip", "sv_warm.hip", "sv_noalloc.hip", "sv_capi.cpp"]
HOST_SOURCES = ["sv_scene.cpp"]

ARCH = os.environ.get("SEEVR_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
GXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-Wall"]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain (ROCm) is required to build libseevr_hip.so")


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")


def build_library(verbose: bool = False) -> Path:
    """Compile every source and link libseevr_hip.so next to this file."""
    hipcc = _hipcc()
    OBJ.mkdir(exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        obj = OBJ / (src + ".o")
        jobs.append([hipcc, *HIPCC_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    for src in HOST_SOURCES:
        obj = OBJ / (src + ".o")
        jobs.append(["g++", *GXX_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    with ThreadPoolExecutor(max_workers=min(len(jobs), 8)) as ex:
        for cmd, fut in zip(jobs, [ex.submit(_run, j) for j in jobs]):
            if verbose:
                print(" ".join(cmd), file=sys.stderr)
            fut.result()
    objs = [str(OBJ / (s + ".o")) for s in HIP_SOURCES + HOST_SOURCES]
    tmp = OUT.with_suffix(".so.tmp")
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", str(tmp)])
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build_library(verbose=True))
