"""Build libmultiscat_amd.so in-tree with hipcc for gfx950.

    python -m multiscat_amd.build [-v]

HIP sources compile in parallel to objects under multiscat_amd/_build/, then link into
multiscat_amd/libmultiscat_amd.so. The packed-FP32 target feature is switched off for all
device code (the compiler forms packed FP32 instructions from ordinary float code by itself).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libmultiscat_amd.so"
ARCH = os.environ.get("MSX_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["msx_kernels.hip", "msx_eig.hip", "msx_eig_dc.hip", "msx_krylov.hip", "msx_dft.hip", "msx_diag_h3.hip", "msx_dft_h3.hip",
               "msx_diag.hip", "msx_profile.hip", "msx_solver.hip", "msx_noalloc.hip"]
CXX_SOURCES = ["msx_host.cpp"]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required)")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    if verbose and res.stderr.strip():
        print(res.stderr, file=sys.stderr)


def build(verbose: bool = False) -> Path:
    hipcc = _hipcc()
    BUILD.mkdir(exist_ok=True)
    common = ["-O3", "-fPIC", "-std=c++17", f"-I{ROOT / 'include'}", f"-I{CSRC}"]
    jobs = []
    objs = []
    for src in HIP_SOURCES:
        obj = BUILD / (Path(src).stem + ".o")
        objs.append(obj)
        jobs.append([hipcc, *common, f"--offload-arch={ARCH}", "-Xclang", "-target-feature",
                     "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument",
                     "-c", str(CSRC / src), "-o", str(obj)])
    for src in CXX_SOURCES:
        obj = BUILD / (Path(src).stem + ".o")
        objs.append(obj)
        jobs.append(["g++", "-O3", "-fPIC", "-std=c++17", f"-I{ROOT / 'include'}", "-c",
                     str(CSRC / src), "-o", str(obj)])
    with ThreadPoolExecutor(max_workers=min(4, len(jobs))) as ex:
        for fut in [ex.submit(_run, j, verbose) for j in jobs]:
            fut.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs)],
         verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
