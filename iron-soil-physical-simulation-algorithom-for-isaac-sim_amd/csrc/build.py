"""Build libmv2.so (HIP, gfx950) in-tree.

    python iron-soil-physical-simulation-algorithom-for-isaac-sim_amd/csrc/build.py

Every translation unit is compiled with -ffp-contract=off (the FP64 path must
round like the NumPy reference, and faces shared by two cells must evaluate
bit-identically) and with packed FP32 disabled.  Objects go to csrc/_build/,
the shared library next to the Python package that loads it.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

CSRC = Path(__file__).resolve().parent
PKG_ROOT = CSRC.parent
REPO = PKG_ROOT.parent
OUT_DIR = PKG_ROOT / "isaac_bulk_hip"
LIB = OUT_DIR / "libmv2.so"
BUILD = CSRC / "_build"
SOURCES = ["mv2_context.hip", "mv2_substep.hip", "mv2_exchange.hip", "mv2_relax.hip",
           "mv2_label.hip", "mv2_batch.hip", "mv2_batch_ops.hip", "mv2_surrogate.hip", "mv2_deposition.hip",
           "mv2_frontier.hip", "mv2_surface.hip", "mv2_patch.hip"]
ARCH = os.environ.get("MV2_OFFLOAD_ARCH", "gfx950")

CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
    f"-I{REPO / 'include'}", f"-I{CSRC}",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libmv2")


def _compile(src: str) -> Path:
    obj = BUILD / (Path(src).stem + ".o")
    cmd = [hipcc(), *CFLAGS, "-c", str(CSRC / src), "-o", str(obj)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{proc.stdout}\n{proc.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    sources = [s for s in SOURCES if (CSRC / s).exists()]
    jobs = min(len(sources), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as pool:
        objs = list(pool.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{proc.stdout}\n{proc.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
