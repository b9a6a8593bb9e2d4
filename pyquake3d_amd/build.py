"""Build the HIP C-ABI library ``libpq3d.so`` for gfx950, in-tree.

The library is compiled with hipcc directly (no JIT cache), so the built file
sits next to the sources and travels with the source tree.

``python -m pyquake3d_amd.build`` rebuilds it.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
LIB = LIBDIR / "libpq3d.so"
INCLUDE = PKG.parent / "include"

SOURCES = [
    "tde_kernels.hip",
    "hmat_kernels.hip",
    "rk_kernels.hip",
    "mlp_kernels.hip",
    "tree_kernels.hip",
]

ARCH = os.environ.get("PQ3D_OFFLOAD_ARCH", "gfx950")

# Packed-FP32 instructions are disabled for all device code; FP contraction is
# off so that FP64 results follow the same rounding sequence as the x86-64
# reference build (which has no FMA contraction).
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def build_library(verbose: bool = False, jobs: int = 4) -> Path:
    """Compile every HIP source to an object, then link libpq3d.so."""
    LIBDIR.mkdir(exist_ok=True)
    objdir = LIBDIR / "obj"
    objdir.mkdir(exist_ok=True)
    cc = hipcc()
    srcs = [CSRC / s for s in SOURCES if (CSRC / s).exists()]
    procs = []
    objs = []
    for src in srcs:
        obj = objdir / (src.stem + ".o")
        objs.append(obj)
        cmd = [cc, *HIPCC_FLAGS, f"-I{CSRC}", f"-I{INCLUDE}", "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _drain(procs)
    _drain(procs)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    if verbose:
        print(" ".join(cmd), flush=True)
    out = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if out.returncode != 0:
        raise RuntimeError(f"link failed:\n{out.stdout}")
    os.replace(tmp, LIB)
    return LIB


def _drain(procs):
    while procs:
        src, p = procs.pop(0)
        out, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src.name}:\n{out.decode(errors='replace')}")


if __name__ == "__main__":
    path = build_library(verbose="-v" in sys.argv)
    print(path)
