"""Build libeemd_mi355x.so (HIP C++ for gfx950) in-tree with hipcc.

    python -m eemd4cuda_amd.build

The library lands in ``eemd4cuda_amd/_build/libeemd_mi355x.so``; it is
git-ignored but travels with the source tree to wherever the tree is copied.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT_DIR = PKG / "_build"
LIB_NAME = "libeemd_mi355x.so"
SOURCES = ["eemd_sift.hip", "eemd_mlp.hip", "eemd_spectra.hip", "eemd_capi.cpp"]

# -ffp-contract=off: the FP64 path reproduces the oracle's rounding (explicit
# fma() only where intended).  Packed FP32 instructions are disabled.
HIPCC_FLAGS = [
    "-x", "hip", "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


# hipFFT for the analytic signal of the Hilbert spectrum
LIBS = ["-L/opt/rocm/lib", "-lhipfft"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libeemd_mi355x.so)")


def lib_path() -> Path:
    return OUT_DIR / LIB_NAME


def build(verbose: bool = False) -> Path:
    """Compile every HIP source for gfx950 into the shared library."""
    OUT_DIR.mkdir(parents=True, exist_ok=True)
    srcs = [str(CSRC / s) for s in SOURCES if (CSRC / s).exists()]
    tmp = OUT_DIR / (LIB_NAME + ".tmp")
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", str(CSRC), "-I", str(PKG.parent / "include"),
                                     "-o", str(tmp)] + srcs + LIBS
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stderr[-8000:])
    os.replace(tmp, lib_path())
    return lib_path()


if __name__ == "__main__":
    p = build(verbose="-v" in sys.argv)
    print(p)
