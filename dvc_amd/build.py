"""Builds the gfx950 library (dvc_amd/libdvc_accel.so) and its cached data.

    python -m dvc_amd.build          # compile + precompute the ladder inverses

Compile flags: device code must keep the reference's single-precision
rounding, so floating-point contraction is off, float division and sqrt are
IEEE-correct, and packed FP32 instructions are disabled.
"""
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libdvc_accel.so"
ARCH = os.environ.get("DVC_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def sources():
    return sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (Path(cand).exists() or cand == "hipcc"):
            return cand
    raise RuntimeError("hipcc not found")


def compile_library(verbose=False):
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", str(tmp), *map(str, sources()), "-ldl"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def precompute_inverses():
    """Writes data/reg<N>.invh (the full-rate GF(2) inverse) with the host-only
    helper so a GPU-less checkout can ship it."""
    from . import _lib
    _lib.prepare_data()


def main(argv=None):
    compile_library(verbose=True)
    precompute_inverses()
    print(f"built {LIB}")


if __name__ == "__main__":
    main(sys.argv[1:])
