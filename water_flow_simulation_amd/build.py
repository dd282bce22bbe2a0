"""Build recipe for libwfs.so (hipcc, gfx950).

    python -m water_flow_simulation_amd.build

compiles csrc/*.hip into water_flow_simulation_amd/libwfs.so in place. The
library is plain HIP with a C ABI (include/wfs.h); it has no PyTorch
dependency and is loaded with ctypes.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libwfs.so")

SOURCES = ["wfs_api.hip"]
HEADERS = ["lattice.h", "closure.h", "step_kernels.h", "aux_kernels.h"]

# -ffp-contract=off: no FMA contraction, so the FP32 kernels round like the
# scalar IEEE formulas they implement. Packed FP32 instructions are disabled.
# MFMA accumulators in ordinary VGPRs (the learned closure's network reads
# every accumulator back for bias, ReLU and the output layer).
HIPFLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mllvm", "-amdgpu-mfma-vgpr-form=true",
    "-Wall", "-Wno-unused-result",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: install ROCm or set HIPCC")


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPFLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC, *srcs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
