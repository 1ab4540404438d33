"""Builds libverion_hip.so in-tree with hipcc for gfx950 (no JIT, no torch extension).

    python -m verion_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libverion_hip.so")
SOURCES = ["vh_api.hip", "vh_bvh.hip", "vh_render.hip", "vh_nrc.hip", "vh_aov.hip", "vh_batch.hip"]
HEADERS = ["vh_internal.h", "vh_math.h", "vh_shade.h", "vh_trace.h", "vh_wavefront.h"]

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # bit-exact float: no contraction, IEEE division and sqrt, denormals kept
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wdouble-promotion",
    "-Werror=double-promotion",
]


def sources():
    return [os.path.join(CSRC, s) for s in SOURCES]


def build(verbose: bool = False) -> str:
    cmd = ["hipcc", *FLAGS, "-o", LIB + ".tmp", *sources()]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
