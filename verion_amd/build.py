"""Builds libverion_hip.so in-tree with hipcc for gfx950 (no JIT, no torch extension).

    python -m verion_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

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


def build(verbose: bool = False, jobs: int = 0) -> str:
    """Compiles every source to an object in parallel (each is its own device code object, as in
    one hipcc call without -fgpu-rdc), then links the shared library."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cflags = [f for f in FLAGS if f != "-shared"]
    cmds = [["hipcc", *cflags, "-c", "-o", os.path.join(objdir, s[:-4] + ".o"), os.path.join(CSRC, s)] for s in SOURCES]
    jobs = jobs or min(len(cmds), max(1, (os.cpu_count() or 1)))
    with ThreadPoolExecutor(jobs) as ex:
        for cmd in cmds:
            if verbose:
                print(" ".join(cmd), flush=True)
        results = list(ex.map(lambda c: subprocess.run(c, capture_output=True, text=True), cmds))
    for cmd, r in zip(cmds, results):
        err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
        if err.strip():
            print(err, file=sys.stderr, flush=True)
        if r.returncode != 0:
            raise subprocess.CalledProcessError(r.returncode, cmd)
    link = ["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB + ".tmp", *[c[c.index("-o") + 1] for c in cmds]]
    if verbose:
        print(" ".join(link), flush=True)
    subprocess.run(link, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    sys.exit(0)
