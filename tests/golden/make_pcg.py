"""Writes tests/golden/pcg_vectors.json: golden values of the PCG hash and the per-path random
stream of shaders/include/pt_core.slang:69-84, 651, computed here with numpy uint32 arithmetic
(an implementation independent of the oracle's C code and the GPU's HIP code).

    python tests/golden/make_pcg.py
"""
import json
import os

import numpy as np


def pcg(v):
    v = np.asarray(v, dtype=np.uint32)
    with np.errstate(over="ignore"):
        s = v * np.uint32(747796405) + np.uint32(2891336453)
        w = ((s >> ((s >> np.uint32(28)) + np.uint32(4))) ^ s) * np.uint32(277803737)
    return (w >> np.uint32(22)) ^ w


def stream(pixel, sample, n):
    """The first n Rng::next() values of camera path (pixel, sample), as float32 bit patterns."""
    with np.errstate(over="ignore"):
        st = pcg(np.uint32(pixel) ^ pcg(np.uint32(sample) * np.uint32(0x9E3779B9) + np.uint32(0x7F4A7C15)))
    out = []
    for _ in range(n):
        st = pcg(st)
        f = np.float32(st >> np.uint32(8)) * np.float32(1.0 / 16777216.0)
        out.append(int(np.array(f, np.float32).view(np.uint32)))
    return int(st), out


def main():
    inputs = [0, 1, 2, 3, 42, 0x7F4A7C15, 0x9E3779B9, 0xFFFFFFFF, 123456789, 2 ** 31]
    data = {"pcg": {str(v): int(pcg(v)) for v in inputs}, "streams": []}
    for pixel, sample in [(0, 0), (1, 0), (0, 1), (65535, 15), (2073599, 63), (12345, 1000)]:
        final, floats = stream(pixel, sample, 8)
        data["streams"].append({"pixel": pixel, "sample": sample, "state_after_8": final, "next_f32_bits": floats})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pcg_vectors.json")
    with open(path, "w") as f:
        json.dump(data, f, indent=1)
    print(path)


if __name__ == "__main__":
    main()
