"""The reference's own throughput benchmark on this backend (scripts/bench.sh): benchmarks/cube.gltf
and the generated 5M-triangle scale scene rendered at 1024x768 with 1,024 spp, GPU seconds the best
of `runs`, the camera framed as the CLI frames a file without one (main.rs:294-296). Both
integrators are timed; the reference's CLI default is the megakernel (main.rs:150). Published GPU
seconds of the same runs (benchmarks/baseline/*.json) are listed beside ours, with the speedup of
our megakernel (the CLI's default integrator) and of our faster integrator over each.

    python -m verion_amd.reference_bench --cube tests/golden/cube.gltf
"""
from __future__ import annotations

import argparse
import json
import sys
import time

from . import gltf, scenes
from .backend import MEGAKERNEL, WAVEFRONT, Context, Renderer

# benchmarks/baseline/*.json: gpu_s of cube and scale, 1024x768, 1024 spp, best of 5
PUBLISHED = {
    "nvidia-geforce-rtx-3080-ac (Vulkan)": {"cube": 0.050603648000000036, "scale": 0.09229087999999999},
    "apple-m4-pro-ac": {"cube": 0.3201, "scale": 0.4198},
    "apple-m5-ac": {"cube": 0.3355, "scale": 0.4017},
}


def load_scenes(cube_path: str):
    cube, _ = gltf.load_gltf(cube_path)
    scale, _ = gltf.load_gltf_bytes(scenes.reference_scale_glb())
    out = []
    for name, s in (("cube", cube), ("scale", scale)):
        if s.camera == type(s.camera)():
            s.frame_camera()
        out.append((name, s))
    return out


def evaluate(cube_path: str, width: int = 1024, height: int = 768, spp: int = 1024, runs: int = 3, log=lambda *a: None) -> dict:
    ctx = Context()
    res = {"settings": f"{width}x{height}, {spp} spp, GPU seconds best of {runs} (scripts/bench.sh uses 5)"}
    for name, s in load_scenes(cube_path):
        t0 = time.perf_counter()
        r = Renderer(ctx, s)
        entry = {"triangles": r.bvh_stats()["triangles"], "scene_and_lbvh_s": round(time.perf_counter() - t0, 2)}
        for label, integ in (("megakernel", MEGAKERNEL), ("wavefront", WAVEFRONT)):
            r.set_integrator(integ)
            warm = r.new_accumulation(width, height)
            r.accumulate(warm, 4)
            warm.close()
            best = None
            for _ in range(runs):
                acc = r.new_accumulation(width, height)
                r.accumulate(acc, spp)
                g = acc.gpu_seconds()
                acc.close()
                best = g if best is None else min(best, g)
            entry[label] = {"gpu_s": round(best, 5), "msamples_per_s": round(width * height * spp / best / 1e6, 1)}
        mk = entry["megakernel"]["gpu_s"]
        best = min(mk, entry["wavefront"]["gpu_s"])
        entry["published_gpu_s"] = {k: v[name] for k, v in PUBLISHED.items()}
        entry["speedup_vs_published_megakernel_default"] = {k: round(v[name] / mk, 3) for k, v in PUBLISHED.items()}
        entry["speedup_vs_published_faster_integrator"] = {k: round(v[name] / best, 3) for k, v in PUBLISHED.items()}
        log(name, entry)
        res[name] = entry
        r.close()
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--cube", default="tests/golden/cube.gltf")
    ap.add_argument("--runs", type=int, default=3)
    a = ap.parse_args()
    print(json.dumps(evaluate(a.cube, runs=a.runs, log=lambda *x: print(*x, file=sys.stderr, flush=True))), flush=True)


if __name__ == "__main__":
    main()
