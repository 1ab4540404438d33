"""ReSTIR on the many-lights workload: per-frame cost and error against the plain wavefront tracer.

Scene: scenes.many_lights() (ground + 41,000 boxes, 502,252 triangles, 10,000 emissive triangles of
lognormal(0, 1) emission). Every method renders 1 spp per frame into an FP64 film at 1920x1080;
the reference is a plain render with `ref_spp` samples and an independent seed. Methods: plain
(deferred shadow rays), ReSTIR DI (`candidates` initial candidates, the reservoirs of the pixel
and `neighbours` random neighbours within 6 px of the previous frame, M capped at `history`:
pt_core.slang:702-900; the workload's 5 neighbours, and the reference's fixed 2 beside it), and
DI + GI (pt_core.slang:901-980, 1566-1617). Each advances along a
doubling frame ladder to `frames` and is scored by relMSE after every rung; times are the GPU
seconds the library reports. Reported per method: ms per frame, relMSE at `frames` frames, and
the plain tracer's time to reach each method's error at `frames` frames (log-log interpolation),
as a ratio to the method's own time (above 1: the method gets there sooner).

    python -m verion_amd.restir_eval [--frames 64] [--ref-spp 1024] [--candidates 32] [--neighbours 5]
    python -m verion_amd.restir_eval --gltf tests/golden/many_lights.gltf   # the reference's own protocol
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import quality, scenes
from .backend import WAVEFRONT, Context, Renderer, Restir


def evaluate(width: int = 1920, height: int = 1080, frames: int = 64, ref_spp: int = 1024, candidates: int = 32,
             history: int = 4, neighbours: int = 5, seed: int = 1, plain_frames: int = 0, methods=None,
             log=lambda *a: None) -> dict:
    """`plain_frames` (default 16 x frames) bounds the plain ladder the methods' errors are timed
    on; `methods` picks among restir_di_2_neighbours, restir_di and restir_di_gi (default all)."""
    ctx = Context()
    r = Renderer(ctx, scenes.many_lights())
    r.set_integrator(WAVEFRONT)
    t0 = time.perf_counter()
    r.set_seed(0x5EED0000 + seed)
    ref_acc = r.new_accumulation(width, height)
    done = 0
    while done < ref_spp:
        k = min(128, ref_spp - done)
        r.accumulate(ref_acc, k)
        done += k
    ref = ref_acc.mean().astype(np.float64)
    ref_s = ref_acc.gpu_seconds()
    ref_acc.close()
    log(f"reference {ref_spp} spp: {ref_s:.1f} s GPU ({time.perf_counter() - t0:.1f} s wall)")
    r.set_seed(seed)

    def run(restir, max_frames):
        r.set_restir(restir)
        acc = r.new_accumulation(width, height)

        def step(k):
            before = acc.gpu_seconds()
            r.accumulate(acc, k)
            return acc.gpu_seconds() - before

        pts = quality.ladder(step, acc.mean, ref, max_frames)
        acc.close()
        r.set_restir(None)
        return pts

    plain = run(None, plain_frames or frames * 16)
    log("plain", [(f, round(t, 4), e) for f, t, e in plain])
    at = {f: (t, e) for f, t, e in plain}
    out = {"plain": {"ms_per_frame": 1e3 * at[frames][0] / frames, "relmse": at[frames][1],
                     "ladder": [[f, round(t, 5), e] for f, t, e in plain]}}
    curve = [(t, e) for _, t, e in plain]
    all_methods = (("restir_di_2_neighbours", Restir(di=True, gi=False, candidates=candidates, history=history)),
               ("restir_di", Restir(di=True, gi=False, candidates=candidates, history=history, neighbours=neighbours)),
               ("restir_di_gi", Restir(di=True, gi=True, candidates=candidates, history=history, neighbours=neighbours)))
    for name, restir in all_methods:
        if methods is not None and name not in methods:
            continue
        pts = run(restir, frames)
        log(name, [(f, round(t, 4), e) for f, t, e in pts])
        t, e = pts[-1][1], pts[-1][2]
        tp = quality.time_to(curve, e)
        out[name] = {"ms_per_frame": 1e3 * t / pts[-1][0], "relmse": e, "plain_seconds_to_this_error": tp,
                     "speedup_to_equal_error": (tp / t) if tp else None,
                     "ladder": [[f, round(tt, 5), ee] for f, tt, ee in pts]}
    return {"scene": "many_lights (502,252 triangles, 10,000 emitters)", "width": width, "height": height,
            "frames": frames, "reference_spp": ref_spp, "candidates": candidates, "history": history,
            "neighbours": neighbours,
            "relmse": "mean over pixels and channels of (x-r)^2/(r^2+0.01)", **out}


def mae_over_mean(x: np.ndarray, ref: np.ndarray) -> float:
    """exr_diff's relative mean absolute error (crates/verion-render-io/examples/exr_diff.rs:72-78):
    the mean over pixels and channels of |x - r| over the mean of r."""
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.mean(np.abs(x - ref)) / np.mean(ref))


def reference_protocol(scene_path: str, width: int = 512, height: int = 384, ref_spp: int = 4096, max_spp: int = 256,
                       candidates: int = 4, history: int = 4, seed: int = 1, log=lambda *a: None) -> dict:
    """The reference's own ReSTIR measurement (scripts/eval_restir.sh:1-59, docs/adr/0016:21-33) on
    one of its scenes: plain path tracing ("reference" mode), ReSTIR DI and DI + GI with the CLI's
    defaults (4 candidates, history 4, the 2 fixed neighbours) render 1, 2, 4, ... 256 spp at
    512x384; each image is scored by mae/mean against a plain render of `ref_spp` spp; every method
    is timed to the error plain tracing has at 128 spp (log-log interpolation), and its speedup is
    plain's 128-spp time over that. Each method runs on whichever integrator is faster for it."""
    from . import gltf

    scene, _ = gltf.load_gltf(scene_path)
    ctx = Context()
    r = Renderer(ctx, scene)
    r.set_integrator(WAVEFRONT)
    r.set_seed(0x5EED0000 + seed)
    ref_acc = r.new_accumulation(width, height)
    done = 0
    while done < ref_spp:
        k = min(256, ref_spp - done)
        r.accumulate(ref_acc, k)
        done += k
    ref = ref_acc.mean().astype(np.float64)
    ref_acc.close()
    r.set_seed(seed)

    def ladder(restir, integrator):
        r.set_integrator(integrator)
        r.set_restir(restir)
        acc = r.new_accumulation(width, height)
        pts, frames, secs, rung = [], 0, 0.0, 1
        while frames < max_spp:
            before = acc.gpu_seconds()
            r.accumulate(acc, rung - frames)
            secs += acc.gpu_seconds() - before
            frames = rung
            pts.append((frames, secs, mae_over_mean(acc.mean(), ref)))
            rung *= 2
        acc.close()
        r.set_restir(None)
        return pts

    from .backend import MEGAKERNEL
    variants = {"reference": None, "interactive": Restir(di=True, gi=False, candidates=candidates, history=history),
                "interactive+gi": Restir(di=True, gi=True, candidates=candidates, history=history)}
    curves, integ = {}, {}
    for name, restir in variants.items():
        best = None
        for integrator, label in ((MEGAKERNEL, "megakernel"), (WAVEFRONT, "wavefront")):
            pts = ladder(restir, integrator)
            if best is None or pts[-1][1] < best[-1][1]:
                best, integ[name] = pts, label
        curves[name] = best
        log(name, integ[name], [(f, round(t, 5), round(e, 5)) for f, t, e in best])
    base = next(p for p in curves["reference"] if p[0] == 128)
    target, base_s = base[2], base[1]
    out = {}
    for name, pts in curves.items():
        t = quality.time_to([(tt, e) for _, tt, e in pts], target)
        out[name] = {"integrator": integ[name], "seconds_to_target": t, "speedup": (base_s / t) if t else None,
                     "ladder": [[f, round(tt, 5), e] for f, tt, e in pts]}
    return {"scene": os.path.basename(scene_path), "width": width, "height": height, "reference_spp": ref_spp,
            "error": "mae/mean (exr_diff.rs)", "target": "plain path tracing's error at 128 spp", "target_error": target,
            "candidates": candidates, "history": history, **out}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--frames", type=int, default=64)
    ap.add_argument("--ref-spp", type=int, default=1024)
    ap.add_argument("--candidates", type=int, default=32)
    ap.add_argument("--history", type=int, default=4)
    ap.add_argument("--neighbours", type=int, default=5)
    ap.add_argument("--gltf", action="append", default=[],
                    help="a scene of the reference (e.g. tests/golden/many_lights.gltf): run its own ReSTIR protocol "
                         "instead (512x384, 4 candidates, 2 neighbours, time to plain's 128-spp error); repeatable")
    a = ap.parse_args()
    if a.gltf:
        res = [reference_protocol(g, log=lambda *x: print(*x, file=sys.stderr, flush=True)) for g in a.gltf]
        print(json.dumps(res), flush=True)
        return
    res = evaluate(a.width, a.height, a.frames, a.ref_spp, a.candidates, a.history, a.neighbours,
                   log=lambda *x: print(*x, file=sys.stderr, flush=True))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
