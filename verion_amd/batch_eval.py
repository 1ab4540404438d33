"""The batched-jobs workload: 32 seeded render jobs of uneven size in one shared wavefront, against
the reference's way of serving them, one job at a time through the session loop in 8-spp chunks
(crates/verion-render/src/session.rs:17, 323-336).

Jobs: resolution with height log-uniform in [256, 1080] and aspect uniform in [1, 16/9] (256x256
to 1920x1080), spp log-uniform in [4, 256], adaptive sampling at relative error 0.01 with the
reference's default min_spp 16 (lib.rs:115-122); scenes drawn from a pool made by the procedural
generator (8 to 1,024 icospheres: 10k to 1.3M triangles) and the many-lights generator (800 to
41,000 boxes with 200 to 10,000 emitters: 10k to 500k triangles). All jobs arrive at time 0; a
job's latency is the GPU time until its film is final. Both ways render the same images (checked).

    python -m verion_amd.batch_eval [--jobs 32] [--seed 1]
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import scenes
from .backend import WAVEFRONT, Adaptive, Batch, Context, Renderer

POOL = [("procedural", 8), ("procedural", 64), ("procedural", 256), ("procedural", 1024),
        ("many_lights", 800), ("many_lights", 8000), ("many_lights", 41000)]


def make_scene(kind: str, size: int):
    if kind == "procedural":
        return scenes.procedural(spheres=size)
    return scenes.many_lights(emitters=max(200, size // 4), clutter_boxes=size)


def make_jobs(count: int, seed: int):
    rng = np.random.default_rng(seed)
    jobs = []
    for _ in range(count):
        h = int(round(math.exp(rng.uniform(math.log(256), math.log(1080)))))
        w = min(1920, int(round(h * rng.uniform(1.0, 16.0 / 9.0))))
        spp = int(round(math.exp(rng.uniform(math.log(4), math.log(256)))))
        jobs.append({"scene": int(rng.integers(len(POOL))), "width": w, "height": h, "spp": spp,
                     "seed": int(rng.integers(1, 1 << 31))})
    return jobs


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs, np.float64), q))


def evaluate(count: int = 32, seed: int = 1, threshold: float = 0.01, min_spp: int = 16, chunk: int = 8,
             log=lambda *a: None) -> dict:
    ctx = Context()
    specs = make_jobs(count, seed)
    used = sorted({j["scene"] for j in specs})
    t0 = time.perf_counter()
    renderers = {}
    for k in used:
        r = Renderer(ctx, make_scene(*POOL[k]))
        r.set_integrator(WAVEFRONT)
        renderers[k] = r
    log(f"{len(used)} scenes ready in {time.perf_counter() - t0:.1f} s")
    ad = Adaptive(threshold, min_spp)
    jobs = [dict(renderer=renderers[j["scene"]], width=j["width"], height=j["height"], spp=j["spp"], adaptive=ad,
                 seed=j["seed"]) for j in specs]

    # batched: all jobs in one wavefront, autotuned, graph-replayed passes
    b = Batch(ctx, jobs)
    b.run(graph=True)
    bst = [b.job_stats(k) for k in range(count)]
    batch_s = bst[0]["batch_s"]
    paths = sum(s["paths"] for s in bst)
    tuning = b.tuning()
    report = b.tuning_report()
    log(f"batched: {batch_s:.3f} s GPU, {paths / batch_s / 1e6:.1f} Msamples/s, tuning {tuning}")

    # one job at a time, 8-spp chunks (the session loop)
    lat_seq, seq_s, seq_paths = [], 0.0, 0
    same = 0
    for k, (j, spec) in enumerate(zip(jobs, specs)):
        r = j["renderer"]
        r.set_seed(spec["seed"])
        r.set_adaptive(ad)
        acc = r.new_accumulation(j["width"], j["height"])
        done = 0
        while done < j["spp"]:
            step = min(chunk, j["spp"] - done)
            r.accumulate(acc, step)
            done += step
        seq_s += acc.gpu_seconds()
        lat_seq.append(seq_s)
        seq_paths += bst[k]["paths"]
        if k < 4:  # the batch rendered the same image
            same += int(np.array_equal(acc.film(), b.film(k)))
        acc.close()
        r.set_adaptive(None)
        r.set_seed(0)
    log(f"sequential: {seq_s:.3f} s GPU")
    lat_b = [s["latency_s"] for s in bst]
    return {
        "jobs": count, "seed": seed, "adaptive": {"threshold": threshold, "min_spp": min_spp},
        "scenes": [{"generator": POOL[k][0], "size": POOL[k][1], "triangles": int(renderers[k].bvh_stats()["triangles"])}
                   for k in used],
        "pixels_total": int(sum(j["width"] * j["height"] for j in jobs)),
        "spp_range": [min(j["spp"] for j in jobs), max(j["spp"] for j in jobs)],
        "batched": {"seconds": batch_s, "msamples_per_s": paths / batch_s / 1e6, "p50_latency_s": percentile(lat_b, 50),
                    "p99_latency_s": percentile(lat_b, 99), "tuning": list(tuning),
                    "tuning_tried": [[list(t), ms] for t, ms in report]},
        "one_at_a_time": {"seconds": seq_s, "msamples_per_s": seq_paths / seq_s / 1e6, "p50_latency_s": percentile(lat_seq, 50),
                          "p99_latency_s": percentile(lat_seq, 99), "chunk_spp": chunk},
        "paths": int(paths), "films_checked_equal": f"{same}/4",
        "per_job": [{"width": j["width"], "height": j["height"], "spp": j["spp"], "scene": POOL[s["scene"]][0] + str(POOL[s["scene"]][1]),
                     "passes": st["passes"], "batched_latency_s": round(st["latency_s"], 4), "alone_latency_s": round(ls, 4)}
                    for j, s, st, ls in zip(jobs, specs, bst, lat_seq)],
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--jobs", type=int, default=32)
    ap.add_argument("--seed", type=int, default=1)
    a = ap.parse_args()
    res = evaluate(a.jobs, a.seed, log=lambda *x: print(*x, file=sys.stderr, flush=True))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
