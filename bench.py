#!/usr/bin/env python3
"""Msamples/s of the plain wavefront path tracer at 1080p on one MI355X.

Workload (the procedural scale scene): 1,024 seeded level-3 icospheres (1,310,720 triangles) in a
20 m cube, GGX roughness U[0.05, 1], metallic Bernoulli(0.3), a sun and a constant environment map;
1920x1080, depth 8, wavefront integrator with compacted queues, FP32 paths and FP64 film.
A step is one accumulate call of 8 samples per pixel (8 x 2,073,600 camera paths), the chunk the
reference's render loop hands its backend (crates/verion-render-cli/src/main.rs:337-341); the
default 8 timed steps are the workload's 64 spp. Inputs (scene tables, LBVH) are resident in HBM
before timing starts.

Prints one JSON line with the throughput (`value`, plain wavefront), a roofline entry for the
dominant kernel (achieved algorithmic bytes per launch over its HIP-event launch time, against
8 TB/s), the CPU oracle's throughput on a bounded sample of the same workload, and `time_to_rmse`:
GPU seconds to relMSE 1e-3 of the neural radiance cache and of the plain tracer on the same scene
at 1080p, 1 spp per frame, against a plain reference render (16,384 spp, the workload's, by default); `batch`:
p50/p99 job latency and aggregate Msamples/s of 32 uneven jobs in one shared wavefront against
one job at a time; and `restir_many_lights`: ms per frame and relMSE of plain, ReSTIR DI and DI + GI
on the many-lights workload at 1080p; and `reference_benchmark`: the reference's own throughput
benchmark (its cube scene and 5M-triangle scale scene, 1024x768, 1,024 spp) beside the GPU seconds
it publishes for RTX 3080, M4 Pro and M5.

    python bench.py [--gpus 1] [--steps 8] [--warmup 2] [--spp-per-step 8]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
NODE_BYTES = 128       # one 4-wide LBVH node
TRI_BYTES = 48         # one leaf triangle record
STATE_WORDS = 17       # slot-indexed path state without the lighting split


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_from_profiles(kernel: str):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, copy-calibrated),
    when a summary for this kernel and workload is committed; None otherwise."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic_procedural_1080p.json")
    try:
        with open(path) as f:
            data = json.load(f)
        return data.get(kernel)
    except (OSError, ValueError):
        return None


def occupancy_from_profiles(kernel: str):
    """Resident waves per CU measured by rocprofv3 --pmc OccupancyPercent (per depth and weighted by
    launch time) and the resource that caps them, when a summary for this kernel is committed."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "occupancy_procedural_1080p.json")
    try:
        with open(path) as f:
            data = json.load(f)
        k = data.get(kernel)
        return dict(k, unit=data["unit"], source=data["source"]) if k else None
    except (OSError, ValueError, KeyError):
        return None


def algorithmic_bytes(c: dict, pixels: int, delta_lights: int, env: bool):
    """Algorithmic bytes of one sample pass, split by kernel, from the traversal counters.

    wf_extend: per live path the ray (24 B read, none at depth 0 where the ray is regenerated) and
               the hit record (16 B write), plus every LBVH node (128 B) and leaf triangle (48 B)
               it visits;
    wf_shade:  per live path the hit (16 B), the state (68 B read beyond depth 0, 68 B write for
               each survivor), the shading tables of a hit (instance 64 + indices 12 + positions 36
               + vertex normals 36 + material 96 = 244 B), 64 B per delta light, 48 B of
               environment sampling tables, one shadow record per shadow ray (40 B + 4 B queue
               entry), per pending path its count, next slot, pixel and total (24 B), and the FP64
               film update (48 B) once per finished path;
    wf_shadow: per shadow ray its queue entry (4 B), origin, direction and tmax (28 B), the
               visibility word (4 B), and the nodes and triangles it visits; the resolve step
               (per shadow ray the visibility and contribution, 16 B, plus the pending record) is
               counted with it;
    wf_trace:  the launch that traces depth d's closest rays and depth d - 1's shadow rays
               (opaque scenes): wf_extend's bytes and wf_shadow's without the resolve step.
    """
    rays = c["closest_rays"]
    hits = c["closest_hits"]
    shadow = c["shadow_rays"]
    later = max(rays - pixels, 0)  # paths that reached depth >= 1
    extend = 24 * later + 16 * rays + NODE_BYTES * c["closest_nodes"] + TRI_BYTES * c["closest_tris"]
    shade = (16 * rays + 4 * STATE_WORDS * later + 4 * STATE_WORDS * later + hits * (244 + 64 * delta_lights + (48 if env else 0))
             + 44 * shadow + 24 * hits + 48 * pixels)
    shadow_rays = (4 + 28 + 4) * shadow + NODE_BYTES * c["shadow_nodes"] + TRI_BYTES * c["shadow_tris"]
    shadow_b = shadow_rays + 16 * shadow + 24 * hits
    return {"wf_extend": extend, "wf_shade": shade, "wf_shadow": shadow_b, "wf_trace": extend + shadow_rays}


def host_cores():
    """The host cores this process may use: its CPU affinity, at most 16 (the GPU box gives one GPU a
    share of 16 cores, whatever the machine's count)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(packed, width, height, gpu_film):
    """The oracle (a scalar C port, OpenMP over pixels) on the host cores: one sample of every pixel
    of the same workload, plus a full-size parity check of that sample against the GPU's."""
    import oracle
    from verion_amd import pack

    threads = host_cores()
    t0 = time.perf_counter()
    o = oracle.OracleScene(packed)
    build_s = time.perf_counter() - t0
    par = pack.base_params(packed.info, packed.view, width, height, False, True)
    t0 = time.perf_counter()
    st = o.render(par, 1, first_sample=0, threads=threads)
    secs = time.perf_counter() - t0
    film = st["film"]
    rel = np.abs(gpu_film - film) / np.maximum(np.abs(film), 1e-6)
    parity = {"max_rel": float(rel.max()), "exact_fraction": float(np.mean(gpu_film == film)), "pixels": width * height,
              "sample": 0}
    base = {"value": round(width * height / secs / 1e6, 4), "unit": "Msamples/s", "cores": threads, "kind": "port",
            "label": "project CPU oracle (not the reference)",
            "sample": f"{width}x{height}, 1 spp (sample 0) of the same scene on the CPU oracle; "
                      f"{secs:.1f} s render, BVH build {build_s:.1f} s excluded",
            "cores_note": "every core of this process's CPU affinity, capped at the 16-core share of one GPU",
            "reference_baseline": "not measured: the reference's Vulkan backend on Mesa lavapipe needs rustc, slangc "
                                  "and a Vulkan ICD with ray queries, none of which is in the image"}
    return base, parity


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--spp-per-step", type=int, default=8, help="samples per pixel of one accumulate call")
    ap.add_argument("--pipeline", type=int, default=3,
                    help="sample passes the wavefront keeps in flight (1: one at a time, e.g. for kernel profiles)")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--integrator", choices=["wavefront", "megakernel"], default="wavefront")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-nrc", action="store_true", help="skip the NRC vs plain time-to-relMSE measurement")
    ap.add_argument("--no-batch", action="store_true", help="skip the 32-job batch measurement")
    ap.add_argument("--no-restir", action="store_true", help="skip the many-lights ReSTIR measurement")
    ap.add_argument("--no-reference-scenes", action="store_true",
                    help="skip the reference's own benchmark scenes (cube and the 5M-triangle scale scene)")
    ap.add_argument("--no-cornell", action="store_true", help="skip the Cornell box measurement (for profiler runs)")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shape (8 blocks per CU, 8x8 tiles)")
    ap.add_argument("--shape", type=str, default=None, help="fixed launch shape BLOCKS,TILE_W,TILE_H instead of tuning")
    ap.add_argument("--nrc-ref-spp", type=int, default=16384,
                    help="spp of the plain reference render the relMSE ladders are scored against")
    ap.add_argument("--no-counting", action="store_true",
                    help="skip the traversal-counting pass (for profiler runs: every launch is then a timed-workload launch)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (the film would be "
              "sharded by sample index and reduced once per render); run it with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from verion_amd import pack, scenes
    from verion_amd.backend import MEGAKERNEL, WAVEFRONT, Context, Renderer

    W, H = args.width, args.height
    pixels = W * H
    t0 = time.perf_counter()
    packed = pack.pack(scenes.procedural())
    ctx = Context()
    r = Renderer(ctx, packed)
    log(f"scene + LBVH ready in {time.perf_counter() - t0:.2f} s: {r.bvh_stats()}")
    r.set_integrator(WAVEFRONT if args.integrator == "wavefront" else MEGAKERNEL)
    r.set_wavefront_pipeline(args.pipeline)

    spp = args.spp_per_step
    warm = r.new_accumulation(W, H)
    for _ in range(args.warmup):
        r.accumulate(warm, spp)
    warm.close()
    tuning = None
    if args.shape:
        blocks, tw, th = (int(x) for x in args.shape.split(","))
        r.set_wavefront_blocks(blocks)
        r.set_wavefront_tiles(tw, th)
        tuning = {"grid_blocks": blocks, "tile": [tw, th], "fixed": True}
    elif args.integrator == "wavefront" and not args.no_tune:
        # launch shape (grid blocks, depth-0 tiles) by the project's autotuner, before timing
        from verion_amd.autotune import tune_wavefront

        best, table = tune_wavefront(r, W, H)
        tuning = {"grid_blocks": best[0], "tile": [best[1], best[2]],
                  "tried_ms_per_pass": [[list(c), round(ms, 3)] for c, ms in table]}
        log(f"autotuned wavefront: {tuning}")

    # timed region: K accumulate calls of `spp` samples from sample 0 (the wavefront keeps
    # `--pipeline` sample passes in flight); accumulate() returns after the GPU finished, so the clock
    # brackets device-complete work on both sides. No events between the launches.
    def timed_region(passes: int, staged: bool):
        r.set_wavefront_pipeline(passes)
        r.set_stage_timing(staged)
        acc = r.new_accumulation(W, H)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            r.accumulate(acc, spp)
        dt = time.perf_counter() - t0
        ks = acc.kernel_seconds() if staged else None
        acc.close()
        r.set_stage_timing(False)
        r.set_wavefront_pipeline(args.pipeline)
        return dt, ks

    elapsed, _ = timed_region(args.pipeline, False)
    value = args.steps * spp * pixels / elapsed / 1e6
    # the same calls with one pass at a time, and again with stage timing: HIP events recorded on
    # the library's stream after every launch (created before the passes) give each kernel's
    # launch time
    elapsed_one, _ = timed_region(1, False)
    elapsed_staged, kernel_s = timed_region(1, True)
    log(f"timed {args.steps} steps of {spp} spp in {elapsed:.3f} s ({args.pipeline} passes in flight; one at a time "
        f"{elapsed_one:.3f} s, with stage events {elapsed_staged:.3f} s); "
        f"extend, shade, shadow, resolve, trace seconds {kernel_s}")

    # traversal counters of one more pass (outside the timed region) for the algorithmic bytes
    if args.no_counting:
        counts = {k: 0 for k in ("closest_rays", "closest_nodes", "closest_tris", "shadow_rays", "shadow_nodes",
                                 "shadow_tris", "closest_hits")}
    else:
        r.set_counters(True)
        cnt_acc = r.new_accumulation(W, H)
        r.accumulate(cnt_acc, 1)
        counts = r.read_counters()
        r.set_counters(False)
    log(f"counters per pass: {counts}")
    delta_lights = packed.info.light_count
    algo = algorithmic_bytes(counts, pixels, delta_lights, packed.info.env_width > 0)
    names = ("wf_extend", "wf_shade", "wf_shadow", "wf_resolve", "wf_trace")
    per = dict(zip(names, kernel_s)) if kernel_s else {k: float("nan") for k in names}
    # launches per pass: one per depth; the trace kernel runs once more for the last depth's shadow rays
    per_pass = {k: 8 for k in names}
    per_pass["wf_trace"] = 9
    ran = [k for k in names if per[k] > 0]
    dom = max([k for k in ("wf_trace", "wf_extend", "wf_shade", "wf_shadow") if k in ran] or ["wf_extend"],
              key=lambda k: per[k])
    passes_timed = args.steps * spp
    dom_s = per[dom] / (passes_timed * per_pass[dom])
    dom_b = algo[dom] / per_pass[dom]
    achieved = dom_b / dom_s / 1e9
    traffic = traffic_from_profiles(dom)
    measured = None
    if traffic:
        gbs = traffic / dom_s / 1e9
        measured = {"gb_per_s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
                    "source": "rocprofv3 FETCH_SIZE + WRITE_SIZE per launch (copy-calibrated, committed profile) "
                              "over this run's launch time"}
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "achieved_is": "algorithmic bytes (every node and triangle fetch counted) over launch time",
                "hbm_measured": measured,
                "occupancy": occupancy_from_profiles(dom),
                "limiter": "instruction issue and dependent-fetch latency of the traversal: node and triangle "
                           "fetches hit L2 and the Infinity Cache, so measured HBM traffic stays far below peak",
                "launch_ms": round(dom_s * 1e3, 4), "algorithmic_bytes_per_launch": int(dom_b),
                "launch_time_from": "HIP events after every launch, one sample pass at a time",
                "ms_per_launch": {k: round(per[k] / (passes_timed * per_pass[k]) * 1e3, 4) for k in ran},
                "algorithmic_gb_per_s": {k: round(algo[k] / (per[k] / passes_timed) / 1e9, 2) for k in algo if k in ran}}

    line = {"metric": "Msamples/s at 1080p + time-to-rMSE 1e-3: NRC vs plain wavefront, 1 MI355X", "value": round(value, 3),
            "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(elapsed / args.steps * 1e3, 3),
            "value_one_pass_at_a_time": round(passes_timed * pixels / elapsed_one / 1e6, 3),
            "value_with_stage_events": round(passes_timed * pixels / elapsed_staged / 1e6, 3),
            "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "procedural 1,024 icospheres (1,310,720 tris), sun + constant env map",
                       "width": W, "height": H, "spp_per_step": spp, "max_depth": 8, "integrator": args.integrator,
                       "passes_in_flight": args.pipeline if args.integrator == "wavefront" else None,
                       "film": "fp64", "launch_shape": tuning},
            "roofline": roofline}
    if not args.no_cpu_baseline:
        gpu1 = r.new_accumulation(W, H)
        r.set_stage_timing(False)
        r.accumulate(gpu1, 1)
        base, parity = cpu_baseline(packed, W, H, gpu1.film())
        line["cpu_baseline"] = base
        line["parity_full_size"] = parity
    if args.integrator == "wavefront":
        line["megakernel_1080p"] = megakernel_same_scene(r, W, H, spp)
    if not args.no_cornell:
        line["cornell_256"] = cornell_256(ctx)
    if not args.no_nrc:
        line["time_to_rmse"] = time_to_rmse(W, H, args.nrc_ref_spp)
    if not args.no_restir:
        line["restir_many_lights"] = restir_many_lights(W, H)
    if not args.no_batch:
        line["batch"] = batch_jobs()
    if not args.no_reference_scenes:
        line["reference_benchmark"] = reference_benchmark()
    print(json.dumps(line), flush=True)


def megakernel_same_scene(r, W, H, spp, steps=8):
    """The other integrator on the same scene and image: the plain megakernel (persistent waves,
    every ray traced wave-cooperatively), `spp` samples per call like the timed steps (a call's
    samples run in parallel, 16 per dispatch at 1080p), after 2 untimed calls; and 1 spp per call."""
    from verion_amd.backend import MEGAKERNEL, WAVEFRONT

    r.set_integrator(MEGAKERNEL)
    r.set_stage_timing(False)
    out = {}
    for per_call, key in ((spp, "msamples_per_s"), (1, "msamples_per_s_1_spp_per_call")):
        acc = r.new_accumulation(W, H)
        for _ in range(2):
            r.accumulate(acc, per_call)
        n = steps if per_call > 1 else 16
        t0 = time.perf_counter()
        for _ in range(n):
            r.accumulate(acc, per_call)
        dt = time.perf_counter() - t0
        acc.close()
        out[key] = round(n * per_call * W * H / dt / 1e6, 1)
    out["spp_per_call"] = spp
    r.set_integrator(WAVEFRONT)
    log(f"megakernel at {W}x{H}: {out}")
    return out


def cornell_256(ctx):
    """The Cornell box workload (36 triangles, 256x256, 16 spp, depth 8): Msamples/s of both
    integrators over 64 accumulations of 16 spp each, after 4 untimed ones."""
    from verion_amd import scenes
    from verion_amd.backend import MEGAKERNEL, WAVEFRONT, Renderer

    r = Renderer(ctx, scenes.cornell())
    out = {"workload": "Cornell box, 256x256, 16 spp per accumulation"}
    for name, integ in (("megakernel", MEGAKERNEL), ("wavefront", WAVEFRONT)):
        r.set_integrator(integ)
        for _ in range(4):
            r.render(256, 256, 16)
        accs = [r.new_accumulation(256, 256) for _ in range(64)]
        t0 = time.perf_counter()
        for a in accs:
            r.accumulate(a, 16)
        dt = time.perf_counter() - t0
        for a in accs:
            a.close()
        out[name + "_msamples_per_s"] = round(64 * 256 * 256 * 16 / dt / 1e6, 1)
    log(f"cornell 256: {out}")
    return out


def batch_jobs():
    """32 seeded jobs (256x256 to 1080p, 4-256 spp, adaptive 0.01, scenes of 10k-1.3M triangles):
    one shared wavefront (autotuned, graph-replayed) against one job at a time in 8-spp chunks."""
    from verion_amd import batch_eval

    res = batch_eval.evaluate(log=log)
    keep = ("seconds", "msamples_per_s", "p50_latency_s", "p99_latency_s")
    return {"jobs": res["jobs"], "pixels_total": res["pixels_total"], "spp_range": res["spp_range"],
            "adaptive": res["adaptive"], "batched": {k: res["batched"][k] for k in keep + ("tuning",)},
            "one_at_a_time": {k: res["one_at_a_time"][k] for k in keep}, "films_checked_equal": res["films_checked_equal"]}


def reference_benchmark():
    """The reference's own throughput benchmark (scripts/bench.sh: benchmarks/cube.gltf and the
    generated 5M-triangle scale scene, 1024x768, 1,024 spp) on this backend, beside the GPU seconds
    the reference publishes for it (RTX 3080 Vulkan, Apple M4 Pro and M5)."""
    from verion_amd import reference_bench

    here = os.path.dirname(os.path.abspath(__file__))
    return reference_bench.evaluate(os.path.join(here, "tests", "golden", "cube.gltf"), log=log)


def restir_many_lights(width, height):
    """The many-lights workload (502,252 triangles, 10,000 emitters, 1 spp per frame over 64 frames,
    32 candidates, temporal + 5-neighbour spatial reuse): ms per frame and relMSE at 64 frames of
    plain, ReSTIR DI and DI + GI against a 1,024-spp plain reference, and plain's time to each
    method's error over the method's time."""
    from verion_amd import restir_eval

    res = restir_eval.evaluate(width, height, frames=64, ref_spp=1024, candidates=32, history=4, neighbours=5,
                               plain_frames=256, methods=("restir_di", "restir_di_gi"), log=log)
    out = {"scene": res["scene"], "frames": res["frames"], "reference_spp": res["reference_spp"],
           "candidates": res["candidates"], "neighbours": res["neighbours"], "relmse": res["relmse"]}
    for k in ("plain", "restir_di", "restir_di_gi"):
        m = res[k]
        out[k] = {"ms_per_frame": round(m["ms_per_frame"], 3), "relmse_64": m["relmse"]}
        if k != "plain":
            out[k]["plain_time_over_method_time"] = m["speedup_to_equal_error"]
    return out


def time_to_rmse(width, height, ref_spp):
    """The second half of the metric: GPU seconds to relMSE 1e-3 of the NRC and of the plain
    wavefront tracer, both 1 spp per frame on the same workload, scored against a plain reference
    render of `ref_spp` spp with an independent seed; and the NRC's time to the plain tracer's
    relMSE at 256 frames as a fraction of the plain tracer's time."""
    from verion_amd import nrc_eval

    res = nrc_eval.evaluate("procedural", width, height, ref_spp=ref_spp, max_frames=256, plain_max=2048,
                            match_frames=256, log=log)
    v = res["nrc"][0]
    tp, tn = res["plain"]["time_to_target_s"], v["time_to_target_s"]
    m = v["at_plain_error"] or {}
    return {"target_relmse": res["target_relmse"], "relmse": "mean over pixels and channels of (x-r)^2/(r^2+0.01)",
            "reference_spp": ref_spp, "seconds": "GPU, HIP events on the library stream (NRC: training included)",
            "plain_s": tp, "nrc_s": tn, "nrc_over_plain": (tn / tp) if (tn and tp) else None,
            "plain_relmse_256": m.get("plain_relmse"), "plain_s_256": m.get("plain_seconds"),
            "nrc_s_to_plain_relmse_256": m.get("nrc_seconds"), "fraction_at_plain_relmse_256": m.get("fraction"),
            "nrc_config": v["config"], "nrc_train_s": v["train_seconds"], "nrc_render_s": v["render_seconds"],
            "plain_ladder": res["plain"]["ladder"], "nrc_ladder": v["ladder"]}


if __name__ == "__main__":
    main()
