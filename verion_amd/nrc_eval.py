"""Time-to-quality of the neural radiance cache against the plain wavefront tracer.

Both render 1 spp per frame into an FP64 film at the same resolution; the reference is a plain
render with many samples. Each method advances along a doubling frame ladder and is scored by
relMSE after every rung; times are the GPU seconds the library reports (the NRC's include its
training paths and Adam steps). Reported: time to relMSE 1e-3 (or "not reached" with the floor),
and the NRC's time to reach the plain tracer's error at `match_frames` frames, as a fraction of
the plain tracer's time.

    python -m verion_amd.nrc_eval --scene cornell --width 1920 --height 1080
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import quality, scenes
from .backend import WAVEFRONT, Context, Nrc, Renderer


DEFAULT_VARIANT = {"train_paths": 65536, "train_steps": 8, "learning_rate": 1e-2, "terminate_depth": 1, "ema_decay": 0.95,
                   "warmup_frames": 16}


def evaluate(scene: str = "cornell", width: int = 1920, height: int = 1080, ref_spp: int = 16384, max_frames: int = 256,
             plain_max: int = 4096, match_frames: int = 256, variants=None, target: float = 1e-3, seed: int = 1,
             log=lambda *a: None) -> dict:
    """Renders the reference, the plain ladder, then one NRC ladder per variant (a dict of
    `Nrc` arguments over DEFAULT_VARIANT), each from freshly initialised parameters."""
    variants = variants or [{}]
    ctx = Context()
    s = scenes.cornell() if scene == "cornell" else scenes.procedural()
    r = Renderer(ctx, s)
    r.set_integrator(WAVEFRONT)
    t0 = time.perf_counter()
    r.set_seed(0x5EED0000 + seed)  # the reference's samples are independent of both methods'
    ref_acc = r.new_accumulation(width, height)
    chunk = 256
    done = 0
    while done < ref_spp:
        r.accumulate(ref_acc, min(chunk, ref_spp - done))
        done += min(chunk, ref_spp - done)
        if done % 1024 == 0 or done == ref_spp:
            log(f"reference: {done}/{ref_spp} spp")
    ref = ref_acc.mean().astype(np.float64)
    ref_s = ref_acc.gpu_seconds()
    ref_acc.close()
    log(f"reference {ref_spp} spp in {ref_s:.1f} s GPU ({time.perf_counter() - t0:.1f} s wall)")
    r.set_seed(seed)

    plain = r.new_accumulation(width, height)

    def plain_step(k):
        before = plain.gpu_seconds()
        r.accumulate(plain, k)
        return plain.gpu_seconds() - before

    pts_plain = quality.ladder(plain_step, plain.mean, ref, plain_max, stop_below=target)
    plain.close()
    log("plain", [(f, round(t, 4), e) for f, t, e in pts_plain])
    tp = quality.time_to([(t, e) for _, t, e in pts_plain], target)
    at = [p for p in pts_plain if p[0] == match_frames]

    results = []
    for v in variants:
        cfg = dict(DEFAULT_VARIANT, **v)
        cache = Nrc(r, **cfg)
        cache.init(seed)
        acc = r.new_accumulation(width, height)

        def nrc_step(k):
            before = acc.gpu_seconds()
            cache.accumulate(acc, k)
            return acc.gpu_seconds() - before

        pts_nrc = quality.ladder(nrc_step, acc.mean, ref, max_frames, stop_below=target)
        acc.close()
        st = cache.stats()
        cache.close()
        log("nrc", cfg, [(f, round(t, 4), e) for f, t, e in pts_nrc], st)
        tn = quality.time_to([(t, e) for _, t, e in pts_nrc], target)
        match = None
        if at:
            e_match, t_match = at[0][2], at[0][1]
            t_nrc = quality.time_to([(t, e) for _, t, e in pts_nrc], e_match)
            match = {"plain_frames": match_frames, "plain_relmse": e_match, "plain_seconds": t_match,
                     "nrc_seconds": t_nrc, "fraction": (t_nrc / t_match) if t_nrc else None}
        results.append({"config": cfg, "time_to_target_s": tn, "floor_relmse": min(e for _, _, e in pts_nrc),
                        "ladder": [[f, round(t, 5), e] for f, t, e in pts_nrc], **st, "at_plain_error": match})
    return {
        "scene": scene, "width": width, "height": height, "reference_spp": ref_spp, "relmse_eps": quality.EPS,
        "target_relmse": target,
        "plain": {"time_to_target_s": tp, "ladder": [[f, round(t, 5), e] for f, t, e in pts_plain],
                  "msamples_per_s": width * height * pts_plain[-1][0] / pts_plain[-1][1] / 1e6},
        "nrc": results,
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--scene", choices=["cornell", "procedural"], default="cornell")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--ref-spp", type=int, default=16384)
    ap.add_argument("--frames", type=int, default=256)
    ap.add_argument("--plain-max", type=int, default=4096)
    ap.add_argument("--variant", action="append", default=[],
                    help='JSON object of Nrc arguments, e.g. \'{"ema_decay": 0.99, "warmup_frames": 16}\'; repeatable')
    a = ap.parse_args()
    variants = [json.loads(v) for v in a.variant] or [{}]
    res = evaluate(a.scene, a.width, a.height, a.ref_spp, a.frames, a.plain_max, min(a.frames, 256), variants,
                   log=lambda *x: print(*x, file=sys.stderr, flush=True))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
