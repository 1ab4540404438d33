"""Scene-edit latency on the 1,024-sphere procedural scene (1,310,720 triangles), after the
reference's update bench (crates/verion-gpu-metal/examples/update_bench.rs:30-90): a full rebuild
(a new renderer) against each kind of edit through `update`, and the first preview after it.

Each edit is packed on the host first (pack.py, not timed: the reference's packer is Rust); the
update column is the wall time of vh_renderer_update (uploads, LBVH refit or rebuild, waits), the
preview column a 512x384 accumulation of `spp` samples on the wavefront. Every edited renderer is
checked against a renderer built from scratch on the edited scene (same image, bit for bit).

    python -m verion_amd.update_eval [--spp 4]
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import sys
import time

import numpy as np

from . import scene as sc
from . import scenes
from .backend import WAVEFRONT, Context, Renderer
from .pack import pack


def preview(r: Renderer, spp: int) -> np.ndarray:
    acc = r.new_accumulation(512, 384)
    r.accumulate(acc, spp)
    return acc.mean()


def moved(inst: sc.Instance, dx: float) -> sc.Instance:
    t = [list(row) for row in inst.transform]
    t[0][3] += dx
    return dataclasses.replace(inst, transform=tuple(tuple(row) for row in t))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--spp", type=int, default=4)
    args = ap.parse_args(argv)
    ctx = Context()
    s = scenes.procedural(spheres=1024)
    s.take_changes()
    t0 = time.perf_counter()
    r = Renderer(ctx, s)
    rebuild = time.perf_counter() - t0
    r.set_integrator(WAVEFRONT)
    preview(r, args.spp)
    rows = [{"edit": "full rebuild (new renderer, pack included)", "update_s": rebuild}]
    rng = np.random.default_rng(3)

    def edit(name: str):
        changes = s.take_changes()
        p = pack(s)
        t = time.perf_counter()
        rep = r.update(p, changes)
        upd = time.perf_counter() - t
        t = time.perf_counter()
        img = preview(r, args.spp)
        first = time.perf_counter() - t
        fresh = Renderer(ctx, p)
        fresh.set_integrator(WAVEFRONT)
        same = bool(np.array_equal(img.view(np.uint32), preview(fresh, args.spp).view(np.uint32)))
        fresh.close()
        row = {"edit": name, "update_s": upd, "preview_s": first, "blases_built": int(rep.blases_built),
               "tlas_built": int(rep.tlas_built), "bvh_work": ["none", "refit", "rebuild"][rep.bvh_work],
               "bytes_uploaded": int(rep.bytes_uploaded), "same_as_fresh": same}
        rows.append(row)
        print(json.dumps(row), flush=True)

    s.set_material(0, dataclasses.replace(s.materials[0], roughness=0.9, base_color=(0.9, 0.1, 0.1)))
    edit("material")
    light = s.lights[0]
    s.set_light(0, dataclasses.replace(light, intensity=light.intensity * 1.5))
    edit("light")
    cam = s.camera
    s.set_camera(dataclasses.replace(cam, transform=moved(sc.Instance(0, 0, cam.transform), 0.5).transform))
    edit("camera")
    s.set_instance(7, moved(s.instances[7], 0.3))
    edit("move one instance")
    for k in rng.choice(1024, 64, replace=False):
        s.set_instance(int(k), moved(s.instances[int(k)], float(rng.uniform(-0.5, 0.5))))
    edit("move 64 instances")
    mesh = s.meshes[0]
    s.set_mesh(0, sc.Mesh(mesh.positions * np.float32(1.02), mesh.indices, mesh.uvs, mesh.normals))
    edit("deform the shared mesh")
    for k in range(1024):
        inst = s.instances[k]
        t = [list(row) for row in inst.transform]
        for a in range(3):
            t[a][3] = float(rng.uniform(-10.0, 10.0))
        s.set_instance(k, dataclasses.replace(inst, transform=tuple(tuple(row) for row in t)))
    edit("scatter every instance")
    print(json.dumps({"rows": rows}))
    return 0 if all(r_.get("same_as_fresh", True) for r_ in rows) else 1


if __name__ == "__main__":
    sys.exit(main())
