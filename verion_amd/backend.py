"""Python mirror of Verion's backend contract over the C ABI of libverion_hip.so.

`Renderer` follows `trait RayTracingBackend` (crates/verion-gpu/src/lib.rs:12-58) and
`Accumulation` follows `trait Accumulation` (lib.rs:74-93), with the same method names, argument
meaning and error behaviour (the reference's `Error` variants become `BackendError.status`).
There is no CPU fallback: without the built library every constructor raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import abi
from .pack import Packed, UpdatePlan, View, pack
from .scene import Scene

# the in-tree build; VERION_HIP_LIBRARY names another build of the same sources (e.g. one with
# different occupancy targets) to load instead
LIB_PATH = os.environ.get("VERION_HIP_LIBRARY") or os.path.join(os.path.dirname(os.path.abspath(__file__)), "libverion_hip.so")

OK, NO_DEVICE, UNSUPPORTED, OUT_OF_MEMORY, KERNEL, SCENE, INVALID_ARGUMENT = range(7)
MEGAKERNEL, WAVEFRONT = 0, 1


class BackendError(RuntimeError):
    def __init__(self, status: int, message: str, needed: int = 0, available: int = 0):
        super().__init__(f"[{status}] {message}")
        self.status = status
        self.message = message
        self.needed = needed
        self.available = available


@dataclass
class Adaptive:
    """lib.rs:110-126 (defaults threshold 0.02, min_spp 16)."""

    threshold: float = 0.02
    min_spp: int = 16

    def active(self, n: float, s: float, s2: float) -> bool:
        """Adaptive::active (lib.rs:154-161)."""
        if n < self.min_spp:
            return True
        mean = s / n
        var = max(s2 / n - mean * mean, 0.0) * n / (n - 1.0)
        return (var / n) ** 0.5 > self.threshold * max(mean, 0.05)


@dataclass
class Restir:
    """lib.rs:131-149 (defaults: DI on, GI off, 4 candidates, history 4)."""

    di: bool = True
    gi: bool = False
    candidates: int = 4
    history: int = 4
    neighbours: int = 0  # random neighbours per reuse; 0 = the reference's 2 (pt_core.slang:706)


_lib = None


def library() -> C.CDLL:
    """Loads libverion_hip.so; raises ImportError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m verion_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, u32, u64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
    u32p, f32p, f64p = C.POINTER(C.c_uint32), C.POINTER(C.c_float), C.POINTER(C.c_double)
    sig = {
        "vh_last_error": (C.c_char_p, []),
        "vh_last_error_needed": (u64, []),
        "vh_last_error_available": (u64, []),
        "vh_context_create": (i32, [C.POINTER(vp)]),
        "vh_context_destroy": (None, [vp]),
        "vh_memory": (i32, [vp, C.POINTER(abi.MemoryUsage)]),
        "vh_renderer_create": (i32, [vp, C.POINTER(abi.SceneTables), C.POINTER(vp)]),
        "vh_renderer_destroy": (None, [vp]),
        "vh_renderer_update": (i32, [vp, C.POINTER(abi.SceneTables), C.POINTER(abi.UpdatePlanS), C.POINTER(abi.UpdateReport)]),
        "vh_set_camera": (i32, [vp, C.POINTER(abi.ViewS)]),
        "vh_set_multiscatter": (i32, [vp, i32]),
        "vh_set_integrator": (i32, [vp, i32]),
        "vh_integrator": (i32, [vp]),
        "vh_set_stage_timing": (i32, [vp, i32]),
        "vh_set_adaptive": (i32, [vp, C.POINTER(abi.Adaptive)]),
        "vh_set_restir": (i32, [vp, C.POINTER(abi.Restir)]),
        "vh_set_seed": (i32, [vp, u32]),
        "vh_set_wavefront_blocks": (i32, [vp, u32]),
        "vh_set_wavefront_tiles": (i32, [vp, u32, u32]),
        "vh_set_wavefront_pipeline": (i32, [vp, u32]),
        "vh_accumulation_create": (i32, [vp, u32, u32, i32, C.POINTER(vp)]),
        "vh_accumulation_destroy": (None, [vp]),
        "vh_accumulate": (i32, [vp, vp, u32]),
        "vh_acc_samples": (u32, [vp]),
        "vh_acc_gpu_seconds": (C.c_double, [vp]),
        "vh_acc_stage_seconds": (i32, [vp, f64p]),
        "vh_acc_kernel_seconds": (i32, [vp, f64p]),
        "vh_acc_wavefront_bytes": (u64, [vp]),
        "vh_acc_restir_bytes": (u64, [vp]),
        "vh_acc_active_pixels": (u64, [vp]),
        "vh_acc_read_film": (i32, [vp, f64p]),
        "vh_acc_read_stats": (i32, [vp, f64p]),
        "vh_acc_read_lighting": (i32, [vp, f32p]),
        "vh_acc_mean": (i32, [vp, f32p]),
        "vh_render_normals": (i32, [vp, u32, u32, f32p]),
        "vh_render_aovs": (i32, [vp, u32, u32, f32p, u32p, u32p, f32p]),
        "vh_primary_hits": (i32, [vp, u32, u32, u32, u32p]),
        "vh_trace_rays": (i32, [vp, u64, f32p, C.c_float, u32p]),
        "vh_bvh_stats": (i32, [vp, C.POINTER(C.c_uint64)]),
        "vh_set_counters": (i32, [vp, i32]),
        "vh_read_counters": (i32, [vp, C.POINTER(C.c_uint64)]),
        "vh_nrc_create": (i32, [vp, C.POINTER(abi.NrcConfig), C.POINTER(vp)]),
        "vh_nrc_destroy": (None, [vp]),
        "vh_nrc_init_params": (i32, [u32, f32p, u64]),
        "vh_trace_limits": (i32, [u32p]),
        "vh_nrc_init": (i32, [vp, u32]),
        "vh_nrc_set_params": (i32, [vp, f32p, u64]),
        "vh_nrc_get_params": (i32, [vp, f32p, u64]),
        "vh_nrc_infer": (i32, [vp, u64, f32p, f32p]),
        "vh_nrc_gradient": (i32, [vp, u64, f32p, f32p, f32p, f32p]),
        "vh_nrc_gradient_noalloc": (i32, [vp, u64, vp, vp, vp, vp]),
        "vh_nrc_train": (i32, [vp, u64, f32p, f32p, u32, f32p]),
        "vh_nrc_accumulate": (i32, [vp, vp, u32]),
        "vh_nrc_stats": (i32, [vp, f64p]),
        "vh_batch_create": (i32, [vp, C.POINTER(abi.BatchJob), u32, C.POINTER(vp)]),
        "vh_batch_destroy": (None, [vp]),
        "vh_batch_reset": (i32, [vp]),
        "vh_batch_set_tuning": (i32, [vp, C.POINTER(abi.BatchTuning)]),
        "vh_batch_get_tuning": (i32, [vp, C.POINTER(abi.BatchTuning)]),
        "vh_batch_run": (i32, [vp, i32, u32]),
        "vh_batch_job_stats": (i32, [vp, u32, f64p]),
        "vh_batch_tuning_report": (i32, [vp, C.POINTER(abi.BatchTuning), f64p, u32, C.POINTER(C.c_uint32)]),
        "vh_batch_read_film": (i32, [vp, u32, f64p]),
        "vh_batch_read_stats": (i32, [vp, u32, f64p]),
        "vh_batch_mean": (i32, [vp, u32, f32p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc: int):
    if rc != OK:
        L = library()
        raise BackendError(rc, L.vh_last_error().decode(), L.vh_last_error_needed(), L.vh_last_error_available())


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(C.POINTER(t))


class _Handle:
    """A library handle. Handles nest (accumulation in renderer in context) and the library wants
    them destroyed children first; freeing a handle frees its live children before it, so the
    order in which the garbage collector finalizes the Python objects does not matter."""

    __slots__ = ("h", "destroy", "kids", "parent")

    def __init__(self, h, destroy, parent: Optional["_Handle"] = None):
        self.h, self.destroy, self.kids, self.parent = h, destroy, [], parent
        if parent is not None:
            parent.kids.append(self)

    def free(self):
        if self.h is None:
            return
        for k in reversed(list(self.kids)):  # newest first: a batch goes before the renderers it uses
            k.free()
        self.destroy(self.h)
        self.h = None
        if self.parent is not None and self in self.parent.kids:
            self.parent.kids.remove(self)


class _Owned:
    _handle: Optional[_Handle] = None

    @property
    def h(self):
        return self._handle.h if self._handle is not None else None

    def close(self):
        if self._handle is not None:
            self._handle.free()

    def __del__(self):
        self.close()


class Context(_Owned):
    """RayTracingBackend::Context: the device and its stream."""

    def __init__(self):
        L = library()
        h = C.c_void_p()
        _check(L.vh_context_create(C.byref(h)))
        self._handle = _Handle(h, L.vh_context_destroy)

    def memory(self) -> abi.MemoryUsage:
        m = abi.MemoryUsage()
        _check(library().vh_memory(self.h, C.byref(m)))
        return m


class Accumulation(_Owned):
    """trait Accumulation (lib.rs:74-93) over the library's running sums."""

    def __init__(self, renderer: "Renderer", width: int, height: int, lighting: bool):
        self.renderer = renderer
        self.width, self.height, self.lighting_on = width, height, lighting
        h = C.c_void_p()
        _check(library().vh_accumulation_create(renderer.h, width, height, int(lighting), C.byref(h)))
        self._handle = _Handle(h, library().vh_accumulation_destroy, renderer._handle)

    @property
    def pixels(self) -> int:
        return self.width * self.height

    def samples(self) -> int:
        return library().vh_acc_samples(self.h)

    def gpu_seconds(self) -> float:
        return library().vh_acc_gpu_seconds(self.h)

    def stage_seconds(self):
        out = np.zeros(2, np.float64)
        rc = library().vh_acc_stage_seconds(self.h, _p(out, C.c_double))
        return None if rc != OK else (float(out[0]), float(out[1]))

    def kernel_seconds(self):
        """(extend, shade, deferred shadow rays, their resolve, trace) GPU seconds, or None; trace
        is closest and shadow rays in one launch (opaque scenes with deferred shadow rays)."""
        out = np.zeros(5, np.float64)
        rc = library().vh_acc_kernel_seconds(self.h, _p(out, C.c_double))
        return None if rc != OK else tuple(float(x) for x in out)

    def wavefront_bytes(self) -> int:
        return library().vh_acc_wavefront_bytes(self.h)

    def restir_bytes(self) -> int:
        return library().vh_acc_restir_bytes(self.h)

    def film(self) -> np.ndarray:
        out = np.zeros(3 * self.pixels, np.float64)
        _check(library().vh_acc_read_film(self.h, _p(out, C.c_double)))
        return out

    def stats(self) -> Optional[np.ndarray]:
        out = np.zeros(3 * self.pixels, np.float64)
        rc = library().vh_acc_read_stats(self.h, _p(out, C.c_double))
        return out if rc == OK else None

    def lighting_sums(self) -> Optional[np.ndarray]:
        out = np.zeros(65 * self.pixels, np.float32)
        rc = library().vh_acc_read_lighting(self.h, _p(out, C.c_float))
        return out if rc == OK else None

    def mean(self) -> np.ndarray:
        """Mean radiance per pixel, (pixels, 3) float32."""
        out = np.zeros(3 * self.pixels, np.float32)
        _check(library().vh_acc_mean(self.h, _p(out, C.c_float)))
        return out.reshape(-1, 3)

    def sample_counts(self) -> Optional[np.ndarray]:
        st = self.stats()
        return None if st is None else st.reshape(-1, 3)[:, 0].astype(np.uint32)

    def active_pixels(self) -> int:
        return library().vh_acc_active_pixels(self.h)


class Renderer(_Owned):
    """RayTracingBackend on one MI355X: upload, LBVH, megakernel and wavefront integrators."""

    def __init__(self, ctx: Context, scene):
        self.ctx = ctx
        self.packed: Packed = scene if isinstance(scene, Packed) else pack(scene)
        self._tables, self._keep = abi.tables_struct(self.packed)
        h = C.c_void_p()
        _check(library().vh_renderer_create(ctx.h, C.byref(self._tables), C.byref(h)))
        self._handle = _Handle(h, library().vh_renderer_destroy, ctx._handle)

    def update(self, scene, changes=None) -> abi.UpdateReport:
        """RayTracingBackend::update (lib.rs:27): `changes` (a scene.Changes, from
        Scene.take_changes) selects what is re-uploaded and redone through pack.UpdatePlan; None
        replaces everything."""
        self.packed = scene if isinstance(scene, Packed) else pack(scene)
        self._tables, self._keep = abi.tables_struct(self.packed)
        plan, blases = None, None
        if changes is not None:
            up = UpdatePlan.new(changes)
            blases = np.array(up.blases, np.uint32)
            plan = abi.UpdatePlanS(int(up.full), int(up.materials), int(up.lights), int(up.emitters), int(up.instances),
                                   int(up.vertices), int(up.tlas), int(up.camera),
                                   abi.ptr(blases, C.c_uint32) if blases.size else None, blases.size)
        rep = abi.UpdateReport()
        _check(library().vh_renderer_update(self.h, C.byref(self._tables), C.byref(plan) if plan else None, C.byref(rep)))
        return rep

    def set_camera(self, view: View):
        v = abi.view_struct(view)
        _check(library().vh_set_camera(self.h, C.byref(v)))

    def set_multiscatter(self, on: bool):
        _check(library().vh_set_multiscatter(self.h, int(on)))

    def set_integrator(self, integrator: int):
        _check(library().vh_set_integrator(self.h, int(integrator)))

    def integrator(self) -> int:
        return library().vh_integrator(self.h)

    def set_stage_timing(self, on: bool):
        _check(library().vh_set_stage_timing(self.h, int(on)))

    def set_adaptive(self, adaptive: Optional[Adaptive]):
        a = abi.Adaptive(adaptive.threshold, adaptive.min_spp) if adaptive else None
        _check(library().vh_set_adaptive(self.h, C.byref(a) if a else None))

    def set_restir(self, restir: Optional[Restir]):
        r = abi.Restir(int(restir.di), int(restir.gi), restir.candidates, restir.history, restir.neighbours) if restir else None
        _check(library().vh_set_restir(self.h, C.byref(r) if r else None))

    def set_seed(self, seed: int):
        _check(library().vh_set_seed(self.h, seed))

    def set_wavefront_blocks(self, blocks: int):
        _check(library().vh_set_wavefront_blocks(self.h, blocks))

    def set_wavefront_tiles(self, tile_width: int, tile_height: int):
        """Depth-0 pixel order of the wavefront: tiles, or (0, 0) for scanlines."""
        _check(library().vh_set_wavefront_tiles(self.h, tile_width, tile_height))

    def set_wavefront_pipeline(self, passes: int):
        """Sample passes of one accumulate call the plain wavefront keeps in flight (1 to 4)."""
        _check(library().vh_set_wavefront_pipeline(self.h, passes))

    def new_accumulation(self, width: int, height: int, lighting: bool = False) -> Accumulation:
        return Accumulation(self, width, height, lighting)

    def accumulate(self, acc: Accumulation, samples: int):
        _check(library().vh_accumulate(self.h, acc.h, samples))

    def render(self, width: int, height: int, spp: int) -> np.ndarray:
        """One accumulation of `spp` samples, its mean (the reference tests' `render`)."""
        acc = self.new_accumulation(width, height)
        self.accumulate(acc, spp)
        return acc.mean()

    def render_normals(self, width: int, height: int) -> np.ndarray:
        """RayTracingBackend::render_normals (lib.rs:55): (pixels, 3) float32, row-major from the top-left."""
        out = np.zeros((width * height, 3), np.float32)
        _check(library().vh_render_normals(self.h, width, height, _p(out, C.c_float)))
        return out

    def render_aovs(self, width: int, height: int) -> dict:
        """RayTracingBackend::render_aovs (lib.rs:57): the fields of `Aovs` (lib.rs:165-176)."""
        n = width * height
        ad = np.zeros((n, 4), np.float32)
        oid = np.zeros(n, np.uint32)
        mid = np.zeros(n, np.uint32)
        cov = np.zeros(n, np.float32)
        _check(library().vh_render_aovs(self.h, width, height, _p(ad, C.c_float), _p(oid, C.c_uint32), _p(mid, C.c_uint32),
                                        _p(cov, C.c_float)))
        return {"albedo": ad[:, :3].copy(), "depth": ad[:, 3].copy(), "object_id": oid, "material_id": mid, "coverage": cov}

    def primary_hits(self, width: int, height: int, sample: int = 0) -> np.ndarray:
        out = np.zeros((width * height, 4), np.uint32)
        _check(library().vh_primary_hits(self.h, width, height, sample, _p(out, C.c_uint32)))
        return out

    def trace_rays(self, rays: np.ndarray, tmax: float = 1.0e30) -> np.ndarray:
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        out = np.zeros((len(rays), 4), np.uint32)
        _check(library().vh_trace_rays(self.h, len(rays), _p(rays, C.c_float), tmax, _p(out, C.c_uint32)))
        return out

    def bvh_stats(self) -> dict:
        out = np.zeros(4, np.uint64)
        _check(library().vh_bvh_stats(self.h, _p(out, C.c_uint64)))
        return {"nodes": int(out[0]), "leaves": int(out[1]), "depth": int(out[2]), "triangles": int(out[3])}

    def set_counters(self, on: bool):
        _check(library().vh_set_counters(self.h, int(on)))

    def read_counters(self) -> dict:
        """Traversal counts since the last read (counting must be on), then cleared."""
        out = np.zeros(8, np.uint64)
        _check(library().vh_read_counters(self.h, _p(out, C.c_uint64)))
        k = ["closest_rays", "closest_nodes", "closest_tris", "shadow_rays", "shadow_nodes", "shadow_tris", "closest_hits"]
        return {name: int(v) for name, v in zip(k, out)}


class Nrc(_Owned):
    """The neural radiance cache of a renderer (vh_nrc_*): a 96-64x5-16 MLP trained online on the
    tracer's own paths with BF16-MFMA kernels and FP32 Adam; `accumulate` is the NRC integrator."""

    PARAMS = 23888

    def __init__(self, renderer: Renderer, train_paths: int = 65536, train_steps: int = 8, learning_rate: float = 1e-2,
                 terminate_depth: int = 1, ema_decay: float = 0.95, warmup_frames: int = 16, minibatch: bool = False):
        self.renderer = renderer
        cfg = abi.NrcConfig(train_paths, train_steps, learning_rate, terminate_depth, ema_decay, warmup_frames, int(minibatch))
        h = C.c_void_p()
        _check(library().vh_nrc_create(renderer.h, C.byref(cfg), C.byref(h)))
        self._handle = _Handle(h, library().vh_nrc_destroy, renderer._handle)

    @staticmethod
    def init_params(seed: int = 0) -> np.ndarray:
        """The library's seeded initial parameters (vh_nrc_init_params; host only, no device)."""
        out = np.zeros(Nrc.PARAMS, np.float32)
        _check(library().vh_nrc_init_params(seed, _p(out, C.c_float), out.size))
        return out

    def init(self, seed: int = 0):
        """Restarts the cache from the library's seeded initialisation (vh_nrc_init)."""
        _check(library().vh_nrc_init(self.h, seed))

    def set_params(self, theta: np.ndarray):
        theta = np.ascontiguousarray(theta, np.float32)
        _check(library().vh_nrc_set_params(self.h, _p(theta, C.c_float), theta.size))

    def params(self) -> np.ndarray:
        out = np.zeros(self.PARAMS, np.float32)
        _check(library().vh_nrc_get_params(self.h, _p(out, C.c_float), out.size))
        return out

    def infer(self, raw: np.ndarray) -> np.ndarray:
        raw = np.ascontiguousarray(raw, np.float32).reshape(-1, 12)
        out = np.zeros((len(raw), 3), np.float32)
        _check(library().vh_nrc_infer(self.h, len(raw), _p(raw, C.c_float), _p(out, C.c_float)))
        return out

    def gradient(self, raw: np.ndarray, target: np.ndarray):
        raw = np.ascontiguousarray(raw, np.float32).reshape(-1, 12)
        target = np.ascontiguousarray(target, np.float32).reshape(-1, 3)
        g = np.zeros(self.PARAMS, np.float32)
        loss = np.zeros(1, np.float32)
        _check(library().vh_nrc_gradient(self.h, len(raw), _p(raw, C.c_float), _p(target, C.c_float), _p(g, C.c_float),
                                         _p(loss, C.c_float)))
        return g, float(loss[0])

    def train(self, raw: np.ndarray, target: np.ndarray, steps: int = 1) -> float:
        raw = np.ascontiguousarray(raw, np.float32).reshape(-1, 12)
        target = np.ascontiguousarray(target, np.float32).reshape(-1, 3)
        loss = np.zeros(1, np.float32)
        _check(library().vh_nrc_train(self.h, len(raw), _p(raw, C.c_float), _p(target, C.c_float), steps, _p(loss, C.c_float)))
        return float(loss[0])

    def accumulate(self, acc: Accumulation, frames: int):
        _check(library().vh_nrc_accumulate(self.h, acc.h, frames))

    def stats(self) -> dict:
        out = np.zeros(4, np.float64)
        _check(library().vh_nrc_stats(self.h, _p(out, C.c_double)))
        return {"loss": float(out[0]), "samples": int(out[1]), "train_seconds": float(out[2]), "render_seconds": float(out[3])}


class Batch(_Owned):
    """Many render jobs in one shared wavefront (vh_batch_*): `jobs` is a list of dicts with
    renderer, width, height, spp and optionally adaptive (Adaptive or None) and seed."""

    def __init__(self, ctx: Context, jobs):
        self.ctx = ctx
        self.jobs = [dict(j) for j in jobs]
        self.renderers = [j["renderer"] for j in self.jobs]  # they must outlive the batch
        arr = (abi.BatchJob * len(self.jobs))()
        for k, j in enumerate(self.jobs):
            a = j.get("adaptive")
            arr[k].renderer = j["renderer"].h
            arr[k].width, arr[k].height, arr[k].spp = j["width"], j["height"], j["spp"]
            arr[k].adaptive = abi.Adaptive(a.threshold, a.min_spp) if a else abi.Adaptive(0.0, 0)
            arr[k].seed = j.get("seed", 0)
        h = C.c_void_p()
        _check(library().vh_batch_create(ctx.h, arr, len(self.jobs), C.byref(h)))
        self._handle = _Handle(h, library().vh_batch_destroy, ctx._handle)

    def reset(self):
        _check(library().vh_batch_reset(self.h))

    def set_tuning(self, tuning):
        """(grid_blocks, tile_width, tile_height), or None for automatic tuning."""
        if tuning is None:
            _check(library().vh_batch_set_tuning(self.h, None))
        else:
            t = abi.BatchTuning(*tuning)
            _check(library().vh_batch_set_tuning(self.h, C.byref(t)))

    def tuning(self):
        t = abi.BatchTuning()
        _check(library().vh_batch_get_tuning(self.h, C.byref(t)))
        return (t.grid_blocks, t.tile_width, t.tile_height)

    def run(self, graph: bool = True, max_passes: int = 0):
        _check(library().vh_batch_run(self.h, int(graph), max_passes))

    def job_stats(self, job: int) -> dict:
        out = np.zeros(4, np.float64)
        _check(library().vh_batch_job_stats(self.h, job, _p(out, C.c_double)))
        return {"latency_s": float(out[0]), "passes": int(out[1]), "paths": int(out[2]), "batch_s": float(out[3])}

    def tuning_report(self):
        cap = 64
        tried = (abi.BatchTuning * cap)()
        ms = np.zeros(cap, np.float64)
        n = C.c_uint32()
        _check(library().vh_batch_tuning_report(self.h, tried, _p(ms, C.c_double), cap, C.byref(n)))
        return [((tried[i].grid_blocks, tried[i].tile_width, tried[i].tile_height), float(ms[i])) for i in range(min(n.value, cap))]

    def film(self, job: int) -> np.ndarray:
        j = self.jobs[job]
        out = np.zeros(3 * j["width"] * j["height"], np.float64)
        _check(library().vh_batch_read_film(self.h, job, _p(out, C.c_double)))
        return out

    def stats(self, job: int) -> np.ndarray:
        j = self.jobs[job]
        out = np.zeros(3 * j["width"] * j["height"], np.float64)
        _check(library().vh_batch_read_stats(self.h, job, _p(out, C.c_double)))
        return out

    def mean(self, job: int) -> np.ndarray:
        j = self.jobs[job]
        out = np.zeros((j["width"] * j["height"], 3), np.float32)
        _check(library().vh_batch_mean(self.h, job, _p(out, C.c_float)))
        return out
