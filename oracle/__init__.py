"""ORACLE - test infrastructure only.

The CPU restatement of Verion's path tracer (pt_oracle.c, or_math.h) and its ctypes wrapper. Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this package, and only to
check or time against it; the product library (verion_amd) never loads it.

The reference's own path (Slang kernels + Rust host) cannot be built here: it needs rustc/cargo,
slangc and a Vulkan ray-query device, none of which exists in this image, so there is no
oracle/_ref build.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
SOURCES = [os.path.join(HERE, "pt_oracle.c"), os.path.join(HERE, "or_math.h")]
CFLAGS = ["-O2", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
          "-Wall", "-Wextra", "-Wdouble-promotion", "-Werror=double-promotion", "-Wno-unused-parameter"]


def build() -> str:
    """Compiles liboracle.so with gcc (the recipe is also oracle/Makefile)."""
    cmd = ["gcc", *CFLAGS, "-o", LIB + ".tmp", SOURCES[0], "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB) or any(os.path.getmtime(s) > os.path.getmtime(LIB) for s in SOURCES if os.path.exists(s)):
            build()
        L = C.CDLL(LIB)
        u32p, f32p, f64p = C.POINTER(C.c_uint32), C.POINTER(C.c_float), C.POINTER(C.c_double)
        L.or_scene_create.restype = C.c_void_p
        L.or_scene_create.argtypes = [C.c_void_p]
        L.or_scene_destroy.argtypes = [C.c_void_p]
        L.or_scene_triangles.restype = C.c_uint32
        L.or_scene_triangles.argtypes = [C.c_void_p]
        L.or_render.restype = C.c_int
        L.or_render.argtypes = [C.c_void_p, u32p, C.c_uint32, C.c_uint32, C.c_uint32, f64p, f64p, f32p, f32p, C.c_int, C.c_int]
        L.or_trace_rays.argtypes = [C.c_void_p, C.c_uint64, f32p, C.c_float, C.c_int, u32p]
        L.or_primary_hits.argtypes = [C.c_void_p, u32p, C.c_uint32, C.c_uint32, C.c_int, u32p, f32p]
        L.or_render_normals.argtypes = [C.c_void_p, u32p, C.c_int, f32p]
        L.or_render_aovs.argtypes = [C.c_void_p, u32p, C.c_int, f32p, u32p, u32p, f32p]
        for name in ("or_sin", "or_cos", "or_acos", "or_log2", "or_exp2"):
            getattr(L, name).restype = C.c_float
            getattr(L, name).argtypes = [C.c_float]
        for name in ("or_atan2", "or_pow", "or_G1", "or_D", "or_fresnel"):
            getattr(L, name).restype = C.c_float
            getattr(L, name).argtypes = [C.c_float, C.c_float]
        L.or_diff_prod.restype = C.c_float
        L.or_diff_prod.argtypes = [C.c_float] * 4
        L.or_sheen_core.restype = C.c_float
        L.or_sheen_core.argtypes = [C.c_float] * 4
        L.or_pcg.restype = C.c_uint32
        L.or_pcg.argtypes = [C.c_uint32]
        L.or_rng_next.restype = C.c_float
        L.or_rng_next.argtypes = [u32p]
        L.or_sample_vndf.argtypes = [f32p, C.c_float, C.c_float, C.c_float, f32p]
        L.or_intersect.restype = C.c_int
        L.or_intersect.argtypes = [f32p, f32p, f32p, C.c_float, f32p]
        L.or_pixel_active.restype = C.c_int
        L.or_pixel_active.argtypes = [f64p, C.c_uint32, C.c_float]
        L.or_gi_reconnect.restype = C.c_int
        L.or_gi_reconnect.argtypes = [f32p] * 6
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class OracleScene:
    """The oracle's copy of a packed scene: world triangles and a median-split BVH."""

    def __init__(self, packed):
        from verion_amd.abi import tables_struct

        self.packed = packed
        self._tables, self._keep = tables_struct(packed)
        self.h = lib().or_scene_create(C.byref(self._tables))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_scene_destroy(self.h)
            self.h = None

    @property
    def triangles(self) -> int:
        return lib().or_scene_triangles(self.h)

    def trace_rays(self, rays: np.ndarray, tmax: float = 1.0e30, brute: bool = False) -> np.ndarray:
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        out = np.zeros((len(rays), 4), np.uint32)
        lib().or_trace_rays(self.h, len(rays), _p(rays, C.c_float), tmax, int(brute), _p(out, C.c_uint32))
        return out

    def primary_hits(self, params: np.ndarray, sample: int, seed: int = 0, brute: bool = False, rays: bool = False):
        params = np.ascontiguousarray(params, np.uint32)
        n = int(params[0]) * int(params[1])
        out = np.zeros((n, 4), np.uint32)
        r = np.zeros((n, 6), np.float32) if rays else None
        lib().or_primary_hits(self.h, _p(params, C.c_uint32), seed, sample, int(brute), _p(out, C.c_uint32),
                              _p(r, C.c_float) if rays else None)
        return (out, r) if rays else out

    def normals(self, params: np.ndarray, brute: bool = False) -> np.ndarray:
        """normals.slang for the image size and camera of `params`: (pixels, 3) float32."""
        params = np.ascontiguousarray(params, np.uint32)
        n = int(params[0]) * int(params[1])
        out = np.zeros((n, 3), np.float32)
        lib().or_render_normals(self.h, _p(params, C.c_uint32), int(brute), _p(out, C.c_float))
        return out

    def aovs(self, params: np.ndarray, brute: bool = False) -> dict:
        """aov.slang: albedo (pixels, 3), depth, object_id, material_id, coverage (pixels,)."""
        params = np.ascontiguousarray(params, np.uint32)
        n = int(params[0]) * int(params[1])
        ad = np.zeros((n, 4), np.float32)
        oid = np.zeros(n, np.uint32)
        mid = np.zeros(n, np.uint32)
        cov = np.zeros(n, np.float32)
        lib().or_render_aovs(self.h, _p(params, C.c_uint32), int(brute), _p(ad, C.c_float), _p(oid, C.c_uint32),
                             _p(mid, C.c_uint32), _p(cov, C.c_float))
        return {"albedo": ad[:, :3].copy(), "depth": ad[:, 3].copy(), "object_id": oid, "material_id": mid, "coverage": cov}

    def render(self, params: np.ndarray, samples: int, first_sample: int = 0, seed: int = 0, threads: int = 0,
               brute: bool = False, state=None):
        """Accumulates `samples` samples into `state` (a dict of the running sums, created when None)
        and returns it: film (FP64, 3/pixel), stats (FP64, 3/pixel), lighting (65/pixel), restir (80/pixel)."""
        params = np.ascontiguousarray(params, np.uint32)
        n = int(params[0]) * int(params[1])
        if state is None:
            state = {"film": np.zeros(3 * n, np.float64),
                     "stats": np.zeros(3 * n, np.float64) if params[32] else None,
                     "lighting": np.zeros(65 * n, np.float32) if params[28] else None,
                     "restir": np.zeros(80 * n, np.float32) if params[36] else None}
        threads = threads or os.cpu_count() or 1
        g = lambda k, t: _p(state[k], t) if state[k] is not None else None
        rc = lib().or_render(self.h, _p(params, C.c_uint32), seed, first_sample, samples, g("film", C.c_double),
                             g("stats", C.c_double), g("lighting", C.c_float), g("restir", C.c_float), threads, int(brute))
        if rc:
            raise ValueError("oracle: unsupported parameter block (ReSTIR GI or a missing buffer)")
        return state


def mean(state, samples: int) -> np.ndarray:
    """readback::inverse_samples + mean (readback.rs:18-38) on the oracle's FP64 sums."""
    film = state["film"].reshape(-1, 3)
    if state.get("stats") is not None:
        inv = 1.0 / np.maximum(state["stats"].reshape(-1, 3)[:, 0], 1.0)
    else:
        inv = np.full(film.shape[0], 1.0 / max(samples, 1))
    return film * inv[:, None]


def gi_reconnect(p, n, pq, x2, n2):
    """ReSTIR GI's reconnection test (pt_core.slang:938-951): (accepted, Jacobian)."""
    arrs = [np.ascontiguousarray(v, np.float32) for v in (p, n, pq, x2, n2)]
    jac = np.zeros(1, np.float32)
    ok = lib().or_gi_reconnect(*[_p(a, C.c_float) for a in arrs], _p(jac, C.c_float))
    return bool(ok), float(jac[0])
