"""ctypes mirror of the structs in include/verion_hip.h, and the conversion of `Packed` tables into them."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .pack import Packed, View

u32p = C.POINTER(C.c_uint32)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)


class SceneInfo(C.Structure):
    _fields_ = [("light_count", C.c_uint32), ("emitter_count", C.c_uint32), ("env_width", C.c_uint32),
                ("env_height", C.c_uint32), ("env_pdf_scale", C.c_float), ("sky", C.c_float * 3),
                ("virtual_light", C.c_uint32), ("transparent_background", C.c_uint32)]


class ViewS(C.Structure):
    _fields_ = [("camera", C.c_float * 16), ("aperture", C.c_float), ("focus", C.c_float), ("shift", C.c_float * 2)]


_TABLES = [("positions", f32p), ("indices", u32p), ("uvs", f32p), ("vnormals", f32p), ("instances", u32p),
           ("materials", f32p), ("lights", u32p), ("emitters", f32p), ("emit_cdf", f32p), ("env_pixels", f32p),
           ("env_marginal", f32p), ("env_cond", f32p), ("texels", f32p), ("tex_info", u32p)]


class SceneTables(C.Structure):
    _fields_ = [f for name, t in _TABLES for f in ((name, t), (name + "_len", C.c_uint64))] + [
        ("meshes", u32p), ("mesh_count", C.c_uint64), ("instance_mesh", u32p), ("instance_count", C.c_uint64),
        ("info", SceneInfo), ("view", ViewS)]


class Adaptive(C.Structure):
    _fields_ = [("threshold", C.c_float), ("min_spp", C.c_uint32)]


class Restir(C.Structure):
    _fields_ = [("di", C.c_uint32), ("gi", C.c_uint32), ("candidates", C.c_uint32), ("history", C.c_uint32),
                ("neighbours", C.c_uint32)]


class NrcConfig(C.Structure):
    _fields_ = [("train_paths", C.c_uint32), ("train_steps", C.c_uint32), ("learning_rate", C.c_float),
                ("terminate_depth", C.c_uint32), ("ema_decay", C.c_float), ("warmup_frames", C.c_uint32),
                ("minibatch", C.c_uint32)]


class BatchJob(C.Structure):
    _fields_ = [("renderer", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32),
                ("adaptive", Adaptive), ("seed", C.c_uint32)]


class BatchTuning(C.Structure):
    _fields_ = [("grid_blocks", C.c_uint32), ("tile_width", C.c_uint32), ("tile_height", C.c_uint32)]


class MemoryUsage(C.Structure):
    _fields_ = [("budget", C.c_uint64), ("allocated", C.c_uint64), ("uploaded", C.c_uint64)]


class UpdateReport(C.Structure):
    _fields_ = [("blases_built", C.c_uint64), ("tlas_built", C.c_uint32), ("bytes_uploaded", C.c_uint64),
                ("bvh_work", C.c_uint32)]


class UpdatePlanS(C.Structure):
    _fields_ = [("full", C.c_uint32), ("materials", C.c_uint32), ("lights", C.c_uint32), ("emitters", C.c_uint32),
                ("instances", C.c_uint32), ("vertices", C.c_uint32), ("tlas", C.c_uint32), ("camera", C.c_uint32),
                ("blases", C.POINTER(C.c_uint32)), ("blas_count", C.c_uint64)]


def ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def view_struct(v: View) -> ViewS:
    s = ViewS()
    for k in range(16):
        s.camera[k] = float(v.camera[k])
    s.aperture = v.aperture
    s.focus = v.focus
    s.shift[0] = v.shift[0]
    s.shift[1] = v.shift[1]
    return s


def tables_struct(p: Packed):
    """Returns (SceneTables, keepalive list). The arrays must outlive every call that reads the struct."""
    keep = []
    t = SceneTables()
    arrays = {
        "positions": p.positions, "indices": p.indices, "uvs": p.uvs, "vnormals": p.vnormals,
        "instances": p.instances, "materials": p.materials, "lights": p.lights, "emitters": p.emitters,
        "emit_cdf": p.emit_cdf, "env_pixels": p.env.pixels, "env_marginal": p.env.marginal,
        "env_cond": p.env.conditional, "texels": p.texels, "tex_info": p.tex_info,
    }
    for name, ctype in _TABLES:
        dt = np.float32 if ctype is f32p else np.uint32
        a = np.ascontiguousarray(arrays[name], dtype=dt)
        if a.size == 0:
            a = np.zeros(1, dt)  # a valid pointer; the length says 0
            n = 0
        else:
            n = a.size
        keep.append(a)
        setattr(t, name, ptr(a, C.c_float if dt == np.float32 else C.c_uint32))
        setattr(t, name + "_len", n)
    meshes = np.ascontiguousarray(p.meshes, np.uint32)
    im = np.ascontiguousarray(p.instance_mesh, np.uint32)
    keep += [meshes, im]
    t.meshes = ptr(meshes, C.c_uint32)
    t.mesh_count = meshes.size // 4
    t.instance_mesh = ptr(im, C.c_uint32)
    t.instance_count = im.size
    i = p.info
    t.info.light_count = i.light_count
    t.info.emitter_count = i.emitter_count
    t.info.env_width = i.env_width
    t.info.env_height = i.env_height
    t.info.env_pdf_scale = i.env_pdf_scale
    for k in range(3):
        t.info.sky[k] = i.sky[k]
    t.info.virtual_light = int(i.virtual_light)
    t.info.transparent_background = int(i.transparent_background)
    t.view = view_struct(p.view)
    return t, keep
