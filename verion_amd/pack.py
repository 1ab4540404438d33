"""Host-side scene packing: the flat word tables every kernel reads and the 43-word parameter block.

A restatement of crates/verion-gpu/src/pack.rs (the reference's single packer, shared by all its
backends): `pack_geometry` (:250-281), `pack_emitters` (:297-355), `pack_instances` (:358-381),
`pack_materials` (:395-473), `mip_chain` (:214-239), `pack_lights` (:582-695) with
`virtual_light` (:493-580), `build_env` (:144-211), `View` (:72-108), `base_params` (:828-874).
Float arithmetic that the reference does in f32 is done in numpy float32 element by element, and
sums the reference folds sequentially are folded sequentially here.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

from . import ies
from . import scene as sc

MAX_DEPTH = 8            # pack.rs:13
PARAM_WORDS = 43         # pack.rs:15
MATERIAL_WORDS = 24      # pack.rs:16
LIGHT_WORDS = 16         # pack.rs:17
ALBEDO_WORDS = 1024 + 256 + 4608  # pack.rs:29
EMITTER_WORDS = 16       # pack.rs:31
INSTANCE_WORDS = 16      # pack.rs:33
KIND_VIRTUAL = 3         # pack.rs:489

_TABLE_DIR = os.path.join(os.path.dirname(__file__), "tables")
_ALBEDO_FILES = ("ggx_albedo.f32", "sheen_albedo.f32", "dielectric_albedo.f32")

F32 = np.float32


def albedo_tables() -> np.ndarray:
    """The three directional-albedo tables as little-endian u32 words, in the order pt_core.slang's
    `albedoAt` expects (pack.rs:21-25; data from crates/verion-gpu/tables/)."""
    words = [np.fromfile(os.path.join(_TABLE_DIR, f), dtype="<u4") for f in _ALBEDO_FILES]
    out = np.concatenate(words).astype(np.uint32)
    assert out.size == ALBEDO_WORDS
    return out


def _f32bits(x) -> int:
    return int(np.array(x, dtype=np.float32).view(np.uint32))


def luminance(c) -> np.float32:
    c = np.asarray(c, dtype=np.float32)
    return F32(0.2126) * c[0] + F32(0.7152) * c[1] + F32(0.0722) * c[2]


@dataclass
class Env:
    pixels: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    marginal: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    conditional: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    width: int = 0
    height: int = 0
    pdf_scale: float = 0.0


@dataclass
class SceneInfo:
    light_count: int
    emitter_count: int
    env_width: int
    env_height: int
    env_pdf_scale: float
    sky: Tuple[float, float, float]
    virtual_light: bool
    transparent_background: bool


@dataclass
class View:
    camera: np.ndarray  # 16 float32
    aperture: float
    focus: float
    shift: Tuple[float, float]

    @staticmethod
    def of(cam: sc.Camera) -> "View":
        camera = np.zeros(16, np.float32)
        camera[2] = F32(math.tan(F32(cam.vertical_fov) * F32(0.5)))
        camera[4:16] = np.asarray(cam.transform, np.float32).reshape(12)
        return View(camera, float(F32(cam.aperture_radius)), float(F32(cam.focus_distance)),
                    (float(F32(cam.shift[0])), float(F32(cam.shift[1]))))


@dataclass
class Packed:
    """pack::Packed (pack.rs:110-133)."""

    positions: np.ndarray
    indices: np.ndarray
    uvs: np.ndarray
    vnormals: np.ndarray
    instances: np.ndarray
    materials: np.ndarray
    lights: np.ndarray
    emitters: np.ndarray
    emit_cdf: np.ndarray
    env: Env
    texels: np.ndarray
    tex_info: np.ndarray
    meshes: np.ndarray          # 4 words per mesh: vertex offset, vertex count, index offset, index count
    instance_mesh: np.ndarray   # mesh index per instance (the TLAS list)
    info: SceneInfo
    view: View

    @property
    def triangle_count(self) -> int:
        return int(sum(int(self.meshes[4 * m + 3]) // 3 for m in self.instance_mesh))


def _world(t: np.ndarray, p: np.ndarray) -> np.ndarray:
    """pack.rs:135-137 in f32: rows of t applied to points p (n, 3), left to right, no fusion."""
    out = np.empty_like(p)
    for r in range(3):
        out[:, r] = t[r, 0] * p[:, 0] + t[r, 1] * p[:, 1] + t[r, 2] * p[:, 2] + t[r, 3]
    return out


def pack_geometry(s: sc.Scene):
    positions, indices, uvs, vnormals, meshes = [], [], [], [], []
    vcount = icount = 0
    for m in s.meshes:
        meshes += [vcount, len(m.positions), icount, m.indices.size]
        positions.append(m.positions.reshape(-1))
        uvs.append(m.uvs.reshape(-1) if m.uvs is not None else np.zeros(2 * len(m.positions), np.float32))
        vnormals.append(m.normals.reshape(-1) if m.normals is not None else np.zeros(3 * len(m.positions), np.float32))
        indices.append(m.indices)
        vcount += len(m.positions)
        icount += m.indices.size
    cat = lambda xs, dt: np.concatenate(xs).astype(dt) if xs else np.zeros(0, dt)
    return (cat(positions, np.float32), cat(indices, np.uint32), cat(uvs, np.float32),
            cat(vnormals, np.float32), np.asarray(meshes, np.uint32))


def pack_emitters(s: sc.Scene):
    """pack.rs:297-355: world-space emissive triangles, selection pdf by power, CDF."""
    words: List[np.ndarray] = []
    powers: List[np.float32] = []
    bases = []
    for i, inst in enumerate(s.instances):
        base = s.materials[inst.material]
        le = np.asarray(base.emission, np.float32)
        radiance = le[0] + le[1] + le[2]
        if radiance <= 0.0 or base.emission_texture is not None:
            bases.append(0)
            continue
        mesh = s.meshes[inst.mesh]
        bases.append(len(powers) + 1)
        t = np.asarray(inst.transform, np.float32)
        tri = mesh.indices.reshape(-1, 3)
        a = _world(t, mesh.positions[tri[:, 0]])
        b = _world(t, mesh.positions[tri[:, 1]])
        c = _world(t, mesh.positions[tri[:, 2]])
        e1 = b - a
        e2 = c - a
        cr = np.stack([e1[:, 1] * e2[:, 2] - e1[:, 2] * e2[:, 1],
                       e1[:, 2] * e2[:, 0] - e1[:, 0] * e2[:, 2],
                       e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]], axis=1)
        area = F32(0.5) * np.sqrt(cr[:, 0] * cr[:, 0] + cr[:, 1] * cr[:, 1] + cr[:, 2] * cr[:, 2])
        rec = np.zeros((len(tri), EMITTER_WORDS), np.float32)
        rec[:, 0:3] = a
        rec[:, 3:6] = b
        rec[:, 6:9] = c
        rec[:, 9:12] = le
        rec[:, 12] = area
        rec[:, 14] = np.array([s.instance_emit_mask(i)], np.uint32).view(np.float32)[0]
        words.append(rec)
        powers.extend(list(area * radiance))
    if not words:
        return np.zeros(0, np.float32), np.zeros(0, np.float32), np.asarray(bases, np.uint32)
    rec = np.concatenate(words)
    total = F32(0.0)
    for p in powers:  # sequential f32 fold, as Iterator::sum
        total = F32(total + p)
    cdf = np.empty(len(powers), np.float32)
    running = F32(0.0)
    for k, p in enumerate(powers):
        sel = F32(p / total)
        rec[k, 13] = sel
        running = F32(running + sel)
        cdf[k] = running
    cdf[-1] = 1.0
    return rec.reshape(-1), cdf, np.asarray(bases, np.uint32)


def pack_instances(s: sc.Scene, meshes: np.ndarray, emitter_bases: np.ndarray):
    table = np.zeros((len(s.instances), INSTANCE_WORDS), np.uint32)
    inst_mesh = np.zeros(len(s.instances), np.uint32)
    for k, inst in enumerate(s.instances):
        table[k, 0] = meshes[4 * inst.mesh]
        table[k, 1] = meshes[4 * inst.mesh + 2]
        table[k, 2] = inst.material
        table[k, 3] = emitter_bases[k]
        table[k, 4:16] = np.asarray(inst.transform, np.float32).reshape(12).view(np.uint32)
        inst_mesh[k] = inst.mesh
    return table.reshape(-1), inst_mesh


def mip_chain(image: sc.Image):
    """pack.rs:214-239: box-filtered levels down to 1x1, level 0 first."""
    w, h = image.width, image.height
    level = image.pixels.reshape(h, w, 3).astype(np.float32)
    out = []
    levels = 0
    while True:
        out.append(level.reshape(-1))
        levels += 1
        if w == 1 and h == 1:
            return np.concatenate(out), levels
        nw, nh = max(w // 2, 1), max(h // 2, 1)
        nxt = np.zeros((nh, nw, 3), np.float32)
        for dx, dy in ((0, 0), (1, 0), (0, 1), (1, 1)):
            ys = np.minimum(2 * np.arange(nh) + dy, h - 1)
            xs = np.minimum(2 * np.arange(nw) + dx, w - 1)
            nxt = nxt + level[ys][:, xs]
        level = nxt / F32(4.0)
        w, h = nw, nh


def pack_materials(s: sc.Scene):
    tex_of = {}
    texels: List[np.ndarray] = []
    tex_info: List[int] = []
    ntexels = 0

    def texture(img: Optional[int]) -> float:
        nonlocal ntexels
        if img is None:
            return 0.0
        if img not in tex_of:
            if not (0 <= img < len(s.images)):
                raise ValueError("material refers to a missing image")
            chain, levels = mip_chain(s.images[img])
            tex_info.extend([ntexels, s.images[img].width, s.images[img].height, levels])
            texels.append(chain)
            ntexels += chain.size // 3
            tex_of[img] = len(tex_of)
        return float(tex_of[img] + 1)

    mats = []
    for b in s.materials:
        t0, t1, t2 = texture(b.base_color_texture), texture(b.metallic_roughness_texture), texture(b.emission_texture)
        tn = texture(b.normal_texture)
        to = texture(b.opacity_texture)
        mats.append([*b.base_color, *b.emission, b.metallic, b.roughness, b.transmission, b.ior, t0, t1, t2, tn,
                     b.normal_scale, b.clearcoat, b.clearcoat_roughness, b.sheen_roughness, *b.sheen_color,
                     b.dispersion, b.opacity, to])
    materials = np.asarray(mats, np.float32).reshape(-1)
    tx = np.concatenate(texels).astype(np.float32) if texels else np.zeros(0, np.float32)
    return materials, tx, np.asarray(tex_info, np.uint32)


def _env_image(s: sc.Scene):
    maps = [l for l in s.lights if isinstance(l, sc.EnvironmentMap)]
    if len(maps) > 1:
        raise ValueError("only one environment map is supported")
    if not maps:
        return None
    if not (0 <= maps[0].image < len(s.images)):
        raise ValueError("environment map refers to a missing image")
    return s.images[maps[0].image], maps[0].scale


def build_env(s: sc.Scene) -> Env:
    """pack.rs:144-211: CDFs over luminance * sin(theta), built in f64 and stored f32."""
    found = _env_image(s)
    if found is None:
        return Env()
    image, scale = found
    w, h = image.width, image.height
    pixels = (image.pixels * F32(scale)).astype(np.float32)
    lumv = luminance(pixels.T).astype(np.float64).reshape(h, w)
    rows = np.zeros(h, np.float64)
    cond = np.zeros(w * h, np.float32)
    for j in range(h):
        sin_t = math.sin((j + 0.5) / h * math.pi)
        wts = lumv[j] * sin_t
        rows[j] = sum(wts.tolist())  # sequential f64 fold, as Iterator::sum
        running = np.cumsum(wts)
        if rows[j] > 0.0:
            cond[j * w:(j + 1) * w] = (running / rows[j]).astype(np.float32)
        else:
            cond[j * w:(j + 1) * w] = (np.arange(1, w + 1) / w).astype(np.float32)
        cond[j * w + w - 1] = 1.0
    total = sum(rows.tolist())
    if total <= 0.0:
        return Env()
    marginal = np.cumsum(rows / total).astype(np.float32)
    marginal[h - 1] = 1.0
    pdf_scale = float(F32((w * h) / (2.0 * math.pi ** 2 * total)))
    return Env(pixels.reshape(-1), marginal, cond, w, h, pdf_scale)


def virtual_light(s: sc.Scene):
    """pack.rs:493-580 in automatic mode: the brightest 1% of the map as a directional light, or None."""
    if not s.virtual_light:
        return None
    found = _env_image(s)
    if found is None:
        return None
    image, scale = found
    w, h = image.width, image.height
    rgb = (image.pixels * F32(scale)).astype(np.float32).reshape(h, w, 3)
    j = np.arange(h, dtype=np.float32)[:, None]
    i = np.arange(w, dtype=np.float32)[None, :]
    theta = (j + F32(0.5)) / F32(h) * F32(math.pi)
    phi = ((i + F32(0.5)) / F32(w) - F32(0.5)) * F32(2.0) * F32(math.pi)
    omega = np.broadcast_to(F32(2.0) * F32(math.pi) * F32(math.pi) / F32(w * h) * np.sin(theta), (h, w)).astype(np.float32)
    dirs = np.stack(np.broadcast_arrays(np.sin(theta) * np.sin(phi), np.cos(theta), -np.sin(theta) * np.cos(phi)), -1)
    lums = luminance(np.moveaxis(rgb, -1, 0)).astype(np.float32)
    total = float(np.sum(lums.astype(np.float64) * omega.astype(np.float64)))
    area = float(np.sum(omega.astype(np.float64)))
    mean = total / area
    flat = np.sort(lums.reshape(-1))[::-1]
    cut = max(flat.size // 100, 1)
    threshold = flat[cut - 1]
    sel = (lums >= threshold) & (lums.astype(np.float64) > mean)
    e = (lums * omega).astype(np.float64)
    energy = float(np.sum(e[sel]))
    if energy <= 0.0:
        return None
    solid = float(np.sum(omega[sel].astype(np.float64)))
    d = np.sum(e[sel][:, None] * dirs[sel].astype(np.float64), axis=0)
    length = float(np.sqrt(np.sum(d * d))) / energy
    cluster_mean = energy / solid
    if not (length >= 0.5 and cluster_mean >= 1.5 * mean):
        return None
    toward = (d / (length * energy)).astype(np.float32)
    cos = np.maximum(np.sum(dirs * toward, axis=-1), 0.0).astype(np.float64)
    irr = np.sum((rgb * omega[..., None]).astype(np.float64) * cos[..., None], axis=(0, 1))
    return toward, irr.astype(np.float32)


def pack_lights(s: sc.Scene):
    """pack.rs:582-695: delta lights, ranked slots, the virtual main light, albedo tables, link masks,
    then the IES profiles of point and spot lights (their frame and table; word 15 = offset + 1)."""
    sky = np.zeros(3, np.float32)
    lights: List[List[int]] = []
    strengths = []
    profiles = []  # (row in the light table, the light) of every light with an IES profile
    for light in s.lights:
        cone = [0.0, 0.0, 0.0, 0.0]
        offset = 0.0
        if isinstance(light, sc.PointLight):
            kind, color, intensity, vector = 0, light.color, light.intensity, light.position
        elif isinstance(light, sc.DirectionalLight):
            kind, color, intensity, vector = 1, light.color, light.intensity, light.direction
        elif isinstance(light, sc.SpotLight):
            d = np.asarray(light.direction, np.float32)
            ln = np.sqrt(np.sum(d * d))
            axis = d / ln
            scale = F32(1.0) / max(F32(math.cos(light.inner_cone)) - F32(math.cos(light.outer_cone)), F32(1.0e-3))
            cone = [axis[0], axis[1], axis[2], scale]
            offset = -F32(math.cos(light.outer_cone)) * scale
            kind, color, intensity, vector = 2, light.color, light.intensity, light.position
        elif isinstance(light, sc.EnvironmentMap):
            continue
        elif isinstance(light, sc.Environment):
            sky = (sky + np.asarray(light.color, np.float32)).astype(np.float32)
            continue
        else:
            raise TypeError(f"unknown light {light!r}")
        words = [0] * LIGHT_WORDS
        words[0] = kind
        words[1:4] = [_f32bits(v) for v in color]
        words[4] = _f32bits(intensity)
        words[5:8] = [_f32bits(v) for v in vector]
        words[8:12] = [_f32bits(v) for v in cone]
        words[12] = _f32bits(offset)
        words[14] = light.link_mask & 0xFFFFFFFF
        if getattr(light, "ies", None) is not None and kind != 1:
            profiles.append((len(lights), light))
        lights.append(words)
        strengths.append(float(luminance(color) * F32(intensity)))
    ranked = sorted(range(len(strengths)), key=lambda k: -strengths[k])  # stable, strongest first
    for rank, k in enumerate(ranked):
        lights[k][13] = min(rank, 2)
    vl = virtual_light(s)
    if vl is not None:
        toward, e = vl
        words = [0] * LIGHT_WORDS
        words[0] = KIND_VIRTUAL
        words[1:4] = [_f32bits(v) for v in e]
        words[4] = _f32bits(1.0)
        words[5:8] = [_f32bits(-v) for v in toward]
        words[13] = 3
        words[14] = 0xFFFFFFFF
        lights.append(words)
    count = len(lights)
    table = [np.asarray(lights, np.uint32).reshape(-1), albedo_tables()]
    masks = np.zeros(2 * len(s.instances), np.uint32)
    for i in range(len(s.instances)):
        masks[2 * i] = s.instance_light_mask(i)
        masks[2 * i + 1] = s.instance_emit_mask(i)
    table.append(masks)
    out = np.concatenate(table).astype(np.uint32)
    if profiles:
        parts = [out]
        size = out.size
        for row, light in profiles:
            profile, frame = light.ies
            frame = np.asarray(frame, np.float32).reshape(9)
            data = np.asarray(profile.table, np.float32).reshape(-1)
            if data.size != ies.HORIZONTAL * ies.VERTICAL:
                raise ValueError(f"an IES table holds {ies.HORIZONTAL} x {ies.VERTICAL} values, not {data.size}")
            out[row * LIGHT_WORDS + 15] = size + 1
            parts += [frame.view(np.uint32), data.view(np.uint32)]
            size += frame.size + data.size
        out = np.concatenate(parts).astype(np.uint32)
    return out, count, sky, vl is not None


def pack(s: sc.Scene) -> Packed:
    """pack.rs:726-756."""
    if not s.instances:
        raise ValueError("scene has no instances")
    positions, indices, uvs, vnormals, meshes = pack_geometry(s)
    emitters, cdf, bases = pack_emitters(s)
    instances, inst_mesh = pack_instances(s, meshes, bases)
    materials, texels, tex_info = pack_materials(s)
    lights, count, sky, virt = pack_lights(s)
    env = build_env(s)
    info = SceneInfo(count, int(cdf.size), env.width, env.height, env.pdf_scale,
                     (float(sky[0]), float(sky[1]), float(sky[2])), virt, bool(s.transparent_background))
    return Packed(positions, indices, uvs, vnormals, instances, materials, lights, emitters, cdf, env, texels,
                  tex_info, meshes, inst_mesh, info, View.of(s.camera))


@dataclass
class UpdatePlan:
    """What a batch of scene changes makes a backend redo (pack.rs:758-825): each flag is a table to
    re-upload or a structure to redo; everything else is reused.

    | change | reused | redone |
    |---|---|---|
    | camera | everything | the view |
    | material | LBVH, vertex tables | material and texture tables, emitters, instance table |
    | light | LBVH, tables | light table, sky, environment map |
    | instance transform | vertex tables | instance table, LBVH refit, emitters |
    | mesh deformation | other tables | vertex tables, LBVH refit, emitters |
    | geometry | nothing | everything |
    """

    full: bool = False
    materials: bool = False
    lights: bool = False
    emitters: bool = False
    instances: bool = False
    vertices: bool = False
    blases: List[int] = field(default_factory=list)
    tlas: bool = False
    camera: bool = False

    @staticmethod
    def new(changes: sc.Changes) -> "UpdatePlan":
        """UpdatePlan::new (pack.rs:785-815)."""
        p = UpdatePlan(camera=changes.camera, lights=changes.lights)
        if changes.geometry:
            return UpdatePlan(full=True, materials=True, lights=True, emitters=True, instances=True, vertices=True,
                              tlas=True, camera=changes.camera)
        if changes.materials:
            p.materials = True
            p.emitters = True
        if changes.instances:
            p.instances = True
            p.tlas = True
            p.emitters = True
        if changes.deformed:
            p.vertices = True
            p.blases = list(changes.deformed)
            p.tlas = True
            p.emitters = True
        p.instances |= p.emitters  # emitters carry the instance table's emitter bases
        return p

    def is_empty(self) -> bool:
        return self == UpdatePlan()


def base_params(info: SceneInfo, view: View, width: int, height: int, lighting: bool, multiscatter: bool,
                adaptive=None, restir=None) -> np.ndarray:
    """pack.rs:828-874: the parameter block with the per-dispatch words (6, 30, 31, 35) zero.
    `adaptive` is (threshold, min_spp) or None; `restir` is (di, gi, candidates, history[, neighbours])
    or None. Neighbours (random neighbours reused per frame, 0 = the reference's fixed 2,
    pt_core.slang:706) go to bits 8-15 of word 36, an extension of this build."""
    p = np.zeros(PARAM_WORDS, np.uint32)
    p[0] = width
    p[1] = height
    p[2] = _f32bits(view.camera[2])
    p[3] = _f32bits(view.aperture)
    p[4] = _f32bits(view.focus)
    p[5] = info.light_count
    p[7] = MAX_DEPTH
    p[8:20] = np.asarray(view.camera[4:16], np.float32).view(np.uint32)
    p[20:23] = np.asarray(info.sky, np.float32).view(np.uint32)
    p[23] = info.emitter_count
    p[24] = info.env_width
    p[25] = info.env_height
    p[26] = _f32bits(info.env_pdf_scale)
    p[27] = _f32bits(F32(2.0) * F32(view.camera[2]) / F32(height))
    p[41] = int(info.virtual_light)
    p[42] = int(info.transparent_background)
    p[39] = _f32bits(view.shift[0])
    p[40] = _f32bits(view.shift[1])
    p[28] = int(bool(lighting))
    p[29] = int(bool(multiscatter))
    if adaptive is not None:
        p[32] = 1
        p[33] = _f32bits(adaptive[0])
        p[34] = int(adaptive[1])
    if restir is not None:
        p[36] = int(bool(restir[0])) | (int(bool(restir[1])) << 1)
        if p[36] and len(restir) > 4:
            p[36] |= int(restir[4]) << 8
        p[37] = int(restir[2])
        p[38] = int(restir[3])
    return p
