"""Seeded generators of the benchmark scenes, plus the small closed-form scenes of the reference tests.

None of these scenes exists in the reference; they are specified by the workload list this backend
is measured on:
  cornell(): a Cornell box in a 1 m cube, 5 walls + 2 boxes + a 0.25 m^2 quad light of radiance
             (17, 12, 4), Lambert red (.63, .065, .05), green (.14, .45, .09), white .73 (~36 triangles);
  procedural(): 1,024 level-3 icospheres (1,280 triangles each, 1,310,720 in all) at uniform
             positions in a 20 m cube, lognormal radii (mu 0, sigma 0.5), GGX roughness U[0.05, 1],
             metallic Bernoulli(0.3), one sun and a constant environment map (importance sampled);
  many_lights(): ground and box clutter with emissive triangles of lognormal emission.
The closed-form scenes restate crates/verion-gpu-metal/tests/pathtrace.rs: `cube` (:38-58),
`plane` (:61-70), `furnace_room` (:87-98), the white furnace (:118-137); `closed_room` restates the
ReSTIR GI test room of crates/verion-gpu-vulkan/tests/restir.rs:152-231.
"""
from __future__ import annotations

import math
from typing import Dict, Tuple

import numpy as np

from . import scene as sc


def cube_mesh(half: float) -> sc.Mesh:
    """Axis-aligned cube of 12 triangles (tests/pathtrace.rs:38-58)."""
    h = half
    pos = [[-h, -h, -h], [h, -h, -h], [h, h, -h], [-h, h, -h], [-h, -h, h], [h, -h, h], [h, h, h], [-h, h, h]]
    idx = [0, 1, 2, 0, 2, 3, 4, 5, 6, 4, 6, 7, 0, 1, 5, 0, 5, 4, 2, 3, 7, 2, 7, 6, 1, 2, 6, 1, 6, 5, 0, 3, 7, 0, 7, 4]
    return sc.Mesh(np.array(pos, np.float32), np.array(idx, np.uint32))


def plane_mesh(half: float, z: float) -> sc.Mesh:
    """Square in the plane z (tests/pathtrace.rs:61-70)."""
    h = half
    return sc.Mesh(np.array([[-h, -h, z], [h, -h, z], [h, h, z], [-h, h, z]], np.float32),
                   np.array([0, 1, 2, 0, 2, 3], np.uint32))


def camera_at_z(z: float) -> sc.Camera:
    return sc.Camera(transform=((1.0, 0.0, 0.0, 0.0), (0.0, 1.0, 0.0, 0.0), (0.0, 0.0, 1.0, z)))


def furnace_room(rho: float, e: float) -> sc.Scene:
    """Camera inside a closed cube whose walls emit e and reflect rho (tests/pathtrace.rs:87-98)."""
    s = sc.Scene()
    m = s.add_mesh(cube_mesh(1.0))
    mat = s.add_material(sc.lambert((rho, rho, rho), (e, e, e)))
    s.add_instance(sc.Instance(m, mat))
    return s


def white_furnace(material: sc.BaseBsdf = None, sky: float = 1.0) -> sc.Scene:
    """Convex object under a uniform sky (tests/pathtrace.rs:118-137); albedo-1 Lambert by default."""
    s = sc.Scene()
    m = s.add_mesh(cube_mesh(0.5))
    mat = s.add_material(material if material is not None else sc.lambert((1.0, 1.0, 1.0)))
    s.add_instance(sc.Instance(m, mat))
    s.add_light(sc.Environment((sky, sky, sky)))
    s.camera = camera_at_z(3.0)
    return s


def floor_scene(light, occluder: bool = False) -> sc.Scene:
    """Albedo-1 Lambert floor lit by one light, seen from z = 3 (tests/pathtrace.rs:232-263); the
    occluder is a small black triangle between the light and the floor point under the camera."""
    s = sc.Scene()
    floor = s.add_mesh(plane_mesh(5.0, 0.0))
    s.add_instance(sc.Instance(floor, s.add_material(sc.lambert((1.0, 1.0, 1.0)))))
    if occluder:
        tri = s.add_mesh(sc.Mesh(np.array([[0.4, -0.3, 0.5], [0.6, -0.3, 0.5], [0.5, 0.3, 0.5]], np.float32),
                                 np.array([0, 1, 2], np.uint32)))
        s.add_instance(sc.Instance(tri, s.add_material(sc.lambert((0.0, 0.0, 0.0)))))
    s.add_light(light)
    s.camera = camera_at_z(3.0)
    return s


def sky_scene(mesh: sc.Mesh, material: sc.BaseBsdf, camera_z: float = 3.0, sky: float = 1.0) -> sc.Scene:
    """One object under a uniform sky, seen from the +z axis (tests/pathtrace.rs:543-560)."""
    s = sc.Scene()
    m = s.add_mesh(mesh)
    s.add_instance(sc.Instance(m, s.add_material(material)))
    s.add_light(sc.Environment((sky, sky, sky)))
    s.camera = camera_at_z(camera_z)
    return s


def closed_room(roughness: float, glass: bool = False, dispersion: float = 8.0) -> sc.Scene:
    """A closed room of rough walls (a 3 m cube) lit by 16 small ceiling panels of radiance 30, with
    a rough 0.8 m block in the middle, seen from inside at z = 1.4: most of what the camera sees is
    indirect light (crates/verion-gpu-vulkan/tests/restir.rs:152-231). `glass` adds a strongly
    dispersive glass block (ior 1.6, dispersion 8 unless given) of half size 0.35 at (-0.7, -0.9, 0)."""
    s = sc.Scene()
    room = s.add_mesh(cube_mesh(1.5))
    ball = s.add_mesh(cube_mesh(0.4))
    pos, idx = [], []
    for k in range(16):
        x, z, h = -0.6 + (k % 4) * 0.4, -0.6 + (k // 4) * 0.4, 0.06
        b = len(pos)
        pos += [[x - h, 1.45, z - h], [x + h, 1.45, z - h], [x + h, 1.45, z + h], [x - h, 1.45, z + h]]
        idx += [b, b + 1, b + 2, b, b + 2, b + 3]
    lamps = s.add_mesh(sc.Mesh(np.array(pos, np.float32), np.array(idx, np.uint32)))
    walls = s.add_material(sc.BaseBsdf(base_color=(0.7, 0.65, 0.6), roughness=roughness, ior=1.0))
    glow = s.add_material(sc.BaseBsdf(base_color=(0.0, 0.0, 0.0), emission=(30.0, 30.0, 30.0), roughness=roughness))
    s.add_instance(sc.Instance(room, walls))
    s.add_instance(sc.Instance(ball, walls))
    s.add_instance(sc.Instance(lamps, glow))
    if glass:
        block = s.add_mesh(cube_mesh(0.35))
        dispersive = s.add_material(sc.BaseBsdf(base_color=(1.0, 1.0, 1.0), transmission=1.0, roughness=0.0,
                                                ior=1.6, dispersion=dispersion))
        s.add_instance(sc.Instance(block, dispersive, ((1.0, 0.0, 0.0, -0.7), (0.0, 1.0, 0.0, -0.9), (0.0, 0.0, 1.0, 0.0))))
    s.camera = camera_at_z(1.4)
    return s


def _quad(p0, p1, p2, p3):
    return [p0, p1, p2, p3], [0, 1, 2, 0, 2, 3]


def _box(cx: float, cz: float, w: float, h: float, angle: float):
    """Closed box of 12 triangles standing on y = 0, rotated about y."""
    c, s_ = math.cos(angle), math.sin(angle)
    corners = []
    for y in (0.0, h):
        for dx, dz in ((-1, -1), (1, -1), (1, 1), (-1, 1)):
            x, z = dx * w / 2, dz * w / 2
            corners.append([cx + c * x + s_ * z, y, cz - s_ * x + c * z])
    idx = [0, 2, 1, 0, 3, 2, 4, 5, 6, 4, 6, 7, 0, 1, 5, 0, 5, 4, 1, 2, 6, 1, 6, 5, 2, 3, 7, 2, 7, 6, 3, 0, 4, 3, 4, 7]
    return corners, idx


def cornell() -> sc.Scene:
    """The Cornell box workload: 36 triangles in a 1 m cube, camera on the open side."""
    s = sc.Scene()
    white = s.add_material(sc.lambert((0.73, 0.73, 0.73)))
    red = s.add_material(sc.lambert((0.63, 0.065, 0.05)))
    green = s.add_material(sc.lambert((0.14, 0.45, 0.09)))
    light = s.add_material(sc.lambert((0.0, 0.0, 0.0), (17.0, 12.0, 4.0)))

    def add(verts, idx, mat):
        m = s.add_mesh(sc.Mesh(np.array(verts, np.float32), np.array(idx, np.uint32)))
        s.add_instance(sc.Instance(m, mat))

    add(*_quad([0, 0, 0], [1, 0, 0], [1, 0, 1], [0, 0, 1]), white)      # floor
    add(*_quad([0, 1, 0], [0, 1, 1], [1, 1, 1], [1, 1, 0]), white)      # ceiling
    add(*_quad([0, 0, 0], [0, 1, 0], [1, 1, 0], [1, 0, 0]), white)      # back wall
    add(*_quad([0, 0, 0], [0, 0, 1], [0, 1, 1], [0, 1, 0]), red)        # left wall
    add(*_quad([1, 0, 0], [1, 1, 0], [1, 1, 1], [1, 0, 1]), green)      # right wall
    add(*_box(0.68, 0.62, 0.3, 0.3, -0.29), white)                       # short box
    add(*_box(0.32, 0.35, 0.3, 0.6, 0.29), white)                        # tall box
    add(*_quad([0.25, 0.999, 0.25], [0.25, 0.999, 0.75], [0.75, 0.999, 0.75], [0.75, 0.999, 0.25]), light)
    s.camera = sc.Camera(transform=((1.0, 0.0, 0.0, 0.5), (0.0, 1.0, 0.0, 0.5), (0.0, 0.0, 1.0, 2.2)))
    return s


def _textured_quad(cx, cz, half, y=0.0):
    """Quad at height y with uvs 0..2 (repeat wrap) and +y vertex normals."""
    v = np.array([[cx - half, y, cz - half], [cx + half, y, cz - half], [cx + half, y, cz + half], [cx - half, y, cz + half]], np.float32)
    uv = np.array([[0, 0], [2, 0], [2, 2], [0, 2]], np.float32)
    return sc.Mesh(v, np.array([0, 2, 1, 0, 3, 2], np.uint32), uvs=uv, normals=np.tile([[0, 1, 0]], (4, 1)).astype(np.float32))


def zoo(seed: int = 7) -> sc.Scene:
    """A material and light test scene: every BSDF feature of pt_core.slang (metal, rough and smooth
    dielectric, dispersion, clear coat, sheen, translucent opacity with an opacity texture, base /
    metallic-roughness / emission / normal textures), a point, a spot and a directional light,
    emissive triangles and an importance-sampled environment map."""
    rng = np.random.default_rng(seed)
    s = sc.Scene()
    n = 16
    checker = np.array([[0.9, 0.2, 0.2] if (i // 4 + j // 4) % 2 else [0.2, 0.8, 0.3] for j in range(n) for i in range(n)], np.float32)
    base_tex = s.add_image(sc.Image(n, n, checker))
    mr = s.add_image(sc.Image(n, n, np.stack([np.zeros(n * n), rng.uniform(0.2, 1.0, n * n), rng.uniform(0, 1, n * n)], 1)))
    em = s.add_image(sc.Image(n, n, rng.uniform(0.0, 2.0, (n * n, 3))))
    nrm = s.add_image(sc.Image(n, n, np.clip(np.stack([0.5 + 0.2 * rng.normal(size=n * n), 0.5 + 0.2 * rng.normal(size=n * n), np.ones(n * n)], 1), 0, 1)))
    opa = s.add_image(sc.Image(n, n, np.repeat(rng.uniform(0.0, 1.0, (n * n, 1)), 3, 1)))
    env = s.add_image(sc.Image(16, 8, rng.uniform(0.1, 1.5, (128, 3))))
    floor = s.add_mesh(_textured_quad(0.0, 0.0, 4.0, 0.0))
    s.add_instance(sc.Instance(floor, s.add_material(sc.BaseBsdf(base_color=(0.8, 0.8, 0.8), base_color_texture=base_tex,
                                                                  normal_texture=nrm, normal_scale=0.7))))
    mats = [
        sc.BaseBsdf(base_color=(0.95, 0.7, 0.3), metallic=1.0, roughness=0.35, metallic_roughness_texture=mr),
        sc.BaseBsdf(base_color=(1.0, 1.0, 1.0), transmission=1.0, roughness=0.0, ior=1.5),
        sc.BaseBsdf(base_color=(0.9, 1.0, 0.95), transmission=1.0, roughness=0.4, ior=1.7, dispersion=1.0),
        sc.BaseBsdf(base_color=(0.2, 0.3, 0.8), roughness=0.6, clearcoat=1.0, clearcoat_roughness=0.1),
        sc.BaseBsdf(base_color=(0.5, 0.1, 0.1), roughness=0.8, sheen_color=(0.9, 0.8, 0.7), sheen_roughness=0.4),
        sc.BaseBsdf(base_color=(0.7, 0.7, 0.2), roughness=0.5, opacity=0.8, opacity_texture=opa),
        sc.BaseBsdf(base_color=(0.1, 0.1, 0.1), emission=(1.0, 1.0, 1.0), emission_texture=em),
        sc.BaseBsdf(base_color=(0.0, 0.0, 0.0), emission=(6.0, 5.0, 4.0), ior=1.0),
    ]
    ico = icosphere(2)
    sphere = s.add_mesh(sc.Mesh(ico.positions, ico.indices, uvs=(ico.positions[:, :2] * 0.5 + 0.5), normals=ico.normals))
    box = s.add_mesh(cube_mesh(0.5))
    for k, m in enumerate(mats):
        mi = s.add_material(m)
        x = -3.0 + 6.0 * (k % 4) / 3.0
        z = -1.2 if k < 4 else 1.2
        mesh = sphere if k % 2 == 0 else box
        r = 0.6
        s.add_instance(sc.Instance(mesh, mi, ((r, 0.0, 0.0, x), (0.0, r, 0.0, 0.7), (0.0, 0.0, r, z))))
    s.add_light(sc.PointLight(position=(0.0, 3.0, 2.0), color=(1.0, 0.9, 0.8), intensity=6.0))
    s.add_light(sc.SpotLight(position=(2.0, 3.0, 0.0), direction=(-0.5, -1.0, 0.0), intensity=8.0, inner_cone=0.2, outer_cone=0.6))
    s.add_light(sc.DirectionalLight(direction=(0.3, -1.0, -0.5), intensity=1.5))
    s.add_light(sc.EnvironmentMap(env, 0.6))
    s.add_light(sc.Environment((0.05, 0.05, 0.08)))
    s.camera = sc.Camera(transform=((1.0, 0.0, 0.0, 0.0), (0.0, 0.9397, 0.342, 3.0), (0.0, -0.342, 0.9397, 6.5)))
    return s


def icosphere(level: int) -> sc.Mesh:
    """Unit icosphere by midpoint subdivision: 20 * 4^level triangles, smooth vertex normals."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
             [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    verts = [list(np.array(v, np.float64) / np.linalg.norm(v)) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache: Dict[Tuple[int, int], int] = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = (np.array(verts[a]) + np.array(verts[b])) / 2.0
                verts.append(list(m / np.linalg.norm(m)))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    v = np.array(verts, np.float32)
    return sc.Mesh(v, np.array(faces, np.uint32).reshape(-1), normals=v.copy())


def procedural(spheres: int = 1024, seed: int = 1, level: int = 3) -> sc.Scene:
    """The procedural scale workload: instanced icospheres in a 20 m cube, a sun and a constant map."""
    rng = np.random.default_rng(seed)
    s = sc.Scene()
    mesh = s.add_mesh(icosphere(level))
    centres = rng.uniform(-10.0, 10.0, size=(spheres, 3))
    radii = rng.lognormal(0.0, 0.5, size=spheres)
    rough = rng.uniform(0.05, 1.0, size=spheres)
    metal = rng.random(spheres) < 0.3
    colour = rng.uniform(0.2, 0.9, size=(spheres, 3))
    for k in range(spheres):
        mat = s.add_material(sc.BaseBsdf(base_color=tuple(float(c) for c in colour[k]),
                                         metallic=1.0 if metal[k] else 0.0, roughness=float(rough[k])))
        r = float(radii[k])
        cx, cy, cz = (float(c) for c in centres[k])
        s.add_instance(sc.Instance(mesh, mat, ((r, 0.0, 0.0, cx), (0.0, r, 0.0, cy), (0.0, 0.0, r, cz))))
    s.add_light(sc.DirectionalLight(direction=(-0.4, -1.0, -0.3), color=(1.0, 0.95, 0.85), intensity=3.0))
    env = s.add_image(sc.Image(32, 16, np.tile(np.array([[0.35, 0.45, 0.6]], np.float32), (32 * 16, 1))))
    s.add_light(sc.EnvironmentMap(env, 1.0))
    s.camera = sc.Camera(transform=((1.0, 0.0, 0.0, 0.0), (0.0, 1.0, 0.0, 0.0), (0.0, 0.0, 1.0, 34.0)))
    return s


def many_lights(emitters: int = 10000, clutter_boxes: int = 41000, seed: int = 1) -> sc.Scene:
    """Many-lights workload: a ground plane and box clutter (12 triangles per box) with `emitters`
    emissive triangles of lognormal(0, 1) emission, all seeded."""
    rng = np.random.default_rng(seed)
    s = sc.Scene()
    ground = s.add_mesh(sc.Mesh(np.array([[-30, 0, -30], [30, 0, -30], [30, 0, 30], [-30, 0, 30]], np.float32),
                                np.array([0, 2, 1, 0, 3, 2], np.uint32)))
    gm = s.add_material(sc.BaseBsdf(base_color=(0.6, 0.6, 0.6), roughness=0.6))
    s.add_instance(sc.Instance(ground, gm))
    box = s.add_mesh(cube_mesh(0.5))
    bm = s.add_material(sc.BaseBsdf(base_color=(0.5, 0.5, 0.55), roughness=0.4))
    # clutter: one mesh of many boxes (non-instanced) so the triangle count is the scene's
    centres = rng.uniform(-25, 25, size=(clutter_boxes, 2))
    sizes = rng.uniform(0.1, 0.8, size=(clutter_boxes, 3))
    cube = cube_mesh(0.5)
    pos = []
    idx = []
    for k in range(clutter_boxes):
        p = cube.positions * sizes[k].astype(np.float32) + np.array([centres[k, 0], sizes[k, 1] / 2, centres[k, 1]], np.float32)
        idx.append(cube.indices + 8 * k)
        pos.append(p)
    clutter = s.add_mesh(sc.Mesh(np.concatenate(pos), np.concatenate(idx)))
    s.add_instance(sc.Instance(clutter, bm))
    del box
    # emissive triangles: small panels floating above the clutter
    tri = []
    for k in range(emitters):
        c = np.array([rng.uniform(-25, 25), rng.uniform(1.5, 4.0), rng.uniform(-25, 25)])
        a = rng.uniform(0, 2 * math.pi)
        e1 = np.array([math.cos(a), 0, math.sin(a)]) * 0.3
        e2 = np.array([-math.sin(a), -0.2, math.cos(a)]) * 0.3
        tri.append([c, c + e1, c + e2])
    emission = rng.lognormal(0.0, 1.0, size=emitters)
    # group emitters into 16 materials by emission quantile so the table stays small
    order = np.argsort(emission)
    groups = np.array_split(order, 16)
    for g in groups:
        lev = float(np.mean(emission[g]))
        mat = s.add_material(sc.lambert((0.0, 0.0, 0.0), (lev, lev * 0.9, lev * 0.8)))
        verts = np.array([tri[k] for k in g], np.float32).reshape(-1, 3)
        m = s.add_mesh(sc.Mesh(verts, np.arange(verts.shape[0], dtype=np.uint32)))
        s.add_instance(sc.Instance(m, mat))
    s.camera = sc.Camera(transform=((1.0, 0.0, 0.0, 0.0), (0.0, 0.9578, 0.2873, 9.0), (0.0, -0.2873, 0.9578, 32.0)))
    return s


def reference_scale_glb(lat: int = 1250, lon: int = 2000, tex: int = 4096) -> bytes:
    """The reference's scale benchmark scene (scripts/gen_scale_scene.py: REQ-PERF-003), built here
    as one GLB document: a UV sphere of 2 lat lon triangles (5.0M) with normals and uvs over
    [0, 8] x [0, 4] at (0, 1, 0) on a 12 m ground quad, four tex x tex PNG textures repeating a
    256 x 256 tile (base colour, normal, metallic-roughness on the sphere; a checker on the ground)
    and a point light of 600 at (3, 6, 4). Load it with verion_amd.gltf.load_gltf_bytes; with no
    camera in the file the CLI frames one (Scene.frame_camera)."""
    import io
    import json
    import struct

    from PIL import Image as PilImage

    def tile_png(fn):
        y, x = np.mgrid[0:256, 0:256].astype(np.float64)
        t = np.stack(fn(x, y), axis=-1)
        full = np.tile(t, (tex // 256, tex // 256, 1)).astype(np.uint8)
        buf = io.BytesIO()
        PilImage.fromarray(full, "RGB").save(buf, "PNG", compress_level=1)
        return buf.getvalue()

    clamp = lambda v: np.clip(np.trunc(v), 0, 255)  # max(0, min(255, int(v)))
    const = lambda c: lambda x, y: np.full_like(x, c)
    images = [
        tile_png(lambda x, y: (clamp(128 + 100 * np.sin(x / 9)), clamp(128 + 100 * np.sin(y / 7)), const(90)(x, y))),
        tile_png(lambda x, y: (clamp(128 + 40 * np.sin(x / 5)), clamp(128 + 40 * np.sin(y / 5)), const(230)(x, y))),
        tile_png(lambda x, y: (const(0)(x, y), clamp(140 + 80 * np.sin((x + y) / 11)), clamp(60 + 40 * np.sin(x / 13)))),
        tile_png(lambda x, y: (np.where((x // 32 + y // 32) % 2 == 1, 200, 90),) * 3),
    ]
    i = np.arange(lat + 1, dtype=np.float64)[:, None]
    j = np.arange(lon + 1, dtype=np.float64)[None, :]
    th, ph = np.pi * i / lat, 2 * np.pi * j / lon
    n = np.stack(np.broadcast_arrays(np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph)), axis=-1).reshape(-1, 3)
    pos = (n + np.array([0.0, 1.0, 0.0])).astype(np.float32)
    nrm = n.astype(np.float32)
    uv = np.stack(np.broadcast_arrays(j / lon * 8 + 0 * i, i / lat * 4 + 0 * j), axis=-1).reshape(-1, 2).astype(np.float32)
    a = (np.arange(lat)[:, None] * (lon + 1) + np.arange(lon)[None, :]).reshape(-1).astype(np.uint32)
    b = a + np.uint32(lon + 1)
    idx = np.stack([a, b, a + 1, a + 1, b, b + 1], axis=-1).reshape(-1).astype(np.uint32)
    gp = np.array([-6, 0, -6, -6, 0, 6, 6, 0, 6, 6, 0, -6], np.float32)
    gn = np.array([0, 1, 0] * 4, np.float32)
    gu = np.array([0, 0, 0, 6, 6, 6, 6, 0], np.float32)
    gi = np.array([0, 1, 2, 0, 2, 3], np.uint32)
    chunks = [c.tobytes() for c in (pos, nrm, uv, idx, gp, gn, gu, gi)] + images
    views, blob = [], b""
    for c in chunks:
        views.append({"buffer": 0, "byteOffset": len(blob), "byteLength": len(c)})
        blob += c + b"\0" * (-len(c) % 4)
    nv = len(pos)
    acc = lambda v, t, count, ct=5126, **kw: {"bufferView": v, "componentType": ct, "count": count, "type": t, **kw}
    doc = {
        "asset": {"version": "2.0"}, "scene": 0, "scenes": [{"nodes": [0, 1, 2]}],
        "extensionsUsed": ["KHR_lights_punctual"],
        "extensions": {"KHR_lights_punctual": {"lights": [{"type": "point", "color": [1, 1, 1], "intensity": 600}]}},
        "nodes": [{"mesh": 0}, {"mesh": 1}, {"translation": [3, 6, 4], "extensions": {"KHR_lights_punctual": {"light": 0}}}],
        "meshes": [
            {"primitives": [{"attributes": {"POSITION": 0, "NORMAL": 1, "TEXCOORD_0": 2}, "indices": 3, "material": 0}]},
            {"primitives": [{"attributes": {"POSITION": 4, "NORMAL": 5, "TEXCOORD_0": 6}, "indices": 7, "material": 1}]},
        ],
        "materials": [
            {"pbrMetallicRoughness": {"baseColorTexture": {"index": 0}, "metallicRoughnessTexture": {"index": 2}},
             "normalTexture": {"index": 1}},
            {"pbrMetallicRoughness": {"baseColorTexture": {"index": 3}, "metallicFactor": 0, "roughnessFactor": 1}},
        ],
        "textures": [{"source": k} for k in range(4)],
        "images": [{"bufferView": 8 + k, "mimeType": "image/png"} for k in range(4)],
        "accessors": [acc(0, "VEC3", nv, min=[-1, 0, -1], max=[1, 2, 1]), acc(1, "VEC3", nv), acc(2, "VEC2", nv),
                      acc(3, "SCALAR", len(idx), 5125), acc(4, "VEC3", 4, min=[-6, 0, -6], max=[6, 0, 6]), acc(5, "VEC3", 4),
                      acc(6, "VEC2", 4), acc(7, "SCALAR", 6, 5125)],
        "bufferViews": views, "buffers": [{"byteLength": len(blob)}],
    }
    js = json.dumps(doc).encode("utf-8")
    js += b" " * (-len(js) % 4)
    total = 12 + 8 + len(js) + 8 + len(blob)
    return (struct.pack("<III", 0x46546C67, 2, total) + struct.pack("<II", len(js), 0x4E4F534A) + js
            + struct.pack("<II", len(blob), 0x004E4942) + blob)
