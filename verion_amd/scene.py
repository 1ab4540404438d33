"""The scene types the packer consumes: the subset of Verion's scene IR that reaches the GPU tables.

Mirrors crates/verion-scene/src/{scene,material,mesh}.rs: `BaseBsdf` (material.rs:14-78, defaults
at :55-78), `Camera` (scene.rs:70-96, default vfov 0.8), `Light` (scene.rs:39-66), `Instance`,
`Image`, and the light-link masks (scene.rs:389-418, all ones by default). Stylization layers and
colour management never reach the GPU and are left out.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Tuple, Union

import numpy as np

IDENTITY = ((1.0, 0.0, 0.0, 0.0), (0.0, 1.0, 0.0, 0.0), (0.0, 0.0, 1.0, 0.0))


@dataclass
class Image:
    """Linear RGB, row-major from the top (verion_scene::Image)."""

    width: int
    height: int
    pixels: np.ndarray  # (height * width, 3) float32

    def __post_init__(self):
        self.pixels = np.ascontiguousarray(self.pixels, dtype=np.float32).reshape(self.width * self.height, 3)


@dataclass
class Mesh:
    """Triangle mesh (mesh.rs:57-198): positions, optional uvs and vertex normals, u32 indices."""

    positions: np.ndarray
    indices: np.ndarray
    uvs: Optional[np.ndarray] = None
    normals: Optional[np.ndarray] = None

    def __post_init__(self):
        self.positions = np.ascontiguousarray(self.positions, dtype=np.float32).reshape(-1, 3)
        self.indices = np.ascontiguousarray(self.indices, dtype=np.uint32).reshape(-1)
        if self.indices.size % 3:
            raise ValueError("index count is not a multiple of 3")
        if self.indices.size and int(self.indices.max()) >= len(self.positions):
            raise ValueError("index out of range")
        if self.uvs is not None:
            self.uvs = np.ascontiguousarray(self.uvs, dtype=np.float32).reshape(-1, 2)
        if self.normals is not None:
            self.normals = np.ascontiguousarray(self.normals, dtype=np.float32).reshape(-1, 3)


@dataclass
class BaseBsdf:
    """Physically based surface (material.rs:14-78), with the reference defaults."""

    base_color: Tuple[float, float, float] = (0.8, 0.8, 0.8)
    metallic: float = 0.0
    roughness: float = 0.5
    transmission: float = 0.0
    ior: float = 1.5
    clearcoat: float = 0.0
    clearcoat_roughness: float = 0.0
    sheen_color: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    sheen_roughness: float = 0.0
    dispersion: float = 0.0
    opacity: float = 1.0
    opacity_texture: Optional[int] = None
    emission: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    base_color_texture: Optional[int] = None
    metallic_roughness_texture: Optional[int] = None
    emission_texture: Optional[int] = None
    normal_texture: Optional[int] = None
    normal_scale: float = 1.0


def lambert(base_color, emission=(0.0, 0.0, 0.0)) -> BaseBsdf:
    """Pure Lambert: ior = 1 switches the dielectric specular layer off (metal tests/pathtrace.rs:21-32)."""
    return BaseBsdf(base_color=tuple(base_color), emission=tuple(emission), ior=1.0)


@dataclass
class Instance:
    mesh: int
    material: int
    transform: Tuple[Tuple[float, ...], ...] = IDENTITY  # 3x4 row-major object-to-world


@dataclass
class PointLight:
    position: Tuple[float, float, float]
    color: Tuple[float, float, float] = (1.0, 1.0, 1.0)
    intensity: float = 1.0
    link_mask: int = 0xFFFFFFFF
    # LightExtra::ies (scene.rs:181-183): an IES profile (verion_amd.ies.IesProfile) and the light's
    # world frame, rows x, y, z; vertical angle 0 points along -z, horizontal angle 0 along +x
    ies: Optional[Tuple[object, Tuple[Tuple[float, float, float], ...]]] = None


@dataclass
class DirectionalLight:
    direction: Tuple[float, float, float]  # the direction the light travels
    color: Tuple[float, float, float] = (1.0, 1.0, 1.0)
    intensity: float = 1.0
    link_mask: int = 0xFFFFFFFF
    ies: Optional[Tuple[object, Tuple[Tuple[float, float, float], ...]]] = None  # kept, not packed (pack.rs:628)


@dataclass
class SpotLight:
    position: Tuple[float, float, float]
    direction: Tuple[float, float, float]
    color: Tuple[float, float, float] = (1.0, 1.0, 1.0)
    intensity: float = 1.0
    inner_cone: float = 0.0
    outer_cone: float = 0.7853981633974483
    link_mask: int = 0xFFFFFFFF
    ies: Optional[Tuple[object, Tuple[Tuple[float, float, float], ...]]] = None  # as PointLight.ies


@dataclass
class Environment:
    """Uniform sky, reached by BSDF sampling only (pt_core.slang:1021-1022)."""

    color: Tuple[float, float, float]


@dataclass
class EnvironmentMap:
    """Lat-long map, importance sampled (pack.rs:144-211)."""

    image: int
    scale: float = 1.0


Light = Union[PointLight, DirectionalLight, SpotLight, Environment, EnvironmentMap]


@dataclass
class Camera:
    """scene.rs:70-96."""

    transform: Tuple[Tuple[float, ...], ...] = IDENTITY
    vertical_fov: float = 0.8
    focus_distance: float = 1.0
    aperture_radius: float = 0.0
    shift: Tuple[float, float] = (0.0, 0.0)


@dataclass
class Changes:
    """What changed since the last `Scene.take_changes` (scene.rs:109-129): a renderer's update
    redoes only that."""

    geometry: bool = False  # meshes or instances added or re-pointed, or a mesh changed shape
    deformed: List[int] = field(default_factory=list)  # meshes with new vertices, same triangles (sorted)
    instances: List[int] = field(default_factory=list)  # instances with a new transform or material (sorted)
    materials: List[int] = field(default_factory=list)  # materials with new parameters (sorted)
    lights: bool = False
    camera: bool = False

    def is_empty(self) -> bool:
        return self == Changes()


def _mark(ids: List[int], i: int) -> None:
    if i not in ids:
        ids.append(i)
        ids.sort()


@dataclass
class Scene:
    meshes: List[Mesh] = field(default_factory=list)
    materials: List[BaseBsdf] = field(default_factory=list)
    instances: List[Instance] = field(default_factory=list)
    lights: List[Light] = field(default_factory=list)
    images: List[Image] = field(default_factory=list)
    camera: Camera = field(default_factory=Camera)
    transparent_background: bool = False
    instance_light_masks: List[int] = field(default_factory=list)
    instance_emit_masks: List[int] = field(default_factory=list)
    virtual_light: bool = True  # StylizationWorld::virtual_light default
    changes: Changes = field(default_factory=Changes, repr=False, compare=False)

    # The setters record what they change the way scene.rs:268-448 does; assigning the fields
    # directly records nothing.

    def take_changes(self) -> Changes:
        """scene.rs:223-226: the changes since the previous call, which are then forgotten."""
        c, self.changes = self.changes, Changes()
        return c

    def add_mesh(self, mesh: Mesh) -> int:
        self.changes.geometry = True
        self.meshes.append(mesh)
        return len(self.meshes) - 1

    def set_mesh(self, i: int, mesh: Mesh) -> None:
        """scene.rs:275-284: the same triangles over the same vertex count is a deformation."""
        old = self.meshes[i]
        if np.array_equal(old.indices, mesh.indices) and len(old.positions) == len(mesh.positions):
            _mark(self.changes.deformed, i)
        else:
            self.changes.geometry = True
        self.meshes[i] = mesh

    def add_material(self, m: BaseBsdf) -> int:
        self.changes.geometry = True
        self.materials.append(m)
        return len(self.materials) - 1

    def set_material(self, i: int, m: BaseBsdf) -> None:
        if self.materials[i] != m:
            _mark(self.changes.materials, i)
        self.materials[i] = m

    def add_instance(self, inst: Instance) -> int:
        if not (0 <= inst.mesh < len(self.meshes)) or not (0 <= inst.material < len(self.materials)):
            raise ValueError("instance refers to a missing mesh or material")
        self.changes.geometry = True
        self.instances.append(inst)
        return len(self.instances) - 1

    def set_instance(self, i: int, inst: Instance) -> None:
        """scene.rs:331-343: another mesh is a geometry change, another transform or material is not."""
        if not (0 <= inst.mesh < len(self.meshes)) or not (0 <= inst.material < len(self.materials)):
            raise ValueError("instance refers to a missing mesh or material")
        old = self.instances[i]
        if old.mesh != inst.mesh:
            self.changes.geometry = True
        elif old != inst:
            _mark(self.changes.instances, i)
        self.instances[i] = inst

    def add_light(self, light: Light) -> int:
        self.changes.lights = True
        self.lights.append(light)
        return len(self.lights) - 1

    def set_light(self, i: int, light: Light) -> None:
        self.changes.lights = True
        self.lights[i] = light

    def frame_camera(self) -> None:
        """scene.rs:453-477 (the CLI's default when the file has no camera, main.rs:294-296): the
        default camera moved back along +z from the centre of the world-space bounds until a sphere
        around them fills the vertical field of view, with 5 % margin; focus at that distance."""
        f = np.float32
        lo = np.full(3, np.finfo(np.float32).max, np.float32)
        hi = np.full(3, np.finfo(np.float32).min, np.float32)
        for inst in self.instances:
            t = np.asarray(inst.transform, np.float32)
            p = self.meshes[inst.mesh].positions.astype(np.float32)
            for a in range(3):
                v = ((t[a, 0] * p[:, 0] + t[a, 1] * p[:, 1]) + t[a, 2] * p[:, 2]) + t[a, 3]
                if v.size:
                    lo[a] = min(lo[a], v.min())
                    hi[a] = max(hi[a], v.max())
        if lo[0] > hi[0]:
            return
        c = [(lo[a] + hi[a]) / f(2.0) for a in range(3)]
        r = f(np.sqrt(sum(f(hi[a] - lo[a]) ** 2 for a in range(3)))) / f(2.0)
        cam = Camera()
        dist = f(r / f(np.sin(f(cam.vertical_fov) / f(2.0)))) * f(1.05)
        rows = [list(row) for row in cam.transform]
        rows[0][3], rows[1][3], rows[2][3] = float(c[0]), float(c[1]), float(c[2] + dist)
        cam.transform = tuple(tuple(row) for row in rows)
        cam.focus_distance = float(dist)
        self.set_camera(cam)

    def set_camera(self, camera: Camera) -> None:
        self.changes.camera = True
        self.camera = camera

    def set_transparent_background(self, on: bool) -> None:
        self.changes.lights = True
        self.transparent_background = on

    def set_instance_light_mask(self, i: int, mask: int) -> None:
        """scene.rs:393-405: link masks live in the light table."""
        if self.instance_light_mask(i) == mask:
            return
        while len(self.instance_light_masks) <= i:
            self.instance_light_masks.append(0xFFFFFFFF)
        self.changes.lights = True
        self.instance_light_masks[i] = mask

    def set_instance_emit_mask(self, i: int, mask: int) -> None:
        """scene.rs:411-425: emitter records carry the mask, so a change rebuilds."""
        if self.instance_emit_mask(i) == mask:
            return
        while len(self.instance_emit_masks) <= i:
            self.instance_emit_masks.append(0xFFFFFFFF)
        self.changes.geometry = True
        self.instance_emit_masks[i] = mask

    def add_image(self, image: Image) -> int:
        self.images.append(image)
        return len(self.images) - 1

    def instance_light_mask(self, i: int) -> int:
        return self.instance_light_masks[i] if i < len(self.instance_light_masks) else 0xFFFFFFFF

    def instance_emit_mask(self, i: int) -> int:
        return self.instance_emit_masks[i] if i < len(self.instance_emit_masks) else 0xFFFFFFFF
