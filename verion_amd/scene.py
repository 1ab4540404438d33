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


@dataclass
class DirectionalLight:
    direction: Tuple[float, float, float]  # the direction the light travels
    color: Tuple[float, float, float] = (1.0, 1.0, 1.0)
    intensity: float = 1.0
    link_mask: int = 0xFFFFFFFF


@dataclass
class SpotLight:
    position: Tuple[float, float, float]
    direction: Tuple[float, float, float]
    color: Tuple[float, float, float] = (1.0, 1.0, 1.0)
    intensity: float = 1.0
    inner_cone: float = 0.0
    outer_cone: float = 0.7853981633974483
    link_mask: int = 0xFFFFFFFF


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

    def add_mesh(self, mesh: Mesh) -> int:
        self.meshes.append(mesh)
        return len(self.meshes) - 1

    def add_material(self, m: BaseBsdf) -> int:
        self.materials.append(m)
        return len(self.materials) - 1

    def add_instance(self, inst: Instance) -> int:
        if not (0 <= inst.mesh < len(self.meshes)) or not (0 <= inst.material < len(self.materials)):
            raise ValueError("instance refers to a missing mesh or material")
        self.instances.append(inst)
        return len(self.instances) - 1

    def add_light(self, light: Light) -> int:
        self.lights.append(light)
        return len(self.lights) - 1

    def add_image(self, image: Image) -> int:
        self.images.append(image)
        return len(self.images) - 1

    def instance_light_mask(self, i: int) -> int:
        return self.instance_light_masks[i] if i < len(self.instance_light_masks) else 0xFFFFFFFF

    def instance_emit_mask(self, i: int) -> int:
        return self.instance_emit_masks[i] if i < len(self.instance_emit_masks) else 0xFFFFFFFF
