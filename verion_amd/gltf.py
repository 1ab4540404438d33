"""glTF 2.0 -> scene IR, for the tables the packer consumes.

Follows crates/verion-render-io/src/gltf_import.rs: `convert` (:138-176; unsupported extensions
become warnings), `node` (:180-205; world = parent x local, meshes become instances with the
row-major 3x4 of the world matrix, the first perspective camera, KHR_lights_punctual lights),
`primitive` (:207-244; triangle lists only, POSITION required, NORMAL and TEXCOORD_0 optional,
implicit indices), `material` (:246-431; metallic-roughness factors, emissive x emissive_strength,
transmission, ior, clearcoat and sheen clamped to [0, 1], dispersion >= 0, textures on UV set 0
only, alpha modes OPAQUE / BLEND / MASK, a texture alpha as a gray image), `camera` (:494-512),
`light` (:514-545; point at the world translation, directional and spot along local -z) and the
image conversions `to_image` / `to_alpha_image` (:557-645). Colours stay linear Rec.709 (the
reference's default working space). The VERION_materials_stylization extension is accepted and not
read: stylization layers never reach the GPU tables.

Buffers may be data URIs, files beside the .gltf, or the binary chunk of a .glb.
"""
from __future__ import annotations

import base64
import io
import json
import math
import os
import struct
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import scene as sc

SUPPORTED = {
    "VERION_materials_stylization",
    "KHR_lights_punctual",
    "KHR_materials_transmission",
    "KHR_materials_ior",
    "KHR_materials_emissive_strength",
    "KHR_materials_clearcoat",
    "KHR_materials_sheen",
    "KHR_materials_dispersion",
}

_COMPONENT = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}
_WIDTH = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT2": 4, "MAT3": 9, "MAT4": 16}


class GltfError(ValueError):
    """Not a valid glTF file, or an invalid value in it."""


def _mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """a x b in float32, each entry summed over k in order (gltf_import.rs:99-107)."""
    out = np.zeros((4, 4), np.float32)
    for r in range(4):
        for c in range(4):
            acc = np.float32(0.0)
            for k in range(4):
                acc = np.float32(acc + np.float32(a[r, k] * b[k, c]))
            out[r, c] = acc
    return out


def _local_matrix(node: dict) -> np.ndarray:
    """The node's local transform as a 4x4 float32 (row-major in memory, column vectors): the
    matrix, or translation x rotation x scale."""
    f = np.float32
    if "matrix" in node:
        return np.array(node["matrix"], np.float32).reshape(4, 4).T  # glTF stores it column-major
    t = [f(v) for v in node.get("translation", [0.0, 0.0, 0.0])]
    x, y, z, w = (f(v) for v in node.get("rotation", [0.0, 0.0, 0.0, 1.0]))
    s = [f(v) for v in node.get("scale", [1.0, 1.0, 1.0])]
    x2, y2, z2 = x + x, y + y, z + z
    xx, xy, xz = x * x2, x * y2, x * z2
    yy, yz, zz = y * y2, y * z2, z * z2
    wx, wy, wz = w * x2, w * y2, w * z2
    one = f(1.0)
    rot = [[one - (yy + zz), xy - wz, xz + wy],
           [xy + wz, one - (xx + zz), yz - wx],
           [xz - wy, yz + wx, one - (xx + yy)]]
    m = np.eye(4, dtype=np.float32)
    for r in range(3):
        for c in range(3):
            m[r, c] = rot[r][c] * s[c]
        m[r, 3] = t[r]
    return m


def _srgb_to_linear(c: np.ndarray) -> np.ndarray:
    c = c.astype(np.float32)
    return np.where(c <= np.float32(0.04045), c / np.float32(12.92),
                    np.power((c + np.float32(0.055)) / np.float32(1.055), np.float32(2.4))).astype(np.float32)


class _Reader:
    def __init__(self, doc: dict, bin_chunk: Optional[bytes], base_dir: Optional[str]):
        self.doc = doc
        self.buffers: List[bytes] = []
        for i, b in enumerate(doc.get("buffers", [])):
            uri = b.get("uri")
            if uri is None:
                if bin_chunk is None:
                    raise GltfError(f"buffer {i} has no uri and there is no binary chunk")
                data = bin_chunk
            elif uri.startswith("data:"):
                data = base64.b64decode(uri.split(",", 1)[1])
            else:
                if base_dir is None:
                    raise GltfError(f"buffer {i} refers to the external file {uri}")
                with open(os.path.join(base_dir, uri), "rb") as f:
                    data = f.read()
            if len(data) < b.get("byteLength", 0):
                raise GltfError(f"buffer {i} is shorter than its byteLength")
            self.buffers.append(data)
        self.base_dir = base_dir

    def view_bytes(self, index: int) -> Tuple[bytes, int, Optional[int]]:
        v = self.doc["bufferViews"][index]
        buf = self.buffers[v["buffer"]]
        off = v.get("byteOffset", 0)
        return buf[off:off + v["byteLength"]], v["byteLength"], v.get("byteStride")

    def accessor(self, index: int) -> np.ndarray:
        a = self.doc["accessors"][index]
        dt = np.dtype(_COMPONENT[a["componentType"]])
        width = _WIDTH[a["type"]]
        count = a["count"]
        if "bufferView" not in a:
            out = np.zeros((count, width), dt)
        else:
            data, _, stride = self.view_bytes(a["bufferView"])
            off = a.get("byteOffset", 0)
            elem = dt.itemsize * width
            stride = stride or elem
            raw = np.frombuffer(data, np.uint8, count=(count - 1) * stride + elem if count else 0, offset=off)
            rows = np.lib.stride_tricks.as_strided(raw, shape=(count, elem), strides=(stride, 1)) if count else raw.reshape(0, elem)
            out = np.ascontiguousarray(rows).view(dt).reshape(count, width)
        if a.get("sparse"):
            raise GltfError(f"accessor {index}: sparse accessors are not supported")
        if a.get("normalized") and dt != np.float32:
            out = out.astype(np.float32) / np.float32(np.iinfo(dt).max)
        return out

    def image(self, index: int):
        """(pixels (h, w, c) uint8/uint16, channels) of image `index`."""
        from PIL import Image as PilImage

        img = self.doc["images"][index]
        if "bufferView" in img:
            data = self.view_bytes(img["bufferView"])[0]
        elif img.get("uri", "").startswith("data:"):
            data = base64.b64decode(img["uri"].split(",", 1)[1])
        else:
            if self.base_dir is None:
                raise GltfError(f"image {index} refers to an external file")
            with open(os.path.join(self.base_dir, img["uri"]), "rb") as f:
                data = f.read()
        p = PilImage.open(io.BytesIO(data))
        p.load()
        mode = p.mode
        if mode in ("L", "LA", "RGB", "RGBA"):
            arr = np.asarray(p)
            arr = arr.reshape(arr.shape[0], arr.shape[1], -1)
            return arr, 255.0
        if mode in ("I;16", "I;16B", "I;16L"):
            arr = np.asarray(p).astype(np.uint16)
            return arr.reshape(arr.shape[0], arr.shape[1], 1), 65535.0
        if mode == "P":
            p = p.convert("RGBA" if "transparency" in p.info else "RGB")
            arr = np.asarray(p)
            return arr.reshape(arr.shape[0], arr.shape[1], -1), 255.0
        raise GltfError(f"image {index}: pixel format {mode} is not supported")


def to_image(pixels: np.ndarray, full: float, srgb: bool) -> sc.Image:
    """gltf_import.rs:557-603: one channel becomes grey, two become (r, g, 0); colour data decodes sRGB."""
    h, w, ch = pixels.shape
    v = pixels.astype(np.float32) / np.float32(full)
    if srgb:
        v = _srgb_to_linear(v)
    if ch == 1:
        rgb = np.repeat(v, 3, axis=2)
    elif ch == 2:
        rgb = np.concatenate([v[:, :, :2], np.zeros((h, w, 1), np.float32)], axis=2)
    else:
        rgb = v[:, :, :3]
    return sc.Image(w, h, rgb.reshape(-1, 3))


def to_alpha_image(pixels: np.ndarray, full: float, threshold: Optional[float]) -> Optional[sc.Image]:
    """gltf_import.rs:606-645: the alpha of a four-channel image as grey (MASK: 0 or 1 at the
    threshold); None without an alpha channel."""
    h, w, ch = pixels.shape
    if ch != 4:
        return None
    a = pixels[:, :, 3].astype(np.float32) / np.float32(full)
    if threshold is not None:
        a = (a >= np.float32(threshold)).astype(np.float32)
    return sc.Image(w, h, np.repeat(a.reshape(-1, 1), 3, axis=1))


class _Importer:
    def __init__(self, doc: dict, reader: _Reader):
        self.doc = doc
        self.r = reader
        self.scene = sc.Scene()
        self.warnings: List[str] = []
        self.meshes: Dict[Tuple[int, int], Tuple[int, int]] = {}
        self.materials: Dict[Optional[int], int] = {}
        self.images: Dict[Tuple[int, bool], Optional[int]] = {}
        self.alpha_images: Dict[Tuple[int, Optional[float]], Optional[int]] = {}
        self.textures = [t.get("source") for t in doc.get("textures", [])]
        self.camera_set = False

    # -- nodes
    def node(self, index: int, parent: np.ndarray, depth: int = 0):
        if depth > 256:
            raise GltfError("node hierarchy too deep (a cycle?)")
        n = self.doc["nodes"][index]
        world = _mul(parent, _local_matrix(n))
        transform = tuple(tuple(float(world[r, c]) for c in range(4)) for r in range(3))
        if "mesh" in n:
            for pi in range(len(self.doc["meshes"][n["mesh"]]["primitives"])):
                ids = self.primitive(index, n["mesh"], pi)
                if ids is not None:
                    self.scene.add_instance(sc.Instance(ids[0], ids[1], transform))
        if "camera" in n:
            self.camera(index, n["camera"], transform)
        light = n.get("extensions", {}).get("KHR_lights_punctual")
        if light is not None:
            self.light(light["light"], world)
        for child in n.get("children", []):
            self.node(child, world, depth + 1)

    def primitive(self, node: int, mesh: int, prim: int):
        key = (mesh, prim)
        if key in self.meshes:
            return self.meshes[key]
        p = self.doc["meshes"][mesh]["primitives"][prim]
        if p.get("mode", 4) != 4:
            self.warnings.append(f"mesh {mesh} primitive {prim}: only triangle lists are supported")
            return None
        attrs = p.get("attributes", {})
        if "POSITION" not in attrs:
            self.warnings.append(f"mesh {mesh} primitive {prim}: no POSITION attribute")
            return None
        pos = self.r.accessor(attrs["POSITION"]).astype(np.float32)
        normals = self.r.accessor(attrs["NORMAL"]).astype(np.float32) if "NORMAL" in attrs else None
        uvs = self.r.accessor(attrs["TEXCOORD_0"]).astype(np.float32) if "TEXCOORD_0" in attrs else None
        idx = (self.r.accessor(p["indices"]).reshape(-1).astype(np.uint32) if "indices" in p
               else np.arange(len(pos), dtype=np.uint32))
        try:
            m = sc.Mesh(pos, idx, uvs=uvs, normals=normals)
        except ValueError as e:
            raise GltfError(f"mesh on node {node}: {e}") from None
        mesh_id = self.scene.add_mesh(m)
        mat_id = self.material(p.get("material"))
        self.meshes[key] = (mesh_id, mat_id)
        return mesh_id, mat_id

    # -- materials
    def material(self, index: Optional[int]) -> int:
        if index in self.materials:
            return self.materials[index]
        m = self.doc["materials"][index] if index is not None else {}
        name = str(index) if index is not None else "default"
        pbr = m.get("pbrMetallicRoughness", {})
        r, g, b, alpha = (float(v) for v in pbr.get("baseColorFactor", [1.0, 1.0, 1.0, 1.0]))
        ext = m.get("extensions", {})
        strength = float(ext.get("KHR_materials_emissive_strength", {}).get("emissiveStrength", 1.0))
        e = [float(v) for v in m.get("emissiveFactor", [0.0, 0.0, 0.0])]
        base = sc.BaseBsdf(
            base_color=(r, g, b),
            metallic=float(pbr.get("metallicFactor", 1.0)),
            roughness=float(pbr.get("roughnessFactor", 1.0)),
            transmission=float(ext.get("KHR_materials_transmission", {}).get("transmissionFactor", 0.0)),
            ior=float(ext.get("KHR_materials_ior", {}).get("ior", 1.5)),
            emission=tuple(float(np.float32(c) * np.float32(strength)) for c in e),
        )
        cc = ext.get("KHR_materials_clearcoat")
        if cc is not None:
            base.clearcoat = min(max(float(cc.get("clearcoatFactor", 0.0)), 0.0), 1.0)
            base.clearcoat_roughness = min(max(float(cc.get("clearcoatRoughnessFactor", 0.0)), 0.0), 1.0)
            for key in ("clearcoatTexture", "clearcoatRoughnessTexture", "clearcoatNormalTexture"):
                if key in cc:
                    self.warnings.append(f"material {name}: {key} is not supported yet and was ignored")
        disp = ext.get("KHR_materials_dispersion")
        if disp is not None:
            base.dispersion = max(float(disp.get("dispersion", 0.0)), 0.0)
        sh = ext.get("KHR_materials_sheen")
        if sh is not None:
            col = sh.get("sheenColorFactor")
            col = col if isinstance(col, list) and len(col) == 3 else [0.0, 0.0, 0.0]
            base.sheen_color = tuple(min(max(float(c), 0.0), 1.0) for c in col)
            base.sheen_roughness = min(max(float(sh.get("sheenRoughnessFactor", 0.0)), 0.0), 1.0)
            for key in ("sheenColorTexture", "sheenRoughnessTexture"):
                if key in sh:
                    self.warnings.append(f"material {name}: {key} is not supported yet and was ignored")
        bct = pbr.get("baseColorTexture")
        if bct is not None:
            base.base_color_texture = self.texture(name, "base color", bct, True)
        mrt = pbr.get("metallicRoughnessTexture")
        if mrt is not None:
            base.metallic_roughness_texture = self.texture(name, "metallic-roughness", mrt, False)
        et = m.get("emissiveTexture")
        if et is not None:
            base.emission_texture = self.texture(name, "emissive", et, True)
        nt = m.get("normalTexture")
        if nt is not None:
            base.normal_scale = float(nt.get("scale", 1.0))
            base.normal_texture = self.texture(name, "normal", nt, False)
        if "occlusionTexture" in m:
            self.warnings.append(f"material {name}: occlusion textures are not supported yet and were ignored")
        if "transmissionTexture" in ext.get("KHR_materials_transmission", {}):
            self.warnings.append(f"material {name}: transmission textures are not supported yet and were ignored")
        mode = m.get("alphaMode", "OPAQUE")
        cutoff = float(m.get("alphaCutoff", 0.5))
        threshold = None
        if mode == "MASK":
            threshold = float(np.float32(cutoff) / np.float32(alpha)) if alpha > 0.0 else math.inf
        alpha_tex = None
        if mode != "OPAQUE" and bct is not None:
            alpha_tex = self.alpha_texture(name, bct, threshold)
        if mode == "OPAQUE":
            base.opacity = 1.0
        elif mode == "BLEND":
            base.opacity = alpha
        else:
            base.opacity = 1.0 if alpha_tex is not None else float(alpha >= cutoff)
        base.opacity_texture = alpha_tex
        mid = self.scene.add_material(base)
        self.materials[index] = mid
        return mid

    def _source(self, info: dict) -> int:
        t = info["index"]
        if t >= len(self.textures) or self.textures[t] is None:
            raise GltfError(f"texture {t} does not exist or has no source")
        return self.textures[t]

    def texture(self, material: str, slot: str, info: dict, srgb: bool) -> Optional[int]:
        tc = info.get("texCoord", 0)
        if tc != 0:
            self.warnings.append(f"material {material}: {slot} texture uses UV set {tc}; only set 0 is supported")
            return None
        src = self._source(info)
        key = (src, srgb)
        if key not in self.images:
            try:
                px, full = self.r.image(src)
                self.images[key] = self.scene.add_image(to_image(px, full, srgb))
            except Exception as e:  # noqa: BLE001 - an unusable texture is a warning, as in the reference
                self.warnings.append(f"material {material}: {slot} texture ignored: {e}")
                self.images[key] = None
        return self.images[key]

    def alpha_texture(self, material: str, info: dict, threshold: Optional[float]) -> Optional[int]:
        tc = info.get("texCoord", 0)
        if tc != 0:
            self.warnings.append(f"material {material}: the base color texture uses UV set {tc}; "
                                 "only set 0 is supported, its alpha was ignored")
            return None
        src = self._source(info)
        key = (src, threshold)
        if key not in self.alpha_images:
            try:
                px, full = self.r.image(src)
                img = to_alpha_image(px, full, threshold)
                self.alpha_images[key] = self.scene.add_image(img) if img is not None else None
            except Exception as e:  # noqa: BLE001
                self.warnings.append(f"material {material}: the alpha of the base color texture ignored: {e}")
                self.alpha_images[key] = None
        return self.alpha_images[key]

    # -- cameras and lights
    def camera(self, node: int, index: int, transform):
        cam = self.doc["cameras"][index]
        if cam.get("type") != "perspective":
            self.warnings.append(f"node {node}: orthographic cameras are not supported")
            return
        if self.camera_set:
            self.warnings.append(f"node {node}: only the first camera is used")
            return
        self.camera_set = True
        self.scene.camera = sc.Camera(transform=transform, vertical_fov=float(cam["perspective"]["yfov"]))

    def light(self, index: int, world: np.ndarray):
        lights = self.doc.get("extensions", {}).get("KHR_lights_punctual", {}).get("lights", [])
        L = lights[index]
        color = tuple(float(c) for c in L.get("color", [1.0, 1.0, 1.0]))
        intensity = float(L.get("intensity", 1.0))
        pos = (float(world[0, 3]), float(world[1, 3]), float(world[2, 3]))
        down = (float(-world[0, 2]), float(-world[1, 2]), float(-world[2, 2]))  # local -z
        kind = L.get("type")
        if kind == "point":
            self.scene.add_light(sc.PointLight(position=pos, color=color, intensity=intensity))
        elif kind == "directional":
            self.scene.add_light(sc.DirectionalLight(direction=down, color=color, intensity=intensity))
        elif kind == "spot":
            spot = L.get("spot", {})
            self.scene.add_light(sc.SpotLight(position=pos, direction=down, color=color, intensity=intensity,
                                              inner_cone=float(spot.get("innerConeAngle", 0.0)),
                                              outer_cone=float(spot.get("outerConeAngle", math.pi / 4))))
        else:
            raise GltfError(f"light {index}: unknown type {kind}")


def _parse(data: bytes) -> Tuple[dict, Optional[bytes]]:
    if data[:4] == b"glTF":
        if len(data) < 20:
            raise GltfError("truncated glb")
        _, version, total = struct.unpack_from("<III", data, 0)
        if version != 2:
            raise GltfError(f"glb version {version}")
        off, doc, bin_chunk = 12, None, None
        while off + 8 <= min(total, len(data)):
            length, kind = struct.unpack_from("<II", data, off)
            chunk = data[off + 8:off + 8 + length]
            if kind == 0x4E4F534A:
                doc = json.loads(chunk.decode("utf-8"))
            elif kind == 0x004E4942 and bin_chunk is None:
                bin_chunk = bytes(chunk)
            off += 8 + length
        if doc is None:
            raise GltfError("glb without a JSON chunk")
        return doc, bin_chunk
    try:
        return json.loads(data.decode("utf-8")), None
    except (UnicodeDecodeError, json.JSONDecodeError) as e:
        raise GltfError(f"glTF load failed: {e}") from None


def load_gltf_bytes(data: bytes, base_dir: Optional[str] = None) -> Tuple[sc.Scene, List[str]]:
    """A .glb or .gltf document in memory -> (scene, warnings) (load_gltf_slice, gltf_import.rs:78-87)."""
    doc, bin_chunk = _parse(data)
    if not isinstance(doc, dict) or doc.get("asset", {}).get("version", "")[:1] != "2":
        raise GltfError("glTF load failed: not a glTF 2.0 document")
    imp = _Importer(doc, _Reader(doc, bin_chunk, base_dir))
    for ext in doc.get("extensionsUsed", []):
        if ext not in SUPPORTED:
            imp.warnings.append(f"extension {ext} is not supported and was ignored")
    scenes = doc.get("scenes", [])
    if scenes:
        root = scenes[doc.get("scene", 0)]
        eye = np.eye(4, dtype=np.float32)
        for n in root.get("nodes", []):
            imp.node(n, eye)
    else:
        imp.warnings.append("file contains no scene")
    return imp.scene, imp.warnings


def load_gltf(path: str) -> Tuple[sc.Scene, List[str]]:
    """A .gltf (embedded or external buffers) or .glb file -> (scene, warnings) (load_gltf, :67-75)."""
    with open(path, "rb") as f:
        data = f.read()
    return load_gltf_bytes(data, os.path.dirname(os.path.abspath(path)))
