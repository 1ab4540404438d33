"""glTF import (verion_amd/gltf.py) against the expectations of the reference importer's own tests
(crates/verion-render-io/src/gltf_import.rs:696-1059), on small documents built here, and the
reference's Cornell box scene file (tests/golden/cornell_box.gltf, the input of its Vulkan GPU
tests, crates/verion-gpu-vulkan/tests/gpu.rs:15-18) through the packer and the oracle."""
import base64
import io
import json
import math
import os
import struct

import numpy as np
import pytest

import oracle
from verion_amd import gltf, pack
from verion_amd import scene as sc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def triangle_doc(**over):
    """One triangle on node 0 (moved +2 x), a camera on node 1 (z = 5), a point light on node 2 (y = 3)."""
    doc = {
        "asset": {"version": "2.0"},
        "extensionsUsed": ["KHR_lights_punctual", "KHR_materials_volume"],
        "scene": 0,
        "scenes": [{"nodes": [0, 1, 2]}],
        "nodes": [{"mesh": 0, "translation": [2, 0, 0]}, {"camera": 0, "translation": [0, 0, 5]},
                  {"translation": [0, 3, 0], "extensions": {"KHR_lights_punctual": {"light": 0}}}],
        "meshes": [{"primitives": [{"attributes": {"POSITION": 0}, "indices": 1, "material": 0}]}],
        "materials": [{"pbrMetallicRoughness": {"baseColorFactor": [1, 0, 0, 1], "metallicFactor": 0.5, "roughnessFactor": 0.25}}],
        "cameras": [{"type": "perspective", "perspective": {"yfov": 0.5, "znear": 0.1}}],
        "extensions": {"KHR_lights_punctual": {"lights": [{"type": "point", "color": [1, 1, 1], "intensity": 10}]}},
        "accessors": [{"bufferView": 0, "componentType": 5126, "count": 3, "type": "VEC3", "min": [0, 0, 0], "max": [1, 1, 0]},
                      {"bufferView": 1, "componentType": 5123, "count": 3, "type": "SCALAR"}],
        "bufferViews": [{"buffer": 0, "byteOffset": 0, "byteLength": 36}, {"buffer": 0, "byteOffset": 36, "byteLength": 6}],
        "buffers": [{"byteLength": 44}],
    }
    doc.update(over)
    return doc


def glb(doc, extra=b""):
    """Binary glTF: the triangle's positions and indices (44 bytes), then `extra`."""
    bin_ = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32).tobytes() + np.array([0, 1, 2], np.uint16).tobytes() + b"\0\0"
    bin_ += extra
    bin_ += b"\0" * (-len(bin_) % 4)
    j = json.dumps(doc).encode()
    j += b" " * (-len(j) % 4)
    total = 12 + 8 + len(j) + 8 + len(bin_)
    return (struct.pack("<III", 0x46546C67, 2, total) + struct.pack("<II", len(j), 0x4E4F534A) + j
            + struct.pack("<II", len(bin_), 0x004E4942) + bin_)


def png_2x2(mode="RGB", pixels=None):
    from PIL import Image

    im = Image.new(mode, (2, 2))
    im.putdata(pixels)
    b = io.BytesIO()
    im.save(b, format="PNG")
    return b.getvalue()


def load(doc, extra=b""):
    return gltf.load_gltf_bytes(glb(doc, extra))


def test_mesh_material_camera_and_light():
    s, warnings = load(triangle_doc())
    assert len(s.meshes) == 1 and list(s.meshes[0].indices) == [0, 1, 2]
    inst = s.instances[0]
    assert inst.transform[0][3] == 2.0
    m = s.materials[inst.material]
    assert (m.base_color, m.metallic, m.roughness) == ((1.0, 0.0, 0.0), 0.5, 0.25)
    assert s.camera.transform[2][3] == 5.0 and s.camera.vertical_fov == 0.5
    assert s.lights == [sc.PointLight(position=(0.0, 3.0, 0.0), color=(1.0, 1.0, 1.0), intensity=10.0)]
    assert warnings == ["extension KHR_materials_volume is not supported and was ignored"]


def test_spot_light_along_local_minus_z():
    doc = triangle_doc()
    doc["extensions"]["KHR_lights_punctual"]["lights"] = [
        {"type": "spot", "color": [1, 1, 1], "intensity": 10, "spot": {"innerConeAngle": 0.1, "outerConeAngle": 0.5}}]
    s, warnings = load(doc)
    L = s.lights[0]
    assert isinstance(L, sc.SpotLight)
    assert L.position == (0.0, 3.0, 0.0) and L.direction == (-0.0, -0.0, -1.0)
    assert (L.intensity, L.inner_cone, L.outer_cone) == (10.0, 0.1, 0.5)
    # a rotated node: a directional light rotated 90 degrees about +x shines along +y
    doc["nodes"][2] = {"rotation": [math.sin(math.pi / 4), 0, 0, math.cos(math.pi / 4)],
                       "extensions": {"KHR_lights_punctual": {"light": 0}}}
    doc["extensions"]["KHR_lights_punctual"]["lights"] = [{"type": "directional"}]
    s, _ = load(doc)
    assert np.allclose(s.lights[0].direction, (0.0, 1.0, 0.0), atol=1e-6)


def test_clearcoat_and_sheen_are_clamped_and_texture_slots_warned():
    doc = triangle_doc(extensionsUsed=["KHR_materials_clearcoat", "KHR_materials_sheen"])
    doc["materials"][0]["extensions"] = {
        "KHR_materials_clearcoat": {"clearcoatFactor": 1.5, "clearcoatRoughnessFactor": 0.2, "clearcoatTexture": {"index": 0}},
        "KHR_materials_sheen": {"sheenColorFactor": [1, 0.5, 2], "sheenRoughnessFactor": 0.3, "sheenRoughnessTexture": {"index": 0}}}
    s, warnings = load(doc)
    m = s.materials[0]
    assert (m.clearcoat, m.clearcoat_roughness) == (1.0, 0.2)
    assert m.sheen_color == (1.0, 0.5, 1.0) and m.sheen_roughness == 0.3
    assert warnings == ["material 0: clearcoatTexture is not supported yet and was ignored",
                        "material 0: sheenRoughnessTexture is not supported yet and was ignored"]


def test_dispersion_transmission_ior_emission_strength():
    doc = triangle_doc(extensionsUsed=["KHR_materials_dispersion"])
    doc["materials"][0]["emissiveFactor"] = [1, 0.5, 0]
    doc["materials"][0]["extensions"] = {"KHR_materials_dispersion": {"dispersion": 0.36},
                                         "KHR_materials_transmission": {"transmissionFactor": 0.75},
                                         "KHR_materials_ior": {"ior": 1.33},
                                         "KHR_materials_emissive_strength": {"emissiveStrength": 4}}
    s, warnings = load(doc)
    m = s.materials[0]
    assert not warnings
    assert (m.dispersion, m.transmission, m.ior, m.emission) == (0.36, 0.75, 1.33, (4.0, 2.0, 0.0))


def test_alpha_modes_give_opacity():
    def opacity(mode, alpha):
        doc = triangle_doc()
        doc["materials"][0]["pbrMetallicRoughness"]["baseColorFactor"] = [1, 0, 0, alpha]
        if mode:
            doc["materials"][0]["alphaMode"] = mode
        return load(doc)[0].materials[0].opacity

    assert opacity(None, 1.0) == 1.0
    assert opacity(None, 0.25) == 1.0  # OPAQUE ignores the alpha
    assert opacity("BLEND", 0.25) == 0.25
    assert opacity("MASK", 0.25) == 0.0  # under the default cutoff 0.5
    assert opacity("MASK", 0.75) == 1.0


def test_texture_alpha_becomes_a_gray_image():
    rgba = np.array([[[255, 0, 0, 255], [0, 255, 0, 128]], [[0, 0, 255, 0], [9, 9, 9, 200]]], np.uint8)
    blend = gltf.to_alpha_image(rgba, 255.0, None)
    assert (blend.width, blend.height) == (2, 2)
    assert np.array_equal(blend.pixels[:, 0], np.array([1.0, 128 / 255, 0.0, 200 / 255], np.float32))
    assert np.all(blend.pixels[:, 0] == blend.pixels[:, 1]) and np.all(blend.pixels[:, 1] == blend.pixels[:, 2])
    mask = gltf.to_alpha_image(rgba, 255.0, 0.5)
    assert list(mask.pixels[:, 0]) == [1.0, 1.0, 0.0, 1.0]  # 128/255 is just over the cutoff
    assert gltf.to_alpha_image(rgba[:, :, :3], 255.0, None) is None


def textured_doc(tex_coord=0):
    doc = triangle_doc(extensionsUsed=[])
    doc["nodes"] = [{"mesh": 0}]
    doc["scenes"] = [{"nodes": [0]}]
    doc["materials"] = [{"pbrMetallicRoughness": {"baseColorTexture": {"index": 0, "texCoord": tex_coord},
                                                  "metallicRoughnessTexture": {"index": 0}},
                         "normalTexture": {"index": 0}}]
    doc["textures"] = [{"source": 0}]
    doc["images"] = [{"bufferView": 2, "mimeType": "image/png"}]
    return doc


def test_textures_import_as_linear_images():
    png = png_2x2("RGB", [(255, 0, 0), (0, 255, 0), (0, 0, 255), (128, 128, 128)])
    doc = textured_doc()
    doc["bufferViews"].append({"buffer": 0, "byteOffset": 44, "byteLength": len(png)})
    doc["buffers"][0]["byteLength"] = 44 + len(png)
    s, warnings = load(doc, png)
    m = s.materials[0]
    assert m.base_color_texture != m.metallic_roughness_texture  # sRGB and linear uses are separate images
    color = s.images[m.base_color_texture]
    assert (color.width, color.height) == (2, 2)
    assert list(color.pixels[0]) == [1.0, 0.0, 0.0]
    assert abs(color.pixels[3][0] - 0.2158) < 1e-3  # sRGB 128/255
    mr = s.images[m.metallic_roughness_texture]
    assert abs(mr.pixels[3][0] - 128 / 255) < 1e-6
    assert m.normal_texture == m.metallic_roughness_texture and m.normal_scale == 1.0
    assert not warnings
    # a texture on UV set 1 is ignored with a warning
    doc1 = textured_doc(1)
    doc1["bufferViews"] = doc["bufferViews"]
    doc1["buffers"] = doc["buffers"]
    s1, w1 = load(doc1, png)
    assert s1.materials[0].base_color_texture is None and s1.materials[0].metallic_roughness_texture is not None
    assert any("UV set 1" in w for w in w1)


def test_invalid_files_are_errors():
    with pytest.raises(gltf.GltfError):
        gltf.load_gltf_bytes(b"not gltf")
    with pytest.raises(gltf.GltfError):
        gltf.load_gltf_bytes(json.dumps({"asset": {"version": "1.0"}}).encode())


def test_data_uri_buffers_equal_the_binary_chunk():
    doc = triangle_doc()
    bin_ = glb(doc)[-44:]
    doc["buffers"] = [{"byteLength": 44, "uri": "data:application/octet-stream;base64," + base64.b64encode(bin_).decode()}]
    a, _ = gltf.load_gltf_bytes(json.dumps(doc).encode())
    b, _ = load(triangle_doc())
    assert np.array_equal(a.meshes[0].positions, b.meshes[0].positions)


def test_reference_cornell_box_imports_and_renders():
    s, warnings = gltf.load_gltf(os.path.join(GOLDEN, "cornell_box.gltf"))
    assert not warnings
    assert len(s.meshes) == 8 and len(s.instances) == 8 and len(s.materials) == 6
    assert len(s.lights) == 1 and isinstance(s.lights[0], sc.PointLight)
    assert s.camera.vertical_fov == pytest.approx(0.9)
    emitters = [m for m in s.materials if max(m.emission) > 0]
    assert len(emitters) == 1 and emitters[0].emission == (14.0, 14.0, 14.0)
    p = pack.pack(s)
    assert p.info.emitter_count == 2 and p.indices.size // 3 == 18444
    # the box fills the frame and the image is lit (the oracle, 24x18, 2 spp)
    o = oracle.OracleScene(p)
    st = o.render(pack.base_params(p.info, p.view, 24, 18, False, True), 2)
    img = oracle.mean(st, 2)
    assert np.isfinite(img).all() and img.mean() > 0.05


def test_frame_camera_fits_the_bounds_like_the_cli():
    # scene.rs:453-477: centre of the world bounds, backed off along +z until the bounding sphere
    # fills the vertical field of view (0.8), 5 % margin; focus at that distance; a scene without
    # instances keeps its camera
    import math

    from verion_amd import scene as sc
    from verion_amd import scenes

    s = sc.Scene()
    m = s.add_mesh(scenes.cube_mesh(0.5))
    s.add_instance(sc.Instance(m, s.add_material(sc.BaseBsdf()), ((2.0, 0.0, 0.0, 1.0), (0.0, 1.0, 0.0, -2.0), (0.0, 0.0, 1.0, 3.0))))
    s.frame_camera()
    r = math.sqrt(2.0 ** 2 + 1.0 + 1.0) / 2.0
    dist = r / math.sin(0.4) * 1.05
    t = s.camera.transform
    assert abs(t[0][3] - 1.0) < 1e-6 and abs(t[1][3] + 2.0) < 1e-6 and abs(t[2][3] - (3.0 + dist)) < 1e-5
    assert abs(s.camera.focus_distance - dist) < 1e-5 and s.camera.vertical_fov == 0.8
    empty = sc.Scene()
    before = empty.camera
    empty.frame_camera()
    assert empty.camera == before
