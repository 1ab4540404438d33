"""The reference's own GPU-test assertions (crates/verion-gpu-metal/tests/pathtrace.rs) restated on
the CPU oracle, for the shading features the closed forms of test_oracle_closed_forms.py leave
open: environment maps (furnace, importance sampling of a bright texel), dispersion, the thin lens,
vertex normals and normal maps, area-light form factors, textures and their mip chain, light slots
and the virtual main light; and an isolating scene for ReSTIR GI's split of first-bounce emission
(pt_core.slang:1162-1168). Same scenes, sizes and tolerances as the reference's tests."""
import math

import numpy as np
import pytest

import oracle
from helpers import center_mean, center_pixel, oracle_render, params
from verion_amd import pack, readback
from verion_amd import scene as sc
from verion_amd import scenes


def hdri(w, h, f):
    """pathtrace.rs:929-935: a lat-long map whose texel (x, y) holds f(x, y) in every channel."""
    px = np.array([[f(x, y)] * 3 for y in range(h) for x in range(w)], np.float32)
    return sc.Image(w, h, px)


def mean_all(img):
    return float(np.asarray(img)[:, 0].mean())


# ------------------------------------------------------------------ environment maps (:937-1013)

def test_constant_environment_map_is_a_white_furnace():
    s = sc.Scene()
    m = s.add_mesh(scenes.cube_mesh(0.5))
    s.add_instance(sc.Instance(m, s.add_material(sc.lambert((1.0, 1.0, 1.0)))))
    s.add_light(sc.EnvironmentMap(s.add_image(hdri(32, 16, lambda x, y: 1.0)), 1.0))
    s.camera = scenes.camera_at_z(3.0)
    img, _ = oracle_render(s, 32, 32, 64)
    assert abs(center_mean(img, 32, 16) - 1.0) < 0.015
    assert abs(mean_all(img) - 1.0) < 0.015


def test_bright_texel_is_importance_sampled():
    w, h, i, j, radiance = 64, 32, 0, 8, 200.0
    s = sc.Scene()
    floor = s.add_mesh(scenes.plane_mesh(25.0, 0.0))
    s.add_instance(sc.Instance(floor, s.add_material(sc.lambert((1.0, 1.0, 1.0)))))
    s.add_light(sc.EnvironmentMap(s.add_image(hdri(w, h, lambda x, y: radiance if (x, y) == (i, j) else 0.0)), 1.0))
    s.camera = scenes.camera_at_z(3.0)
    n = 64
    e = 0.0
    for a in range(n):
        for b in range(n):
            u = (i + (a + 0.5) / n) / w
            v = (j + (b + 0.5) / n) / h
            phi, theta = (u - 0.5) * 2.0 * math.pi, v * math.pi
            d_omega = math.sin(theta) * (math.pi / (h * n)) * (2.0 * math.pi / (w * n))
            e += radiance * (-math.sin(theta) * math.cos(phi)) * d_omega
    expected = e / math.pi
    assert expected > 0.1
    img, _ = oracle_render(s, 65, 65, 64)
    got = center_mean(img, 65, 5)
    assert abs(got - expected) / expected < 0.03, (got, expected)


# ------------------------------------------------------------------ dispersion (:701-760)

def test_dispersion_gives_each_channel_its_own_fresnel_reflection():
    s = sc.Scene()
    glass = s.add_material(sc.BaseBsdf(base_color=(1.0, 1.0, 1.0), roughness=0.0, transmission=1.0, dispersion=4.0))
    floor = s.add_material(sc.BaseBsdf(base_color=(0.0, 0.0, 0.0), metallic=1.0, roughness=1.0))
    for mesh, mat in ((scenes.plane_mesh(5.0, 0.0), glass), (scenes.plane_mesh(5.0, -1.0), floor)):
        s.add_instance(sc.Instance(s.add_mesh(mesh), mat))
    s.add_light(sc.Environment((1.0, 1.0, 1.0)))
    s.camera = scenes.camera_at_z(3.0)
    img, _ = oracle_render(s, 32, 32, 512)
    got = [center_mean(img, 32, 16, c) for c in range(3)]
    for c, want in enumerate((0.0293, 0.04, 0.0465)):
        assert abs(got[c] - want) < 0.006, got
    assert got[0] < got[1] < got[2], got


def test_dispersive_glass_furnace_conserves_energy():
    sphere = scenes.icosphere(2)
    sphere = sc.Mesh(sphere.positions * np.float32(0.5), sphere.indices)
    m = sc.BaseBsdf(base_color=(1.0, 1.0, 1.0), roughness=0.0, transmission=1.0, dispersion=2.0)
    img, _ = oracle_render(scenes.sky_scene(sphere, m, 3.0), 32, 32, 256)
    for c in range(3):
        got = center_mean(img, 32, 16, c)
        assert 0.97 < got < 1.03, (c, got)


# ------------------------------------------------------------------ thin lens (:446-465)

def test_thin_lens_keeps_the_focal_plane_sharp():
    s = sc.Scene()
    s.add_instance(sc.Instance(s.add_mesh(scenes.plane_mesh(5.0, 0.0)), s.add_material(sc.lambert((0.0, 0.0, 0.0), (1.0, 1.0, 1.0)))))
    s.camera = sc.Camera(transform=((1.0, 0.0, 0.0, 0.0), (0.0, 1.0, 0.0, 0.0), (0.0, 0.0, 1.0, 3.0)),
                         focus_distance=3.0, aperture_radius=0.2)
    img, _ = oracle_render(s, 32, 32, 8)
    assert np.all(np.abs(img[:, 0] - 1.0) < 1e-4)


def test_thin_lens_blurs_off_the_focal_plane():
    # the converse, from the same lens: a sharp emissive edge off the focal plane spreads over pixels
    # that a pinhole leaves fully dark or fully lit
    def render(aperture):
        s = sc.Scene()
        half = sc.Mesh(np.array([[0, -5, 0], [5, -5, 0], [5, 5, 0], [0, 5, 0]], np.float32), np.array([0, 1, 2, 0, 2, 3], np.uint32))
        s.add_instance(sc.Instance(s.add_mesh(half), s.add_material(sc.lambert((0.0, 0.0, 0.0), (1.0, 1.0, 1.0)))))
        s.camera = sc.Camera(transform=((1.0, 0.0, 0.0, 0.0), (0.0, 1.0, 0.0, 0.0), (0.0, 0.0, 1.0, 3.0)),
                             focus_distance=1.0, aperture_radius=aperture)
        return oracle_render(s, 32, 32, 64)[0][:, 0].reshape(32, 32)
    pin, lens = render(0.0), render(0.2)
    partial = lambda im: int(np.sum((im > 0.05) & (im < 0.95)))
    assert partial(lens) > 4 * max(partial(pin), 1)
    assert abs(lens.mean() - pin.mean()) < 0.05  # the blur moves light, it does not add any


# ------------------------------------------------------------------ vertex normals, normal maps (:332-444)

def shaded_floor(normal, transform=sc.IDENTITY):
    h = 5.0
    mesh = sc.Mesh(np.array([[-h, -h, 0], [h, -h, 0], [h, h, 0], [-h, h, 0]], np.float32), np.array([0, 1, 2, 0, 2, 3], np.uint32),
                   normals=np.array([normal] * 4, np.float32))
    s = sc.Scene()
    s.add_instance(sc.Instance(s.add_mesh(mesh), s.add_material(sc.lambert((1.0, 1.0, 1.0))), transform))
    s.add_light(sc.DirectionalLight(direction=(0.0, 0.0, -1.0), intensity=math.pi))
    s.camera = scenes.camera_at_z(3.0)
    return center_pixel(oracle_render(s, 65, 65, 64)[0], 65)


def test_vertex_normals_drive_the_shading():
    assert abs(shaded_floor((0.8660254, 0.0, 0.5)) - 0.5) < 0.01


def test_vertex_normals_follow_a_non_uniform_scale():
    scale = ((2.0, 0.0, 0.0, 0.0), (0.0, 1.0, 0.0, 0.0), (0.0, 0.0, 1.0, 0.0))
    assert abs(shaded_floor((1.0, 0.0, 1.0), scale) - 0.894427) < 0.01


def normal_mapped_floor(tangent_normal, scale, to_light):
    h = 5.0
    mesh = sc.Mesh(np.array([[-h, -h, 0], [h, -h, 0], [h, h, 0], [-h, h, 0]], np.float32), np.array([0, 1, 2, 0, 2, 3], np.uint32),
                   uvs=np.array([[0, 1], [1, 1], [1, 0], [0, 0]], np.float32))
    s = sc.Scene()
    texel = [c * 0.5 + 0.5 for c in tangent_normal]
    nmap = s.add_image(sc.Image(1, 1, np.array([texel], np.float32)))
    mat = s.add_material(sc.BaseBsdf(base_color=(1.0, 1.0, 1.0), ior=1.0, normal_texture=nmap, normal_scale=scale))
    s.add_instance(sc.Instance(s.add_mesh(mesh), mat))
    s.add_light(sc.DirectionalLight(direction=tuple(-c for c in to_light), intensity=math.pi))
    s.camera = scenes.camera_at_z(3.0)
    return center_pixel(oracle_render(s, 65, 65, 64)[0], 65)


def test_normal_map_tilts_the_shading_normal():
    t0, t1 = 0.8660254, 0.5
    x = normal_mapped_floor((t0, 0.0, t1), 1.0, (t0, 0.0, t1))
    y = normal_mapped_floor((0.0, t0, t1), 1.0, (0.0, t0, t1))
    assert abs(x - 1.0) < 0.01 and abs(y - 1.0) < 0.01, (x, y)
    away = normal_mapped_floor((0.0, -t0, t1), 1.0, (0.0, t0, t1))
    assert away < 0.01, away


def test_normal_scale_flattens_the_tilt():
    assert abs(normal_mapped_floor((0.8660254, 0.0, 0.5), 0.0, (0.0, 0.0, 1.0)) - 1.0) < 0.01


# ------------------------------------------------------------------ area lights (:843-927)

def floor_with_emitter(floor_material, light_mesh, le):
    s = sc.Scene()
    s.add_instance(sc.Instance(s.add_mesh(scenes.plane_mesh(25.0, 0.0)), s.add_material(floor_material)))
    up = ((1.0, 0.0, 0.0, 0.0), (0.0, 1.0, 0.0, 0.0), (0.0, 0.0, 1.0, 1.0))
    s.add_instance(sc.Instance(s.add_mesh(light_mesh), s.add_material(sc.lambert((0.0, 0.0, 0.0), (le, le, le))), up))
    s.camera = scenes.camera_at_z(0.5)
    return s


def test_large_area_light_over_a_floor_gives_rho_times_radiance():
    s = floor_with_emitter(sc.lambert((0.5, 0.5, 0.5)), scenes.plane_mesh(25.0, 0.0), 2.0)
    got = center_mean(oracle_render(s, 65, 65, 64)[0], 65, 5)
    assert abs(got - 0.5 * 2.0 * 0.999) < 0.015, got


def test_small_area_light_matches_the_square_form_factor():
    a = 0.1
    root = math.sqrt(1.0 + a * a)
    corner = (a / root) * math.atan(a / root) / math.pi
    rho, le = 1.0, 50.0
    expected = rho * 4.0 * corner * le
    s = floor_with_emitter(sc.lambert((rho, rho, rho)), scenes.plane_mesh(0.1, 0.0), le)
    got = center_mean(oracle_render(s, 65, 65, 64)[0], 65, 5)
    assert abs(got - expected) / expected < 0.03, (got, expected)


@pytest.mark.parametrize("roughness,lo", [(0.0, 0.995), (0.3, 0.95), (0.5, 0.9)])
def test_glossy_metal_reflects_a_large_area_light_without_gain_or_loss(roughness, lo):
    metal = sc.BaseBsdf(base_color=(1.0, 1.0, 1.0), metallic=1.0, roughness=roughness)
    got = center_mean(oracle_render(floor_with_emitter(metal, scenes.plane_mesh(25.0, 0.0), 1.0), 65, 65, 64)[0], 65, 5)
    assert lo < got < 1.005, got


# ------------------------------------------------------------------ textures (:1045-1170)

def textured_plane(half, repeat):
    h = half
    return sc.Mesh(np.array([[-h, -h, 0], [h, -h, 0], [h, h, 0], [-h, h, 0]], np.float32), np.array([0, 1, 2, 0, 2, 3], np.uint32),
                   uvs=np.array([[0, repeat], [repeat, repeat], [repeat, 0], [0, 0]], np.float32))


def texture_scene(mesh, image, base):
    s = sc.Scene()
    img = s.add_image(image)
    s.add_instance(sc.Instance(s.add_mesh(mesh), s.add_material(base(img))))
    s.add_light(sc.Environment((1.0, 1.0, 1.0)))
    s.camera = scenes.camera_at_z(3.0)
    return s


def lambert_tex(base, tex):
    return sc.BaseBsdf(base_color=base, ior=1.0, base_color_texture=tex)


def test_base_color_texels_land_where_the_uvs_say():
    tex = sc.Image(2, 2, np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [0.5, 0.5, 0.5]], np.float32))
    img, _ = oracle_render(texture_scene(textured_plane(0.5, 1.0), tex, lambda t: lambert_tex((1.0, 1.0, 1.0), t)), 65, 65, 64)
    at = lambda x, y: img.reshape(65, 65, 3)[y, x]
    close = lambda a, b: np.all(np.abs(np.asarray(a) - np.asarray(b)) < 0.1)
    assert close(at(26, 26), [1, 0, 0]) and close(at(39, 26), [0, 1, 0])
    assert close(at(26, 39), [0, 0, 1]) and close(at(39, 39), [0.5, 0.5, 0.5])


def test_minified_checkerboard_averages_through_the_mip_chain():
    px = np.array([[1.0] * 3 if (i % 8 + i // 8) % 2 == 0 else [0.0] * 3 for i in range(64)], np.float32)
    s = texture_scene(textured_plane(5.0, 256.0), sc.Image(8, 8, px), lambda t: lambert_tex((1.0, 1.0, 1.0), t))
    img, _ = oracle_render(s, 64, 64, 1)
    assert np.abs(img[:, 0] - 0.5).max() < 0.05


def test_metallic_roughness_texture_scales_the_factors():
    tex = sc.Image(1, 1, np.array([[0.0, 0.0, 1.0]], np.float32))
    s = texture_scene(textured_plane(5.0, 1.0), tex, lambda t: sc.BaseBsdf(base_color=(0.5, 0.5, 0.5), metallic=1.0, roughness=1.0,
                                                                         metallic_roughness_texture=t))
    assert abs(center_mean(oracle_render(s, 32, 32, 16)[0], 32, 16) - 0.5) < 0.01


def test_emission_texture_scales_the_emitted_radiance():
    tex = sc.Image(1, 1, np.array([[0.25, 0.25, 0.25]], np.float32))
    s = texture_scene(textured_plane(5.0, 1.0), tex, lambda t: sc.BaseBsdf(base_color=(0.0, 0.0, 0.0), emission=(1.0, 1.0, 1.0), ior=1.0,
                                                                         emission_texture=t))
    assert abs(center_mean(oracle_render(s, 32, 32, 8)[0], 32, 16) - 0.25) < 1e-3


# ------------------------------------------------------------------ light slots, virtual main light (:1387-1592)

def lighting_at_center(s, spp):
    img, st = oracle_render(s, 65, 65, spp, lighting=True)
    inv = readback.inverse_samples(None, 65 * 65, spp)
    return img, readback.lighting(st["lighting"], inv), 32 * 65 + 32


POINT = dict(position=(1.0, 0.0, 1.0), intensity=math.pi * 2.0 * math.sqrt(2.0))


def test_lighting_split_files_light_by_source():
    _, l, i = lighting_at_center(scenes.floor_scene(sc.PointLight(**POINT)), 64)
    assert abs(l.direct_diffuse[i][0] - 1.0) < 0.01
    assert l.specular[i][0] < 0.05 and l.emission[i][0] == 0.0
    assert abs(l.visibility[i] - 1.0) < 1e-6
    _, l, i = lighting_at_center(scenes.floor_scene(sc.PointLight(**POINT), occluder=True), 64)
    assert l.direct_diffuse[i][0] < 1e-3 and l.visibility[i] < 1e-3


def test_the_strongest_lights_get_the_own_light_slots():
    s = scenes.floor_scene(sc.PointLight(position=(1.0, 0.0, 1.0), intensity=1.0))
    s.add_light(sc.PointLight(position=(-1.0, 0.0, 1.0), intensity=10.0))
    s.add_light(sc.PointLight(position=(0.0, 1.0, 1.0), intensity=20.0))
    s.add_light(sc.PointLight(position=(0.0, -1.0, 1.0), intensity=5.0))
    _, l, i = lighting_at_center(s, 16)
    d = l.light_dir[i]
    unit = lambda v: np.asarray(v, np.float64) / np.linalg.norm(v)
    assert np.all(np.abs(d[0] - unit([0, 1, 1])) < 5e-3), d[0]
    assert np.all(np.abs(d[1] - unit([-1, 0, 1])) < 5e-3), d[1]
    assert d[2][1] < -0.1 and d[2][0] > 0.0, d[2]


def test_light_slots_hold_potential_diffuse_and_response():
    _, l, i = lighting_at_center(scenes.floor_scene(sc.PointLight(**POINT)), 64)
    u, r, dd, d, lit, spec = l.light_u[i], l.light_response[i], l.direct_diffuse[i][0], l.light_dir[i], l.light_lit[i], l.light_s[i]
    to_light = np.array([math.sqrt(0.5), 0.0, math.sqrt(0.5)])
    assert np.all(np.abs(d[0] - to_light) < 1e-3) and abs(lit[0] - 1.0) < 1e-3
    assert np.all(np.abs(spec[0]) < 1e-6)
    assert np.all(d[1] == 0.0) and lit[1] == 0.0
    assert abs(r[0] - math.sqrt(0.5)) < 1e-3 and abs(u[0][0] * r[0] - dd) < 1e-3
    assert u[1][0] == 0.0 and u[3][0] == 0.0
    _, l, i = lighting_at_center(scenes.floor_scene(sc.PointLight(**POINT), occluder=True), 64)
    assert l.direct_diffuse[i][0] < 1e-3 and l.light_response[i][0] == 0.0
    assert l.light_u[i][0][0] > 0.1
    assert abs(l.light_dir[i][0][0] - to_light[0]) < 1e-3 and l.light_lit[i][0] == 0.0


def test_virtual_main_light_fills_slot_three_and_leaves_the_beauty():
    def build(virtual):
        s = sc.Scene()
        s.add_instance(sc.Instance(s.add_mesh(scenes.plane_mesh(25.0, 0.0)), s.add_material(sc.lambert((1.0, 1.0, 1.0)))))
        s.add_light(sc.EnvironmentMap(s.add_image(hdri(64, 32, lambda x, y: 50.0 if x < 4 and 8 <= y < 12 else 0.1)), 1.0))
        s.camera = scenes.camera_at_z(3.0)
        s.virtual_light = virtual
        return s

    def run(s):
        img, l, i = lighting_at_center(s, 256)
        beauty = center_pixel(oracle_render(s, 65, 65, 64)[0], 65)
        return beauty, l.light_dir[i][3], l.light_lit[i][3], l.light_response[i][3]

    on, d, lit, resp = run(build(True))
    off, _, _, _ = run(build(False))
    assert abs(on - off) < 0.03 * off, (on, off)
    phi, theta = (2.0 / 64.0 - 0.5) * 2.0 * math.pi, 10.0 / 32.0 * math.pi
    want = np.array([math.sin(theta) * math.sin(phi), math.cos(theta), -math.sin(theta) * math.cos(phi)])
    assert float(np.dot(d, want) / np.linalg.norm(d)) > 0.99, (d, want)
    assert lit > 0.0 and resp > 0.0


# ------------------------------------------------------------------ ReSTIR GI's first-bounce emission split

RESTIR_STRIDE, RESTIR_HALF = 80, 28


def test_gi_keeps_first_bounce_emission_out_of_its_reservoirs():
    # A Lambert floor under a black emissive ceiling and no other light: everything the first bounce
    # brings back is emission seen at depth 1, which counts with the first hit's own light
    # (pt_core.slang:1162-1168) and not in the GI sample. So every first-frame GI reservoir is empty
    # (valid 0, W 0, radiance 0), and with the split left out it would hold the emitter's radiance.
    s = sc.Scene()
    s.add_instance(sc.Instance(s.add_mesh(scenes.plane_mesh(5.0, 0.0)), s.add_material(sc.lambert((0.8, 0.8, 0.8)))))
    up = ((1.0, 0.0, 0.0, 0.0), (0.0, 1.0, 0.0, 0.0), (0.0, 0.0, 1.0, 1.0))
    s.add_instance(sc.Instance(s.add_mesh(scenes.plane_mesh(5.0, 0.0)), s.add_material(sc.lambert((0.0, 0.0, 0.0), (2.0, 2.0, 2.0))), up))
    s.camera = sc.Camera(transform=((1.0, 0.0, 0.0, 0.0), (0.0, 1.0, 0.0, 0.0), (0.0, 0.0, 1.0, 0.5)))
    p = pack.pack(s)
    o = oracle.OracleScene(p)
    par = params(p, 32, 32, restir=(False, True, 8, 4))
    st = o.render(par, 1)
    R = st["restir"].reshape(-1, RESTIR_STRIDE)[:, 0:RESTIR_HALF]  # frame 0 writes half 0
    floor_px = R[:, 27].view(np.uint32) == 0
    assert np.all(floor_px), np.count_nonzero(~floor_px)
    assert np.all(R[:, 25] == 0.0) and np.all(R[:, 22:25] == 0.0)
    # the image itself is the path-traced one: first-bounce emission and NEE are all in d0
    plain = o.render(params(p, 32, 32), 1)["film"]
    np.testing.assert_allclose(st["film"], plain, rtol=1e-5, atol=1e-7)
    assert np.mean(plain.reshape(-1, 3)[:, 0]) > 0.5
