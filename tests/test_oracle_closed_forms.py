"""The oracle against closed forms: the assertions of the reference's GPU tests
(crates/verion-gpu-metal/tests/pathtrace.rs), restated on the CPU oracle. These pin the radiance
path (shadePath, finishPath, the BSDF lobes, light sampling and MIS) independently of the GPU."""
import math

import numpy as np
import pytest

from helpers import center_mean, center_pixel, oracle_render
from verion_amd import scene as sc
from verion_amd import scenes


def test_white_furnace_returns_the_sky_exactly():
    # pathtrace.rs:118-137: albedo-1 convex object under a uniform sky, every pixel 1 within 1e-4
    img, _ = oracle_render(scenes.white_furnace(), 32, 32, 4)
    assert np.all(np.abs(img - 1.0) < 1.0e-4)


@pytest.mark.parametrize("rho", [0.25, 0.5, 0.8])
def test_furnace_generalises_to_albedo_rho(rho):
    # A Lambert convex object of albedo rho under sky 1: one bounce then escape -> rho on the object.
    img, _ = oracle_render(scenes.white_furnace(sc.lambert((rho, rho, rho))), 32, 32, 4)
    centre = img.reshape(32, 32, 3)[11:21, 11:21]  # fully covered by the cube's front face
    assert np.all(np.abs(centre - rho) < 1.0e-4)


def test_closed_emissive_room_matches_the_series_sum():
    # pathtrace.rs:104-115: L = E (1 - rho^8) / (1 - rho) within 1%
    img, _ = oracle_render(scenes.furnace_room(0.5, 1.0), 64, 64, 64)
    want = 1.0 * (1.0 - 0.5 ** 8) / (1.0 - 0.5)
    assert abs(img[:, 0].mean() - want) / want < 0.01


def test_variance_decreases_with_samples():
    v = [oracle_render(scenes.furnace_room(0.5, 1.0), 32, 32, n)[0][:, 0].var() for n in (4, 16, 64)]
    assert v[0] > v[1] > v[2]


def test_point_light_matches_lambert_and_shadows():
    # pathtrace.rs:276-289: L = rho/pi * I cos / d^2 = 1 at the origin; shadowed < 1e-3
    light = sc.PointLight(position=(1.0, 0.0, 1.0), intensity=math.pi * 2.0 * math.sqrt(2.0))
    lit = center_pixel(oracle_render(scenes.floor_scene(light), 65, 65, 64)[0], 65)
    assert abs(lit - 1.0) < 0.01
    shadowed = center_pixel(oracle_render(scenes.floor_scene(light, occluder=True), 65, 65, 64)[0], 65)
    assert shadowed < 1.0e-3


def _spot(off):
    a = math.pi / 4 + off
    return sc.SpotLight(position=(1.0, 0.0, 1.0), direction=(-3.0 * math.sin(a), 0.0, -3.0 * math.cos(a)),
                        intensity=math.pi * 2.0 * math.sqrt(2.0), inner_cone=0.1, outer_cone=0.5)


def test_spot_light_falls_off_between_its_cone_angles():
    # pathtrace.rs:306-317
    lit = lambda off: center_pixel(oracle_render(scenes.floor_scene(_spot(off)), 65, 65, 64)[0], 65)
    assert abs(lit(0.0) - 1.0) < 0.01
    c, i, o = math.cos(0.3), math.cos(0.1), math.cos(0.5)
    assert abs(lit(0.3) - ((c - o) / (i - o)) ** 2) < 0.01
    assert lit(0.6) < 1.0e-3


def test_directional_light_matches_lambert():
    # pathtrace.rs:320-330: sun straight down, L = rho/pi * E = 1
    light = sc.DirectionalLight(direction=(0.0, 0.0, -1.0), intensity=math.pi)
    assert abs(center_pixel(oracle_render(scenes.floor_scene(light), 65, 65, 64)[0], 65) - 1.0) < 0.01


def test_smooth_dielectric_reflects_normal_incidence_fresnel():
    # pathtrace.rs:563-577: black mirror base: F0 = ((1.5-1)/(1.5+1))^2 = 0.04
    m = sc.BaseBsdf(base_color=(0.0, 0.0, 0.0), roughness=0.0)
    img, _ = oracle_render(scenes.sky_scene(scenes.plane_mesh(5.0, 0.0), m), 32, 32, 64)
    assert abs(center_mean(img, 32, 16) - 0.04) < 0.003


@pytest.mark.parametrize("roughness,lo", [(0.0, 0.99), (0.3, 0.96), (0.5, 0.9)])
def test_metal_furnace_conserves_energy(roughness, lo):
    # pathtrace.rs:580-597
    m = sc.BaseBsdf(base_color=(1.0, 1.0, 1.0), metallic=1.0, roughness=roughness)
    img, _ = oracle_render(scenes.sky_scene(scenes.cube_mesh(0.5), m), 32, 32, 64)
    got = center_mean(img, 32, 16)
    assert lo < got < 1.005


def test_clearcoat_adds_a_second_fresnel_layer():
    # pathtrace.rs:600-616: 0.04 + 0.96 * 0.04 = 0.0784
    m = sc.BaseBsdf(base_color=(0.0, 0.0, 0.0), roughness=0.0, clearcoat=1.0)
    img, _ = oracle_render(scenes.sky_scene(scenes.plane_mesh(5.0, 0.0), m), 32, 32, 128)
    assert abs(center_mean(img, 32, 16) - 0.0784) < 0.004


@pytest.mark.parametrize("coat_roughness", [0.0, 0.3])
def test_clearcoat_furnace_conserves_energy(coat_roughness):
    m = sc.BaseBsdf(base_color=(1.0, 1.0, 1.0), roughness=0.5, clearcoat=1.0, clearcoat_roughness=coat_roughness)
    img, _ = oracle_render(scenes.sky_scene(scenes.cube_mesh(0.5), m), 32, 32, 64)
    assert 0.95 < center_mean(img, 32, 16) < 1.01


def test_sheen_adds_its_directional_albedo():
    # pathtrace.rs:643-661: 0.137 + (1 - 0.137) * 0.04
    m = sc.BaseBsdf(base_color=(0.0, 0.0, 0.0), roughness=0.0, sheen_color=(1.0, 1.0, 1.0), sheen_roughness=1.0)
    img, _ = oracle_render(scenes.sky_scene(scenes.plane_mesh(5.0, 0.0), m), 32, 32, 128)
    want = 0.137 + (1.0 - 0.137) * 0.04
    assert abs(center_mean(img, 32, 16) - want) < 0.01


@pytest.mark.parametrize("sheen_roughness", [0.3, 1.0])
def test_sheen_furnace_conserves_energy(sheen_roughness):
    m = sc.BaseBsdf(base_color=(1.0, 1.0, 1.0), roughness=0.5, sheen_color=(1.0, 1.0, 1.0), sheen_roughness=sheen_roughness)
    img, _ = oracle_render(scenes.sky_scene(scenes.cube_mesh(0.5), m), 32, 32, 64)
    assert 0.95 < center_mean(img, 32, 16) < 1.02


def _sphere_furnace(base, multiscatter=True, spp=1024):
    # pathtrace.rs:1182-1191: a level-2 icosphere (flat normals) of radius 1, camera at z = 4
    ico = scenes.icosphere(2)
    mesh = sc.Mesh(ico.positions, ico.indices)
    img, _ = oracle_render(scenes.sky_scene(mesh, base, 4.0), 32, 32, spp, multiscatter=multiscatter)
    return center_mean(img, 32, 8)


def _dielectric(transmission, roughness):
    return sc.BaseBsdf(transmission=transmission, roughness=roughness, base_color=(1.0, 1.0, 1.0))


@pytest.mark.parametrize("base,lo", [(_dielectric(1.0, 0.0), 0.99), (_dielectric(0.0, 0.1), 0.99),
                                     (_dielectric(0.0, 0.5), 0.99)])
def test_smooth_dielectrics_conserve_energy_in_a_furnace(base, lo):
    # pathtrace.rs:1202-1213
    assert lo <= _sphere_furnace(base) <= 1.005


@pytest.mark.parametrize("roughness,lo", [(0.3, 0.95), (0.6, 0.70)])
def test_rough_glass_without_compensation_loses_but_never_gains(roughness, lo):
    # pathtrace.rs:1217-1224
    assert lo <= _sphere_furnace(_dielectric(1.0, roughness), multiscatter=False) <= 1.005


@pytest.mark.parametrize("roughness", [0.3, 0.5, 0.7, 1.0])
def test_compensation_restores_the_rough_glass_furnace(roughness):
    # pathtrace.rs:1228-1236
    assert 0.97 <= _sphere_furnace(_dielectric(1.0, roughness)) <= 1.01


def test_multiscatter_compensation_restores_rough_metal():
    m = sc.BaseBsdf(base_color=(1.0, 1.0, 1.0), metallic=1.0, roughness=1.0)
    on = center_mean(oracle_render(scenes.sky_scene(scenes.cube_mesh(0.5), m), 32, 32, 64)[0], 32, 16)
    off = center_mean(oracle_render(scenes.sky_scene(scenes.cube_mesh(0.5), m), 32, 32, 64, multiscatter=False)[0], 32, 16)
    assert off < 0.9 < 0.97 < on < 1.03


def test_lighting_parts_sum_to_the_beauty():
    # pathtrace.rs:1346-1384: the five radiance parts sum to the beauty image
    s = scenes.cornell()
    s.add_light(sc.PointLight(position=(0.5, 0.8, 0.6), intensity=0.3))
    _, st = oracle_render(s, 32, 32, 8, lighting=True)
    L = st["lighting"].reshape(-1, 65)
    parts = sum(L[:, 3 * k:3 * k + 3].astype(np.float64) for k in range(5))
    film = st["film"].reshape(-1, 3)
    assert np.allclose(parts, film, rtol=1e-4, atol=1e-4)
    vis = L[:, 15] / np.maximum(L[:, 16], 1e-30)
    assert np.all((vis >= 0) & (vis <= 1.0 + 1e-6))


def test_cornell_is_plausible():
    img, _ = oracle_render(scenes.cornell(), 32, 32, 16)
    assert np.all(np.isfinite(img)) and img.mean() > 0.1
