"""IES profiles on the GPU against the oracle (iesFactor, pt_core.slang:199-216, applied to point
and spot lights at :1252-1253): a floor and a box under a point light with an asymmetric profile in
a rotated frame and a spot light with a narrow downlight profile, both integrators, with and
without the lighting split; films within 1e-4 relative of the oracle's."""
import math

import numpy as np
import pytest

import oracle
from helpers import params
from test_ies import DOWNLIGHT, asymmetric_profile
from verion_amd import ies, scenes
from verion_amd import scene as sc
from verion_amd.backend import MEGAKERNEL, WAVEFRONT, Renderer

pytestmark = pytest.mark.gpu
TOL = 1.0e-4


def ies_scene():
    s = sc.Scene()
    floor = s.add_mesh(scenes.plane_mesh(4.0, 0.0))
    s.add_instance(sc.Instance(floor, s.add_material(sc.lambert((0.8, 0.8, 0.8)))))
    box = s.add_mesh(scenes.cube_mesh(0.3))
    s.add_instance(sc.Instance(box, s.add_material(sc.BaseBsdf(base_color=(0.7, 0.3, 0.2), roughness=0.4)),
                               ((1.0, 0.0, 0.0, 0.6), (0.0, 1.0, 0.0, -0.4), (0.0, 0.0, 1.0, 0.3))))
    c, si = math.cos(0.5), math.sin(0.5)
    tilted = ((c, 0.0, -si), (0.0, 1.0, 0.0), (si, 0.0, c))  # rows x, y, z: -z tipped toward +x
    s.add_light(sc.PointLight(position=(-0.5, 0.2, 1.6), color=(1.0, 0.9, 0.8), intensity=12.0,
                              ies=(asymmetric_profile(), tilted)))
    s.add_light(sc.SpotLight(position=(0.8, 0.8, 2.0), direction=(-0.2, -0.3, -1.0), intensity=20.0, outer_cone=0.6,
                             ies=(ies.IesProfile.parse(DOWNLIGHT), ((1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0)))))
    s.add_light(sc.Environment((0.05, 0.06, 0.08)))
    s.camera = sc.Camera(transform=((1.0, 0.0, 0.0, 0.0), (0.0, 0.8, -0.6, -3.0), (0.0, 0.6, 0.8, 3.2)))
    return s


def rel(got, want):
    got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
    return float((np.abs(got - want) / np.maximum(np.abs(want), 1.0e-6)).max())


@pytest.fixture(scope="module")
def scene(ctx):
    r = Renderer(ctx, ies_scene())
    return r, oracle.OracleScene(r.packed)


@pytest.mark.parametrize("integrator", [MEGAKERNEL, WAVEFRONT])
def test_ies_lights_match_the_oracle(scene, integrator):
    r, o = scene
    r.set_integrator(integrator)
    acc = r.new_accumulation(80, 60)
    r.accumulate(acc, 8)
    want = o.render(params(r.packed, 80, 60), 8)["film"]
    got = acc.film()
    assert rel(got, want) <= TOL, f"exact {np.mean(got == want):.4f}"
    # the profiles shape the image: not a uniform scale of the bare lights
    assert np.ptp(got) > 0


@pytest.mark.parametrize("integrator", [MEGAKERNEL, WAVEFRONT])
def test_ies_lights_in_the_lighting_split(scene, integrator):
    r, o = scene
    r.set_integrator(integrator)
    acc = r.new_accumulation(48, 36, lighting=True)
    r.accumulate(acc, 4)
    want = o.render(params(r.packed, 48, 36, lighting=True), 4)
    assert rel(acc.film(), want["film"]) <= TOL
    assert rel(acc.lighting_sums(), want["lighting"]) <= TOL
