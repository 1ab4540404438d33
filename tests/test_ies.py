"""IES profiles (crates/verion-scene/src/ies.rs; packed after the link masks, pack.rs:628-629,
679-686; applied to point and spot lights by iesFactor, pt_core.slang:199-216, 1252-1253).

- the parser: ies.rs's own unit tests (:164-222) restated;
- the packer: the profile's frame and table sit past the link masks, word 15 = offset + 1;
- the oracle: an isotropic profile draws the bare light bit for bit; a floor lit by a point light
  with an asymmetric profile matches, sample by sample, the closed form rho/pi I f(theta, phi)
  cos / r^2 at the sample's own floor point (f from the profile's bilinear lookup); and the
  reference's GPU test of a downlight and light linking (pathtrace.rs:1740-1790) on the oracle.
"""
import math

import numpy as np
import pytest

import oracle
from helpers import center_pixel, oracle_render, params
from verion_amd import ies, pack
from verion_amd import scene as sc
from verion_amd import scenes

DOWNLIGHT = "IESNA:LM-63-2002\n[TEST] x\nTILT=NONE\n1 1000 1 4 1 1 2 0 0 0\n1 1 10\n0 30 60 90\n0\n100 80 40 0\n"
IDENTITY_FRAME = ((1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0))


def test_symmetric_downlight_is_normalised_and_round():
    # ies.rs:168-184
    p = ies.IesProfile.parse(DOWNLIGHT)
    assert p.table.shape == (ies.HORIZONTAL, ies.VERTICAL)
    assert abs(p.at(0.0, 0.0) - 1.0) < 1e-6
    assert abs(p.at(30.0, 123.0) - 0.8) < 1e-6  # one plane is the same all around
    assert abs(p.at(15.0, 0.0) - 0.9) < 1e-6    # linear between samples
    assert p.at(120.0, 0.0) == 0.0              # nothing above the measured range
    assert p.peak == 100.0
    assert abs(p.scaled(10.0).at(0.0, 0.0) - 10.0) < 1e-5


def test_quadrant_symmetry_mirrors_the_planes():
    # ies.rs:187-196
    s = "IESNA:LM-63-2002\nTILT=NONE\n1 1000 1 2 2 1 2 0 0 0\n1 1 10\n0 90\n0 90\n100 100 50 50\n"
    p = ies.IesProfile.parse(s)
    assert abs(p.at(45.0, 0.0) - 1.0) < 1e-6
    assert abs(p.at(45.0, 90.0) - 0.5) < 1e-6
    assert abs(p.at(45.0, 180.0) - 1.0) < 1e-6  # mirrored to 0
    assert abs(p.at(45.0, 270.0) - 0.5) < 1e-6  # mirrored to 90


def test_bad_files_are_errors():
    # ies.rs:198-214
    with pytest.raises(ies.IesError) as e:
        ies.IesProfile.parse("nothing")
    assert e.value.kind == "truncated"
    with pytest.raises(ies.IesError) as e:
        ies.IesProfile.parse(DOWNLIGHT.replace("4 1 1 2", "4 1 2 2"))
    assert e.value.kind == "photometric" and e.value.detail == "2"
    with pytest.raises(ies.IesError) as e:
        ies.IesProfile.parse(DOWNLIGHT.replace("80", "x"))
    assert e.value.kind == "number"
    with pytest.raises(ies.IesError) as e:
        ies.IesProfile.parse(DOWNLIGHT.replace("100 80 40 0", "0 0 0 0"))
    assert e.value.kind == "empty"


def test_bilateral_and_full_symmetries_and_tilt_include():
    # planes up to 180 mirror 180..360 onto 180..0; a full set (to 360) is read as given; TILT=INCLUDE
    # skips its block of lamp geometry, angles and multipliers
    bil = "IESNA:LM-63-2002\nTILT=NONE\n1 1000 1 2 3 1 2 0 0 0\n1 1 10\n0 90\n0 90 180\n100 100 50 50 20 20\n"
    p = ies.IesProfile.parse(bil)
    assert abs(p.at(10.0, 180.0) - 0.2) < 1e-6 and abs(p.at(10.0, 270.0) - 0.5) < 1e-6
    assert abs(p.at(10.0, 315.0) - 0.75) < 1e-6  # 315 -> 45: halfway between 1 and 0.5
    full = ("IESNA:LM-63-2002\nTILT=INCLUDE\n1\n2\n0 90\n1 1\n1 1000 2 2 4 1 2 0 0 0\n1 1 10\n0 90\n0 90 180 360\n"
            "10 10 20 20 40 40 10 10\n")
    q = ies.IesProfile.parse(full)
    assert q.peak == 80.0  # candela times the multiplier 2
    assert abs(q.at(0.0, 0.0) - 0.25) < 1e-6 and abs(q.at(0.0, 180.0) - 1.0) < 1e-6
    assert abs(q.at(0.0, 270.0) - 0.625) < 1e-6  # between 180 (1.0) and 360 (0.25)


def asymmetric_profile():
    """Full horizontal set, each plane its own level and slope over the vertical angle."""
    rows = {0: (100, 90, 60, 20), 90: (25, 25, 20, 10), 180: (50, 45, 35, 15), 270: (75, 60, 40, 30), 360: (100, 90, 60, 20)}
    h = sorted(rows)
    text = ("IESNA:LM-63-2002\nTILT=NONE\n1 1000 1 4 5 1 2 0 0 0\n1 1 10\n0 30 60 90\n" + " ".join(map(str, h)) + "\n"
            + "\n".join(" ".join(map(str, rows[k])) for k in h) + "\n")
    return ies.IesProfile.parse(text)


def test_packer_places_frame_and_table_after_the_link_masks():
    prof = asymmetric_profile()
    frame = ((0.0, 1.0, 0.0), (-1.0, 0.0, 0.0), (0.0, 0.0, 1.0))
    s = scenes.floor_scene(sc.PointLight(position=(0.0, 0.0, 1.0), intensity=4.0, ies=(prof, frame)))
    s.add_light(sc.DirectionalLight(direction=(0.0, 0.0, -1.0), ies=(prof, frame)))  # ignored for directional lights
    s.add_light(sc.SpotLight(position=(1.0, 0.0, 1.0), direction=(0.0, 0.0, -1.0), ies=(ies.isotropic(), IDENTITY_FRAME)))
    p = pack.pack(s)
    L = p.lights
    base = p.info.light_count * 16 + pack.ALBEDO_WORDS + 2 * len(s.instances)
    po = int(L[15]) - 1
    assert po == base
    np.testing.assert_array_equal(L[po:po + 9].view(np.float32), np.asarray(frame, np.float32).reshape(-1))
    np.testing.assert_array_equal(L[po + 9:po + 9 + 361 * 181].view(np.float32), prof.table.reshape(-1))
    assert L[16 + 15] == 0  # the directional light takes no profile (pack.rs:628)
    po2 = int(L[32 + 15]) - 1
    assert po2 == po + 9 + 361 * 181 and L.size == po2 + 9 + 361 * 181
    assert np.all(L[po2 + 9:].view(np.float32) == 1.0)


def test_isotropic_profile_draws_the_bare_light_bit_for_bit():
    light = dict(position=(0.3, -0.2, 1.0), intensity=5.0)
    bare, _ = oracle_render(scenes.floor_scene(sc.PointLight(**light)), 33, 33, 8)
    prof, _ = oracle_render(scenes.floor_scene(sc.PointLight(**light, ies=(ies.isotropic(), IDENTITY_FRAME))), 33, 33, 8)
    np.testing.assert_array_equal(bare, prof)
    spot = dict(position=(0.3, -0.2, 1.0), direction=(0.1, 0.0, -1.0), intensity=5.0, outer_cone=0.9)
    bare, _ = oracle_render(scenes.floor_scene(sc.SpotLight(**spot)), 33, 33, 8)
    prof, _ = oracle_render(scenes.floor_scene(sc.SpotLight(**spot, ies=(ies.isotropic(), IDENTITY_FRAME))), 33, 33, 8)
    np.testing.assert_array_equal(bare, prof)


@pytest.mark.parametrize("frame", [IDENTITY_FRAME, ((0.0, 1.0, 0.0), (-1.0, 0.0, 0.0), (0.0, 0.0, 1.0))])
def test_asymmetric_profile_matches_the_closed_form_per_sample(frame):
    # Lambert floor (albedo 1) at z = 0, point light 1 m above the origin, black sky: a sample's
    # radiance is its floor point's direct light, 1/pi I f(theta, phi) cos / r^2, with
    # (theta, phi) the direction from the light in the light's frame (vertical angle from -z) and
    # the floor point lifted by the shader's 1e-4 offset (pt_core.slang:1221-1222).
    prof = asymmetric_profile()
    I = 7.0
    s = scenes.floor_scene(sc.PointLight(position=(0.0, 0.0, 1.0), intensity=I, ies=(prof, frame)))
    p = pack.pack(s)
    o = oracle.OracleScene(p)
    w = h = 33
    par = params(p, w, h)
    F = np.asarray(frame, np.float64)
    for sample in (0, 5):
        st = o.render(par, 1, first_sample=sample)
        got = st["film"].reshape(-1, 3)[:, 0]
        hits, rays = o.primary_hits(par, sample, rays=True)
        t = hits[:, 3].view(np.float32).astype(np.float64)
        x = rays[:, :3].astype(np.float64) + t[:, None] * rays[:, 3:].astype(np.float64)
        x[:, 2] += 1.0e-4  # NEE leaves from the hit point offset along the facing normal (+z)
        d = x - np.array([0.0, 0.0, 1.0])
        r = np.linalg.norm(d, axis=1)
        l = (d / r[:, None]) @ F.T
        theta = np.degrees(np.arccos(np.clip(-l[:, 2], -1.0, 1.0)))
        phi = np.degrees(np.arctan2(l[:, 1], l[:, 0])) % 360.0
        f = np.array([prof.at(a, b) for a, b in zip(theta, phi)])
        cos = -d[:, 2] / r
        want = I * f * cos / (r * r) / math.pi
        assert np.all(hits[:, 0] == 1)
        np.testing.assert_allclose(got, want, rtol=2e-5, atol=1e-7)
        # the profile shapes the image: the same samples without it differ by the factor f
        assert np.ptp(f) > 0.3


def _downlight_at(ies_on, mask, x):
    # pathtrace.rs:1740-1790: a point light 1 m above the floor point under the camera, a narrow
    # downlight (0 at 40 degrees); the camera moves along x; mask = the floor's light group
    profile = ies.IesProfile.parse("IESNA:LM-63-2002\nTILT=NONE\n1 1000 1 4 1 1 2 0 0 0\n1 1 10\n0 20 40 90\n0\n100 50 0 0\n")
    light = sc.PointLight(position=(0.0, 0.0, 1.0), intensity=40.0, link_mask=1,
                          ies=(profile, IDENTITY_FRAME) if ies_on else None)
    s = scenes.floor_scene(light)
    s.set_instance_light_mask(0, mask)
    s.camera = sc.Camera(transform=((1.0, 0.0, 0.0, x), (0.0, 1.0, 0.0, 0.0), (0.0, 0.0, 1.0, 3.0)))
    img, _ = oracle_render(s, 33, 33, 32)
    return center_pixel(img, 33)


def test_reference_downlight_and_light_linking_on_the_oracle():
    plain, with_ = _downlight_at(False, 1, 0.0), _downlight_at(True, 1, 0.0)
    assert abs(plain - with_) < 0.08 * plain  # straight below: the profile's peak
    off_plain, off = _downlight_at(False, 1, 0.84), _downlight_at(True, 1, 0.84)
    assert off_plain > 0.05 and off < 0.05 * off_plain  # 40 degrees off the axis: the profile says 0
    assert _downlight_at(False, 2, 0.0) < 0.01 * plain  # the floor is not in the light's group
