"""Host packing and readback: the reference's own unit tests (crates/verion-gpu/src/pack.rs:892-1011,
readback.rs:139-166, lib.rs:256-269) restated on the Python restatements, plus the scene generators."""
import math

import numpy as np

from verion_amd import pack, readback, scenes
from verion_amd import scene as sc
from verion_amd.backend import Adaptive


def test_albedo_tables_follow_the_light_table_and_precede_the_masks():
    s = sc.Scene()
    s.add_light(sc.PointLight(position=(0.0, 0.0, 0.0)))
    words, count, _, _ = pack.pack_lights(s)
    at = count * pack.LIGHT_WORDS
    assert pack.albedo_tables().size == pack.ALBEDO_WORDS
    assert words.size == at + pack.ALBEDO_WORDS
    assert np.array(words[at], np.uint32).view(np.float32) == 1.0


def _env_scene(f):
    w, h = 64, 32
    px = np.array([[f(i, j)] * 3 for j in range(h) for i in range(w)], np.float32)
    s = sc.Scene()
    s.add_light(sc.EnvironmentMap(s.add_image(sc.Image(w, h, px)), 1.0))
    return s


def test_virtual_light_follows_the_brightest_patch():
    s = _env_scene(lambda i, j: 50.0 if 0 <= i < 4 and 8 <= j < 12 else 0.1)
    d, e = pack.virtual_light(s)
    phi = ((2.0 + 0.5) / 64.0 - 0.5) * 2.0 * math.pi
    theta = 10.0 / 32.0 * math.pi
    want = [math.sin(theta) * math.sin(phi), math.cos(theta), -math.sin(theta) * math.cos(phi)]
    assert float(np.dot(d, want)) > 0.99
    assert e[0] > 0.1 and e[0] == e[1]
    assert pack.virtual_light(_env_scene(lambda i, j: 1.0)) is None
    off = _env_scene(lambda i, j: 50.0 if 0 <= i < 4 and 8 <= j < 12 else 0.1)
    off.virtual_light = False
    assert pack.virtual_light(off) is None


def test_mip_chain_ends_at_one_texel_and_averages():
    img = sc.Image(2, 2, np.array([[1.0] * 3, [3.0] * 3, [5.0] * 3, [7.0] * 3], np.float32))
    texels, levels = pack.mip_chain(img)
    assert levels == 2 and texels.size == (4 + 1) * 3
    assert np.all(texels[12:] == 4.0)


def test_base_params_carry_the_view_and_leave_dispatch_words_clear():
    info = pack.SceneInfo(2, 5, 0, 0, 0.0, (0.1, 0.2, 0.3), False, False)
    view = pack.View.of(sc.Camera())
    p = pack.base_params(info, view, 64, 32, True, False, (0.02, 16), (True, False, 4, 4))
    assert (p[0], p[1], p[5], p[23]) == (64, 32, 2, 5)
    assert (p[28], p[29], p[32], p[34]) == (1, 0, 1, 16)
    assert (p[6], p[30], p[31], p[35]) == (0, 0, 0, 0)
    assert (p[36], p[37], p[38]) == (1, 4, 4)
    p = pack.base_params(info, view, 64, 32, False, False, None, (True, True, 4, 4))
    assert p[36] == 3
    assert np.array(p[21], np.uint32).view(np.float32) == np.float32(0.2)
    assert p[7] == pack.MAX_DEPTH == 8
    assert np.array(p[2], np.uint32).view(np.float32) == np.float32(math.tan(np.float32(0.8) * np.float32(0.5)))


def test_emitter_cdf_and_selection_pdf():
    p = pack.pack(scenes.cornell())
    assert p.info.emitter_count == 2
    rec = p.emitters.reshape(-1, 16)
    assert np.allclose(rec[:, 12], 0.125)  # two halves of the 0.25 m^2 light
    assert p.emit_cdf[-1] == 1.0 and np.all(np.diff(p.emit_cdf) >= 0)
    assert abs(rec[:, 13].sum() - 1.0) < 1e-6
    inst = p.instances.reshape(-1, 16)
    assert inst[-1, 3] == 1 and np.all(inst[:-1, 3] == 0)  # emitter base + 1 on the light only


def test_env_cdfs_are_monotone_and_end_at_one():
    p = pack.pack(scenes.procedural(spheres=4))
    e = p.env
    assert e.width == 32 and e.height == 16
    assert e.marginal[-1] == 1.0 and np.all(np.diff(e.marginal) >= 0)
    cond = e.conditional.reshape(16, 32)
    assert np.all(cond[:, -1] == 1.0) and np.all(np.diff(cond, axis=1) >= 0)
    # a constant map: pdf_scale * lum = 1 / (4 pi) times the sin-weighted normalisation
    assert e.pdf_scale > 0


def test_scene_generators():
    c = pack.pack(scenes.cornell())
    assert c.triangle_count == 36
    pr = pack.pack(scenes.procedural())
    assert pr.triangle_count == 1024 * 1280 == 1310720
    ico = scenes.icosphere(3)
    assert len(ico.positions) == 642 and ico.indices.size == 3 * 1280
    assert np.allclose(np.linalg.norm(ico.positions, axis=1), 1.0, atol=1e-6)


def test_readback_means_divide_by_each_pixels_own_count():
    stats = np.array([4.0, 0.0, 0.0, 2.0, 0.0, 0.0])
    inv = readback.inverse_samples(stats, 2, 99)
    assert inv.tolist() == [0.25, 0.5]
    assert readback.mean(np.array([4.0, 8.0, 12.0, 2.0, 2.0, 2.0]), inv).tolist() == [[1.0, 2.0, 3.0], [1.0, 1.0, 1.0]]
    assert readback.inverse_samples(None, 2, 4).tolist() == [0.25, 0.25]
    assert readback.sample_counts(stats).tolist() == [4, 2]


def test_lighting_reads_parts_and_defaults_empty_slots():
    rec = np.zeros(65, np.float32)
    rec[3:6] = [2.0, 4.0, 6.0]
    l = readback.lighting(rec, np.array([0.5]))
    assert l.indirect_diffuse.tolist() == [[1.0, 2.0, 3.0]]
    assert l.visibility.tolist() == [1.0]
    assert np.all(l.light_response == 0.0)


def test_adaptive_stops_after_the_minimum_once_the_error_is_small():
    a = Adaptive(threshold=0.1, min_spp=4)
    assert a.active(3.0, 3.0, 3.0)
    assert not a.active(4.0, 4.0, 4.0)
    assert a.active(4.0, 4.0, 8.0)
    assert not a.active(16.0, 0.0, 0.0)
