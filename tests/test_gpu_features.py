"""GPU parity beyond the plain image: every shading feature, adaptive sampling, the lighting split,
ReSTIR DI, seeds, progressive accumulation and the error paths of the backend contract.
Oracle comparisons hold to 1e-4 relative per value (the build reproduces the oracle's bits);
integrator comparisons are bit-exact (crates/verion-gpu-metal/tests/pathtrace.rs:151-230,
restir.rs:111-137, tests/gpu.rs:52-88)."""
import numpy as np
import pytest

import oracle
from helpers import params
from verion_amd import scenes
from verion_amd.backend import (INVALID_ARGUMENT, MEGAKERNEL, SCENE, WAVEFRONT, Adaptive, BackendError,
                                Renderer, Restir)

pytestmark = pytest.mark.gpu
TOL = 1.0e-4


def rel(got, want):
    got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
    return float((np.abs(got - want) / np.maximum(np.abs(want), 1.0e-6)).max())


@pytest.fixture(scope="module")
def zoo(ctx):
    r = Renderer(ctx, scenes.zoo())
    return r, oracle.OracleScene(r.packed)


@pytest.mark.parametrize("integrator", [MEGAKERNEL, WAVEFRONT])
def test_zoo_matches_oracle(zoo, integrator):
    r, o = zoo
    r.set_integrator(integrator)
    acc = r.new_accumulation(96, 64)
    r.accumulate(acc, 8)
    want = o.render(params(r.packed, 96, 64), 8)["film"]
    got = acc.film()
    assert rel(got, want) <= TOL, f"exact {np.mean(got == want):.4f}"
    r.set_integrator(MEGAKERNEL)


@pytest.mark.parametrize("integrator", [MEGAKERNEL, WAVEFRONT])
def test_zoo_without_multiscatter(zoo, integrator):
    r, o = zoo
    r.set_integrator(integrator)
    r.set_multiscatter(False)
    acc = r.new_accumulation(48, 32)
    r.accumulate(acc, 4)
    want = o.render(params(r.packed, 48, 32, multiscatter=False), 4)["film"]
    r.set_multiscatter(True)
    r.set_integrator(MEGAKERNEL)
    assert rel(acc.film(), want) <= TOL


@pytest.mark.parametrize("integrator", [MEGAKERNEL, WAVEFRONT])
def test_lighting_split_matches_oracle(zoo, integrator):
    r, o = zoo
    r.set_integrator(integrator)
    acc = r.new_accumulation(64, 48, lighting=True)
    r.accumulate(acc, 4)
    st = o.render(params(r.packed, 64, 48, lighting=True), 4)
    r.set_integrator(MEGAKERNEL)
    assert rel(acc.film(), st["film"]) <= TOL
    got = acc.lighting_sums()
    assert np.abs(got - st["lighting"]).max() <= TOL * np.maximum(np.abs(st["lighting"]), 1.0).max()


def test_lighting_split_integrators_agree_bit_for_bit(zoo):
    r, _ = zoo
    out = []
    for integ in (MEGAKERNEL, WAVEFRONT):
        r.set_integrator(integ)
        acc = r.new_accumulation(64, 48, lighting=True)
        r.accumulate(acc, 3)
        out.append((acc.film(), acc.lighting_sums()))
    r.set_integrator(MEGAKERNEL)
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("integrator", [MEGAKERNEL, WAVEFRONT])
def test_adaptive_matches_oracle_with_counts(ctx, integrator):
    r = Renderer(ctx, scenes.furnace_room(0.5, 1.0))
    r.set_integrator(integrator)
    ad = Adaptive(threshold=0.05, min_spp=4)
    r.set_adaptive(ad)
    acc = r.new_accumulation(16, 16)
    r.accumulate(acc, 32)
    st = oracle.OracleScene(r.packed).render(params(r.packed, 16, 16, adaptive=(0.05, 4)), 32)
    assert np.array_equal(acc.stats(), st["stats"])
    assert rel(acc.film(), st["film"]) <= TOL
    counts = acc.sample_counts()
    assert (counts < 32).any() and (counts == 32).any()


def test_adaptive_gives_the_same_image_with_both_integrators(ctx):
    r = Renderer(ctx, scenes.furnace_room(0.5, 1.0))
    r.set_adaptive(Adaptive(threshold=0.05, min_spp=4))
    accs = []
    for integ in (MEGAKERNEL, WAVEFRONT):
        r.set_integrator(integ)
        a = r.new_accumulation(16, 16)
        r.accumulate(a, 32)
        accs.append(a)
    assert np.array_equal(accs[0].mean(), accs[1].mean())
    assert np.array_equal(accs[0].sample_counts(), accs[1].sample_counts())


def test_adaptive_with_a_zero_threshold_is_uniform(ctx):
    r = Renderer(ctx, scenes.furnace_room(0.5, 1.0))
    r.set_integrator(WAVEFRONT)
    uniform = r.render(16, 16, 8)
    r.set_adaptive(Adaptive(threshold=0.0, min_spp=2))
    acc = r.new_accumulation(16, 16)
    r.accumulate(acc, 8)
    assert np.array_equal(acc.mean(), uniform)
    assert acc.sample_counts().tolist() == [8] * 256
    assert acc.active_pixels() == 256


def test_adaptive_stops_converged_pixels_at_the_minimum(ctx):
    r = Renderer(ctx, scenes.furnace_room(0.5, 1.0))
    r.set_adaptive(Adaptive(threshold=1e9, min_spp=4))
    acc = r.new_accumulation(16, 16)
    r.accumulate(acc, 8)
    assert acc.sample_counts().tolist() == [4] * 256
    assert acc.active_pixels() == 0
    assert abs(acc.mean()[:, 0].mean() - 1.99) < 0.2
    r.set_adaptive(None)
    with pytest.raises(BackendError) as e:
        r.accumulate(acc, 1)
    assert e.value.status == SCENE


@pytest.fixture(scope="module")
def many(ctx):
    s = scenes.many_lights(emitters=600, clutter_boxes=400)
    r = Renderer(ctx, s)
    return r, oracle.OracleScene(r.packed)


@pytest.mark.parametrize("integrator", [MEGAKERNEL, WAVEFRONT])
def test_restir_di_matches_oracle(many, integrator):
    r, o = many
    r.set_integrator(integrator)
    r.set_restir(Restir(di=True, gi=False, candidates=8, history=4))
    acc = r.new_accumulation(64, 48)
    for _ in range(4):  # frames: each reuses the previous frame's reservoirs
        r.accumulate(acc, 1)
    st = o.render(params(r.packed, 64, 48, restir=(True, False, 8, 4)), 4)
    r.set_restir(None)
    r.set_integrator(MEGAKERNEL)
    assert rel(acc.film(), st["film"]) <= TOL, f"exact {np.mean(acc.film() == st['film']):.4f}"


@pytest.mark.parametrize("integrator", [MEGAKERNEL, WAVEFRONT])
def test_restir_di_with_the_lighting_split_matches_oracle(many, integrator):
    # the lighting split under ReSTIR DI (pt_core.slang:1306-1332: the reservoir's light is tracked
    # in the first-hit slots); GI stays off with the split (:1296), even when asked for
    r, o = many
    r.set_integrator(integrator)
    r.set_restir(Restir(di=True, gi=True, candidates=8, history=4))
    acc = r.new_accumulation(64, 48, lighting=True)
    for _ in range(3):
        r.accumulate(acc, 1)
    st = o.render(params(r.packed, 64, 48, lighting=True, restir=(True, True, 8, 4)), 3)
    r.set_restir(None)
    r.set_integrator(MEGAKERNEL)
    assert rel(acc.film(), st["film"]) <= TOL, f"exact {np.mean(acc.film() == st['film']):.4f}"
    got = acc.lighting_sums()
    assert np.abs(got - st["lighting"]).max() <= TOL * np.maximum(np.abs(st["lighting"]), 1.0).max()


def test_restir_integrators_agree(many):
    r, _ = many
    r.set_restir(Restir(candidates=8))
    out = []
    for integ in (MEGAKERNEL, WAVEFRONT):
        r.set_integrator(integ)
        a = r.new_accumulation(64, 48)
        r.accumulate(a, 3)
        out.append(a.film())
    r.set_restir(None)
    r.set_integrator(MEGAKERNEL)
    assert np.array_equal(out[0], out[1])


def test_restir_switch_mid_accumulation_is_an_error(many):
    r, _ = many
    acc = r.new_accumulation(8, 8)
    r.accumulate(acc, 1)
    r.set_restir(Restir())
    with pytest.raises(BackendError) as e:
        r.accumulate(acc, 1)
    r.set_restir(None)
    assert e.value.status == SCENE


def test_progressive_accumulation_equals_one_shot(ctx):
    # pathtrace.rs:151-161
    r = Renderer(ctx, scenes.cornell())
    a = r.new_accumulation(32, 32)
    r.accumulate(a, 8)
    r.accumulate(a, 8)
    b = r.new_accumulation(32, 32)
    r.accumulate(b, 16)
    assert a.samples() == 16 and a.gpu_seconds() > 0.0
    assert np.array_equal(a.film(), b.film())


def test_seed_zero_is_the_reference_stream_and_seeds_differ(ctx):
    r = Renderer(ctx, scenes.cornell())
    o = oracle.OracleScene(r.packed)
    r.set_seed(1)
    acc = r.new_accumulation(32, 32)
    r.accumulate(acc, 2)
    assert rel(acc.film(), o.render(params(r.packed, 32, 32), 2, seed=1)["film"]) <= TOL
    r.set_seed(0)
    acc0 = r.new_accumulation(32, 32)
    r.accumulate(acc0, 2)
    assert not np.array_equal(acc.film(), acc0.film())
    assert rel(acc0.film(), o.render(params(r.packed, 32, 32), 2)["film"]) <= TOL


def test_error_paths(ctx):
    r = Renderer(ctx, scenes.cornell())
    with pytest.raises(BackendError) as e:
        r.new_accumulation(0, 16)
    assert e.value.status == INVALID_ARGUMENT
    from verion_amd import scene as sc

    empty = sc.Scene()
    with pytest.raises((BackendError, ValueError)):
        Renderer(ctx, empty)
    with pytest.raises(BackendError) as e:
        r.set_integrator(7)
    assert e.value.status == INVALID_ARGUMENT


def test_wavefront_memory_is_reported(ctx):
    r = Renderer(ctx, scenes.cornell())
    r.set_integrator(WAVEFRONT)
    acc = r.new_accumulation(64, 32)
    assert acc.wavefront_bytes() == 0
    r.accumulate(acc, 1)
    n = 64 * 32
    # state ping-pong, hits, depth-0 queue, counters, and the deferred shadow rays: per slot K
    # records of 10 words + visibility + queue entry (K = delta lights + 1), and 6 pending words
    k = r.packed.info.light_count + 1
    assert acc.wavefront_bytes() == 2 * 4 * 17 * n + 16 * n + 4 * n + 128 + 4 * 12 * k * n + 4 * 6 * n
    lit = r.new_accumulation(64, 32, lighting=True)  # the lighting split shades fused
    r.accumulate(lit, 1)
    assert lit.wavefront_bytes() == 2 * 4 * 34 * n + 16 * n + 4 * n + 128
    m = ctx.memory()
    assert m.budget > m.allocated > 0 and m.uploaded > 0


@pytest.mark.parametrize("adaptive", [False, True])
def test_depth0_order_and_grid_do_not_change_the_image(ctx, adaptive):
    from verion_amd.backend import Adaptive

    r = Renderer(ctx, scenes.zoo())
    r.set_integrator(WAVEFRONT)
    if adaptive:
        r.set_adaptive(Adaptive(0.05, 2))
    films = []
    for blocks, tw, th in [(2048, 0, 0), (2048, 8, 8), (512, 16, 4), (4096, 5, 3), (64, 64, 64)]:
        r.set_wavefront_blocks(blocks)
        r.set_wavefront_tiles(tw, th)
        acc = r.new_accumulation(37, 29)
        r.accumulate(acc, 3)
        films.append((acc.film(), acc.stats() if adaptive else None))
    for f, s in films[1:]:
        assert np.array_equal(f, films[0][0])
        if adaptive:
            assert np.array_equal(s, films[0][1])


@pytest.mark.parametrize("scene", ["zoo", "cornell", "spheres", "lights"])
def test_passes_in_flight_do_not_change_the_film(ctx, scene):
    # 1 to 4 sample passes in flight, over calls of 3 and 4 samples: the same film bit for bit (each
    # pass's film is added in sample order), on a translucent scene (separate extend and shadow
    # launches), opaque ones (merged trace launches) and one that shades its shadow rays fused
    from verion_amd import scene as sc

    if scene == "lights":
        s = scenes.floor_scene(sc.PointLight(position=(0.0, 0.0, 2.0), color=(1.0, 1.0, 1.0), intensity=2.0), occluder=True)
        for k in range(8):
            s.add_light(sc.PointLight(position=(0.3 * k - 1.0, 0.5, 1.5), color=(0.5, 0.6, 0.7), intensity=1.0))
    else:
        s = {"zoo": scenes.zoo, "cornell": scenes.cornell, "spheres": lambda: scenes.procedural(spheres=16)}[scene]()
    r = Renderer(ctx, s)
    r.set_integrator(WAVEFRONT)
    films = []
    for passes in (1, 2, 3, 4):
        r.set_wavefront_pipeline(passes)
        acc = r.new_accumulation(40, 28)
        r.accumulate(acc, 3)
        r.accumulate(acc, 4)
        assert acc.samples() == 7
        films.append(acc.film())
        acc.close()
    r.set_wavefront_pipeline(3)
    for f in films[1:]:
        assert np.array_equal(f, films[0])
    want = oracle.OracleScene(r.packed).render(params(r.packed, 40, 28), 7)["film"]
    assert rel(films[-1], want) <= TOL
    with pytest.raises(BackendError) as e:
        r.set_wavefront_pipeline(0)
    assert e.value.status == INVALID_ARGUMENT
    with pytest.raises(BackendError):
        r.set_wavefront_pipeline(5)


def test_many_samples_per_call_split_into_megakernel_dispatches_and_wavefront_passes(ctx):
    # 128 x 128 at 2,100 spp in one call: 34.4M (pixel, sample) paths, more than one persistent
    # megakernel dispatch holds (32M), and 2,100 wavefront passes 3 at a time; the same film bit
    # for bit from both integrators, and from calls of 1,000 + 1,100 samples
    r = Renderer(ctx, scenes.cornell())
    films = []
    for integ, calls in ((MEGAKERNEL, (2100,)), (WAVEFRONT, (2100,)), (MEGAKERNEL, (1000, 1100))):
        r.set_integrator(integ)
        acc = r.new_accumulation(128, 128)
        for k in calls:
            r.accumulate(acc, k)
        assert acc.samples() == 2100
        films.append(acc.film())
        acc.close()
    r.set_integrator(MEGAKERNEL)
    assert (films[0] > 0).any()
    for f in films[1:]:
        assert np.array_equal(f, films[0])


def test_many_delta_lights_shade_fused_and_match_the_oracle(ctx):
    # more than 7 delta lights: the wavefront shades with its shadow rays inline (no deferred records)
    from verion_amd import scene as sc

    s = scenes.floor_scene(sc.PointLight(position=(0.0, 0.0, 2.0), color=(1.0, 1.0, 1.0), intensity=2.0), occluder=True)
    for k in range(8):
        s.add_light(sc.PointLight(position=(0.3 * k - 1.0, 0.5, 1.5), color=(0.5, 0.6, 0.7), intensity=1.0))
    r = Renderer(ctx, s)
    assert r.packed.info.light_count >= 9
    r.set_integrator(WAVEFRONT)
    acc = r.new_accumulation(24, 24)
    r.accumulate(acc, 2)
    # no shadow records; the 2 samples run as 2 passes in flight, each with its buffers and FP64 film
    assert acc.wavefront_bytes() == 2 * (2 * 4 * 17 * 576 + 16 * 576 + 4 * 576 + 128 + 24 * 576)
    st = oracle.OracleScene(r.packed).render(params(r.packed, 24, 24), 2)
    assert np.array_equal(acc.film(), st["film"])
    # the persistent megakernel likewise traces them in the shading (its non-deferred form)
    r.set_integrator(MEGAKERNEL)
    mega = r.new_accumulation(24, 24)
    r.accumulate(mega, 2)
    r.set_integrator(WAVEFRONT)
    assert np.array_equal(mega.film(), st["film"])


def test_environment_maps_with_dark_rows_and_columns_match_the_oracle(ctx):
    # an importance-sampled map whose CDFs have plateaus: black rows (the marginal CDF repeats),
    # black columns and all-black rows (a row's CDF stays at 0), a bright texel; the library
    # starts each CDF search from guide-table cutpoints and must pick the oracle's texels
    from verion_amd import scene as sc

    w, h = 64, 24
    px = np.full((h, w, 3), 0.2, np.float32)
    px[:5] = 0.0
    px[15:19] = 0.0
    px[:, 10:30] = 0.0
    px[8, 40] = (400.0, 300.0, 200.0)
    s = scenes.sky_scene(scenes.icosphere(2), sc.BaseBsdf(base_color=(0.7, 0.6, 0.5), roughness=0.4))
    s.lights.clear()
    s.add_light(sc.EnvironmentMap(s.add_image(sc.Image(w, h, px.reshape(-1, 3))), 1.0))
    r = Renderer(ctx, s)
    o = oracle.OracleScene(r.packed)
    want = o.render(params(r.packed, 48, 48), 16)["film"]
    for integrator in (MEGAKERNEL, WAVEFRONT):
        r.set_integrator(integrator)
        acc = r.new_accumulation(48, 48)
        r.accumulate(acc, 16)
        got = acc.film()
        assert np.array_equal(got, want), f"{integrator}: exact {np.mean(got == want):.4f}, rel {rel(got, want):.2e}"


def test_stage_timing_keeps_the_image_and_books_each_launch(ctx, zoo):
    """Stage timing puts an event after every wavefront launch and changes nothing in the image.
    An opaque scene traces depth d's closest rays with depth d - 1's shadow rays in one launch
    (trace); a translucent one (zoo: glass) runs extend and shadow on their own."""
    cases = ((Renderer(ctx, scenes.cornell()), True), (zoo[0], False))
    for r, merged in cases:
        r.set_integrator(WAVEFRONT)
        films = []
        for timing in (False, True):
            r.set_stage_timing(timing)
            acc = r.new_accumulation(64, 48)
            r.accumulate(acc, 3)
            films.append(acc.film())
            if timing:
                ext, shade, shadow, resolve, trace = acc.kernel_seconds()
                assert shade > 0 and resolve > 0
                if merged:
                    assert trace > 0 and ext == 0 and shadow == 0
                else:
                    assert trace == 0 and ext > 0 and shadow > 0
                tracing, shading = acc.stage_seconds()
                assert tracing == pytest.approx(ext + trace) and shading == pytest.approx(shade + shadow + resolve)
            else:
                assert acc.kernel_seconds() is None
        r.set_stage_timing(False)
        r.set_integrator(MEGAKERNEL)
        assert np.array_equal(films[0], films[1])
