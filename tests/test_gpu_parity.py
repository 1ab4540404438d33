"""GPU parity: the HIP path against the CPU oracle on the same seeded inputs.

Gates (restated from the reference's GPU tests, crates/verion-gpu-metal/tests/pathtrace.rs):
  - primary-hit (instance, primitive, t) bit-exact against the oracle's brute force;
  - per-pixel radiance within 1e-4 relative of the oracle at equal seed and spp (FP32 path,
    FP64 film; the library is built to reproduce the oracle's bits, so the test also reports
    the exact-match count);
  - wavefront == megakernel bit for bit (pathtrace.rs:170-176);
  - white furnace within 1e-4 (pathtrace.rs:118-137).
"""
import numpy as np
import pytest

import oracle
from verion_amd import pack, scenes
from verion_amd.backend import MEGAKERNEL, WAVEFRONT, Renderer

pytestmark = pytest.mark.gpu

REL_TOL = 1.0e-4


def params(packed, w, h, lighting=False, adaptive=None, restir=None):
    return pack.base_params(packed.info, packed.view, w, h, lighting, True, adaptive, restir)


def rel_close(got, want, tol=REL_TOL):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    err = np.abs(got - want) / np.maximum(np.abs(want), 1.0e-6)
    return float(err.max()) if err.size else 0.0


@pytest.fixture(scope="module")
def cornell(ctx):
    r = Renderer(ctx, scenes.cornell())
    return r, oracle.OracleScene(r.packed)


@pytest.fixture(scope="module")
def procedural(ctx):
    r = Renderer(ctx, scenes.procedural())
    return r, oracle.OracleScene(r.packed)


@pytest.mark.parametrize("sample", [0, 1, 7])
def test_primary_hits_bit_exact_cornell(cornell, sample):
    r, o = cornell
    got = r.primary_hits(256, 256, sample)
    want = o.primary_hits(params(r.packed, 256, 256), sample, brute=True)
    assert np.array_equal(got, want), f"{int((got != want).any(1).sum())} of {len(got)} pixels differ"


def test_primary_hits_bit_exact_procedural(procedural):
    r, o = procedural
    w, h = 480, 270
    got = r.primary_hits(w, h, 3)
    want = o.primary_hits(params(r.packed, w, h), 3)
    assert np.array_equal(got, want), f"{int((got != want).any(1).sum())} of {len(got)} pixels differ"
    assert got[:, 0].mean() > 0.3  # the view is mostly spheres


def test_bvh_shape(procedural):
    r, _ = procedural
    s = r.bvh_stats()
    assert s["triangles"] == 1024 * 1280
    assert 0 < s["nodes"] < s["triangles"]
    assert s["depth"] <= 20


@pytest.mark.parametrize("integrator", [MEGAKERNEL, WAVEFRONT])
def test_radiance_matches_oracle_cornell(cornell, integrator):
    r, o = cornell
    r.set_integrator(integrator)
    acc = r.new_accumulation(64, 64)
    r.accumulate(acc, 16)
    film = acc.film()
    want = o.render(params(r.packed, 64, 64), 16)["film"]
    exact = int(np.sum(film == want))
    assert rel_close(film, want) <= REL_TOL, f"exact {exact}/{film.size}"
    r.set_integrator(MEGAKERNEL)


@pytest.mark.parametrize("integrator", [MEGAKERNEL, WAVEFRONT])
def test_radiance_matches_oracle_procedural(procedural, integrator):
    r, o = procedural
    r.set_integrator(integrator)
    acc = r.new_accumulation(96, 54)
    r.accumulate(acc, 2)
    film = acc.film()
    want = o.render(params(r.packed, 96, 54), 2)["film"]
    exact = int(np.sum(film == want))
    assert rel_close(film, want) <= REL_TOL, f"exact {exact}/{film.size}"
    r.set_integrator(MEGAKERNEL)


def test_wavefront_equals_megakernel(procedural):
    r, _ = procedural
    r.set_integrator(MEGAKERNEL)
    a = r.new_accumulation(160, 90)
    r.accumulate(a, 4)
    r.set_integrator(WAVEFRONT)
    b = r.new_accumulation(160, 90)
    r.accumulate(b, 4)
    r.set_integrator(MEGAKERNEL)
    assert np.array_equal(a.film(), b.film())


def test_white_furnace(ctx):
    r = Renderer(ctx, scenes.white_furnace())
    img = r.render(32, 32, 4)
    assert np.all(np.abs(img - 1.0) < 1.0e-4)


def test_closed_room_series(ctx):
    r = Renderer(ctx, scenes.furnace_room(0.5, 1.0))
    got = float(r.render(64, 64, 64)[:, 0].mean())
    want = 1.0 * (1.0 - 0.5 ** 8) / (1.0 - 0.5)
    assert abs(got - want) / want < 0.01


@pytest.mark.parametrize("ntri", [1, 2, 3, 4, 5])
def test_tiny_scenes_trace_and_render(ctx, ntri):
    # scenes of at most 4 triangles have a single BVH node with empty child slots
    from verion_amd import scene as sc

    rng = np.random.default_rng(ntri)
    s = sc.Scene()
    v = (rng.random((3 * ntri, 3)) * 2.0 - 1.0).astype(np.float32)
    v[:, 2] = rng.random(3 * ntri).astype(np.float32) * 0.5
    m = s.add_mesh(sc.Mesh(v, np.arange(3 * ntri, dtype=np.uint32)))
    s.add_instance(sc.Instance(m, s.add_material(sc.lambert((0.7, 0.6, 0.5), (0.2, 0.2, 0.2)))))
    s.add_light(sc.DirectionalLight(direction=(0.3, -0.2, -1.0), color=(1.0, 1.0, 1.0), intensity=2.0))
    s.camera = scenes.camera_at_z(3.0)
    r = Renderer(ctx, s)
    p = r.packed
    o = oracle.OracleScene(p)
    par = pack.base_params(p.info, p.view, 24, 20, False, True)
    assert np.array_equal(r.primary_hits(24, 20, 0), o.primary_hits(par, 0, brute=True))
    r.set_integrator(WAVEFRONT)
    acc = r.new_accumulation(24, 20)
    r.accumulate(acc, 2)
    assert np.array_equal(acc.film(), o.render(par, 2)["film"])
