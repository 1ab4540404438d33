"""The neural radiance cache on the GPU against the FP64 numpy oracle (oracle/nrc.py).

- inference: the fused BF16-MFMA forward against the oracle's bfloat16-rounding forward;
- gradients: the fused forward/backward kernel against the oracle's analytic backward
  (relative L2 error of the whole gradient and per-layer cosine);
- Adam: one step against its closed form with the GPU's own gradient;
- training lowers the loss on a smooth target;
- the integrator: a trained cache gives a lower error than plain path tracing at equal frames.
"""
import numpy as np
import pytest

from oracle import nrc
from verion_amd import scenes
from verion_amd.backend import WAVEFRONT, Nrc, Renderer

pytestmark = pytest.mark.gpu


def batch(n, seed):
    rng = np.random.default_rng(seed)
    raw = np.zeros((n, 12), np.float32)
    raw[:, 0:3] = rng.random((n, 3))
    raw[:, 3] = rng.random(n)
    d = rng.normal(size=(n, 3))
    raw[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    nn = rng.normal(size=(n, 3))
    raw[:, 8:11] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    target = (0.5 + 0.4 * np.sin(3 * raw[:, 0:3]) * raw[:, 4:7]).astype(np.float32)
    return raw, target


@pytest.fixture(scope="module")
def cache(ctx):
    r = Renderer(ctx, scenes.cornell())
    return r, Nrc(r, train_paths=4096)


def test_parameters_round_trip(cache):
    _, c = cache
    theta = nrc.init(1)
    c.set_params(theta)
    assert np.array_equal(c.params(), theta)


def test_inference_matches_the_bf16_oracle(cache):
    _, c = cache
    theta = nrc.init(2)
    # larger biases and output weights so the outputs are far from zero
    layers = nrc.unpack(theta.astype(np.float64))
    theta = theta.copy()
    theta[nrc.param_count() - 16:nrc.param_count() - 13] = 0.7
    c.set_params(theta)
    raw, _ = batch(4096, 3)
    got = c.infer(raw)
    want = nrc.forward(theta, raw, bf16_round=True)[0][:, :3]
    err = np.abs(got - want)
    assert np.median(err / (np.abs(want) + 1e-2)) < 4e-3
    assert err.max() < 0.02 * np.abs(want).max()
    del layers


def test_gradient_matches_the_oracle(cache):
    _, c = cache
    theta = nrc.init(4)
    c.set_params(theta)
    raw, target = batch(2048, 5)
    g, loss = c.gradient(raw, target)
    want = nrc.backward(theta, raw, target, bf16_round=True)
    rel = np.linalg.norm(g - want) / np.linalg.norm(want)
    assert rel < 0.03, rel
    o = 0
    for k, n in nrc.LAYERS:
        gw, ww = g[o:o + k * n], want[o:o + k * n]
        if np.linalg.norm(ww) > 0:
            assert float(gw @ ww / (np.linalg.norm(gw) * np.linalg.norm(ww))) > 0.999
        o += k * n + n
    y = nrc.forward(theta, raw, bf16_round=True)[0]
    assert abs(loss - nrc.loss(y, target)) < 0.01 * nrc.loss(y, target)


def test_adam_step_matches_its_closed_form(cache):
    _, c = cache
    theta = nrc.init(6)
    c.set_params(theta)
    raw, target = batch(1024, 7)
    g, _ = c.gradient(raw, target)
    c.train(raw, target, steps=1)
    got = c.params()
    want, _, _ = nrc.adam(theta.astype(np.float64), np.zeros(theta.size), np.zeros(theta.size), g.astype(np.float64), 1, lr=1e-2)
    assert np.allclose(got, want, rtol=1e-5, atol=1e-6)


def test_training_lowers_the_loss(cache):
    _, c = cache
    c.set_params(nrc.init(8))
    raw, target = batch(8192, 9)
    first = c.train(raw, target, steps=1)
    last = c.train(raw, target, steps=300)
    assert last < 0.2 * first, (first, last)


def relmse(x, ref):
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + 0.01)))


def test_trained_cache_beats_plain_tracing_at_equal_frames(ctx):
    # Cornell at 128x128: a 4096-spp plain render is the reference; after 16 frames of 1 spp the
    # cache-terminated image has a lower relMSE than the plain one
    r = Renderer(ctx, scenes.cornell())
    r.set_integrator(WAVEFRONT)
    ref = r.render(128, 128, 4096).astype(np.float64)
    plain = r.new_accumulation(128, 128)
    r.accumulate(plain, 16)
    c = Nrc(r, train_paths=8192, train_steps=4)
    c.set_params(nrc.init(10))
    warm = r.new_accumulation(128, 128)
    c.accumulate(warm, 64)  # warm-up: the cache learns the scene
    acc = r.new_accumulation(128, 128)
    c.accumulate(acc, 16)
    e_plain = relmse(plain.mean(), ref)
    e_nrc = relmse(acc.mean(), ref)
    st = c.stats()
    assert st["samples"] > 8192
    assert np.all(np.isfinite(acc.mean()))
    assert e_nrc < e_plain, (e_nrc, e_plain, st)


def test_inference_uses_the_bias_corrected_moving_average(ctx):
    # after three Adam steps from theta0 the integrator's weights are
    # (1-d)(d^2 theta1 + d theta2 + theta3) / (1 - d^3), d = 0.95
    r = Renderer(ctx, scenes.cornell())
    c = Nrc(r, train_paths=1024, ema_decay=0.95, learning_rate=0.05)
    c.set_params(nrc.init(12))
    raw, target = batch(2048, 13)
    thetas = []
    for _ in range(3):
        c.train(raw, target, steps=1)
        thetas.append(c.params().astype(np.float64))
    d = 0.95
    avg = (1 - d) * (d * d * thetas[0] + d * thetas[1] + thetas[2]) / (1 - d ** 3)
    got = c.infer(raw)
    want = nrc.forward(avg.astype(np.float32), raw, bf16_round=True)[0][:, :3]
    last = nrc.forward(thetas[2].astype(np.float32), raw, bf16_round=True)[0][:, :3]
    err = np.abs(got - want).max()
    assert err < 0.02 * np.abs(want).max() + 1e-3, err
    # and it is the average, not the last parameters
    assert err < 0.5 * np.abs(got - last).max(), (err, np.abs(got - last).max())


def test_minibatch_steps_take_disjoint_slices(ctx):
    # one step per frame: the one slice is every sample, so both modes draw the same film (up to
    # the order the training samples arrive in, which is not fixed); with four steps the minibatch
    # mode trains on quarter slices and still renders finite images
    r = Renderer(ctx, scenes.cornell())
    r.set_integrator(WAVEFRONT)
    films = []
    for minibatch in (False, True):
        c = Nrc(r, train_paths=4096, train_steps=1, warmup_frames=2, minibatch=minibatch)
        c.set_params(nrc.init(10))
        acc = r.new_accumulation(64, 64)
        c.accumulate(acc, 4)
        films.append(acc.film())
    diff = np.abs(films[0] - films[1]) / np.maximum(np.abs(films[1]), 1e-3)
    assert diff.mean() < 1e-3, diff.mean()
    c = Nrc(r, train_paths=4096, train_steps=4, warmup_frames=2, minibatch=True)
    c.set_params(nrc.init(10))
    acc = r.new_accumulation(64, 64)
    c.accumulate(acc, 4)
    assert np.all(np.isfinite(acc.mean())) and c.stats()["samples"] > 0


def test_training_paths_defer_their_shadow_rays_without_changing_the_samples(ctx):
    # The same training paths with their shadow rays deferred (at most 7 delta lights) and traced
    # inline (8 more point lights of zero intensity force the fused shading): a frame's training
    # paths leave the same number of samples, and training on them gives the same loss
    from verion_amd import scene as sc
    base = scenes.cornell()
    fused = scenes.cornell()
    for k in range(8):
        fused.add_light(sc.PointLight(position=(0.1 * k, 0.5, 0.5), color=(1.0, 1.0, 1.0), intensity=0.0))
    stats = []
    for s in (base, fused):
        r = Renderer(ctx, s)
        r.set_integrator(WAVEFRONT)
        c = Nrc(r, train_paths=4096, train_steps=1, warmup_frames=0)
        c.set_params(nrc.init(10))
        acc = r.new_accumulation(32, 32)
        c.accumulate(acc, 1)
        stats.append(c.stats())
    assert stats[0]["samples"] == stats[1]["samples"] > 0, stats
    assert abs(stats[0]["loss"] - stats[1]["loss"]) <= 1e-4 * abs(stats[1]["loss"]), stats


def test_a_fresh_cache_starts_from_the_library_initialisation_and_trains(ctx):
    # the library stands alone: vh_nrc_create starts from vh_nrc_init_params(0), vh_nrc_init(seed)
    # restarts from another seed, and Adam lowers the loss from there (a ReLU MLP at zero would not)
    r = Renderer(ctx, scenes.cornell())
    c = Nrc(r, train_paths=1024)
    np.testing.assert_array_equal(c.params(), Nrc.init_params(0))
    c.init(5)
    np.testing.assert_array_equal(c.params(), Nrc.init_params(5))
    raw, target = batch(8192, 9)
    first = c.train(raw, target, steps=1)
    last = c.train(raw, target, steps=300)
    assert last < 0.2 * first, (first, last)


def test_merged_trace_launches_render_the_nrc_frames_of_the_fused_path(ctx):
    # Opaque scenes with at most 7 delta lights trace a depth's closest rays together with the
    # previous depth's shadow rays (one launch) and resolve afterwards; 8 more point lights of zero
    # intensity force the fused shading and separate launches. From the same parameters both draw
    # the same film up to the order the training samples arrive in (mean relative difference, as
    # the minibatch test)
    from verion_amd import scene as sc
    base = scenes.cornell()
    fused = scenes.cornell()
    for k in range(8):
        fused.add_light(sc.PointLight(position=(0.1 * k, 0.5, 0.5), color=(1.0, 1.0, 1.0), intensity=0.0))
    films = []
    for s in (base, fused):
        r = Renderer(ctx, s)
        r.set_integrator(WAVEFRONT)
        c = Nrc(r, train_paths=4096, train_steps=1, warmup_frames=2)
        c.init(10)
        acc = r.new_accumulation(64, 64)
        c.accumulate(acc, 4)
        films.append(acc.film())
    diff = np.abs(films[0] - films[1]) / np.maximum(np.abs(films[1]), 1e-3)
    assert np.all(np.isfinite(films[0])) and diff.mean() < 1e-3, diff.mean()
