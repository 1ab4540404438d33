"""Batched jobs on the GPU (vh_batch_*): every job's film equals rendering it alone with
vh_accumulate (wavefront) bit for bit, whatever the tuning, with or without the captured graph,
across graph replays with eager work in between; a job's film also equals the oracle's; the job
statistics are consistent."""
import numpy as np
import pytest

import oracle
from verion_amd import pack, scenes
from verion_amd.backend import WAVEFRONT, Adaptive, Batch, BackendError, Renderer

from helpers import params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def renderers(ctx):
    rs = {
        "cornell": Renderer(ctx, scenes.cornell()),
        "procedural": Renderer(ctx, scenes.procedural(spheres=24)),
        "zoo": Renderer(ctx, scenes.zoo()),
        "lights": Renderer(ctx, scenes.many_lights(emitters=300, clutter_boxes=400)),
    }
    for r in rs.values():
        r.set_integrator(WAVEFRONT)
    return rs


def job_list(rs):
    return [
        dict(renderer=rs["cornell"], width=64, height=48, spp=5),
        dict(renderer=rs["procedural"], width=37, height=29, spp=3, seed=7),
        dict(renderer=rs["zoo"], width=80, height=45, spp=6, adaptive=Adaptive(0.05, 2)),
        dict(renderer=rs["lights"], width=50, height=50, spp=4, seed=3),
        dict(renderer=rs["cornell"], width=16, height=16, spp=9, adaptive=Adaptive(0.02, 3), seed=11),
    ]


def alone(job):
    r = job["renderer"]
    r.set_seed(job.get("seed", 0))
    r.set_adaptive(job.get("adaptive"))
    acc = r.new_accumulation(job["width"], job["height"])
    r.accumulate(acc, job["spp"])
    film = acc.film()
    st = acc.stats() if job.get("adaptive") else None
    r.set_adaptive(None)
    r.set_seed(0)
    return film, st


@pytest.fixture(scope="module")
def expected(renderers):
    return [alone(j) for j in job_list(renderers)]


def check(batch, expected, jobs):
    for k, (film, st) in enumerate(expected):
        assert np.array_equal(batch.film(k), film), k
        if st is not None:
            assert np.array_equal(batch.stats(k), st), k


def test_batch_films_equal_single_job_renders(ctx, renderers, expected):
    jobs = job_list(renderers)
    b = Batch(ctx, jobs)
    b.run(graph=True)
    check(b, expected, jobs)
    report = b.tuning_report()
    assert 1 <= len(report) <= 3  # the batch's smallest spp bounds the fair tuning window
    assert b.tuning() in [t for t, _ in report]
    for k, j in enumerate(jobs):
        s = b.job_stats(k)
        assert 0 < s["latency_s"] <= s["batch_s"] + 1e-9
        if not j.get("adaptive"):
            assert s["passes"] == j["spp"] and s["paths"] == j["spp"] * j["width"] * j["height"]
        else:
            assert 1 <= s["passes"] <= j["spp"]
    lat = {k: b.job_stats(k)["latency_s"] for k in range(len(jobs))}
    assert lat[1] <= lat[0] <= lat[4]  # 3, 5 and 9 passes


@pytest.mark.parametrize("tuning,graph", [((512, 0, 0), False), ((2048, 16, 4), True), ((64, 8, 8), True)])
def test_tuning_and_graph_do_not_change_the_image(ctx, renderers, expected, tuning, graph):
    jobs = job_list(renderers)
    b = Batch(ctx, jobs)
    b.set_tuning(tuning)
    b.run(graph=graph)
    check(b, expected, jobs)
    assert b.tuning() == tuning


def test_graph_replays_with_eager_work_between(ctx, renderers, expected):
    jobs = job_list(renderers)
    b = Batch(ctx, jobs)
    b.set_tuning((1024, 8, 8))
    for _ in range(4):
        b.run(graph=True, max_passes=2)
        # the program's own eager work between replays
        r = renderers["cornell"]
        acc = r.new_accumulation(32, 32)
        r.accumulate(acc, 1)
        assert np.isfinite(acc.mean()).all()
    b.run(graph=True)
    check(b, expected, jobs)
    b.reset()
    b.run(graph=True)
    check(b, expected, jobs)


def test_a_job_equals_the_oracle(ctx, renderers):
    job = dict(renderer=renderers["cornell"], width=40, height=30, spp=4)
    b = Batch(ctx, [job, dict(renderer=renderers["procedural"], width=20, height=20, spp=2)])
    b.run()
    p = renderers["cornell"].packed
    st = oracle.OracleScene(p).render(params(p, 40, 30), 4)
    assert np.array_equal(b.film(0), st["film"])
    assert np.allclose(b.mean(0), oracle.mean(st, 4), rtol=1e-7, atol=0)


def test_bad_jobs_are_rejected(ctx, renderers):
    r = renderers["cornell"]
    with pytest.raises(BackendError):
        Batch(ctx, [dict(renderer=r, width=0, height=8, spp=1)])
    with pytest.raises(BackendError):
        Batch(ctx, [dict(renderer=r, width=8, height=8, spp=0)])
    with pytest.raises(BackendError):
        Batch(ctx, [])
    b = Batch(ctx, [dict(renderer=r, width=8, height=8, spp=1)])
    with pytest.raises(BackendError):
        b.set_tuning((0, 8, 8))
    with pytest.raises(BackendError):
        b.set_tuning((256, 8, 0))
