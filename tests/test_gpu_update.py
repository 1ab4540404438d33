"""Differential scene updates on the GPU, restating crates/verion-gpu-vulkan/tests/update.rs:1-208:
an updated renderer draws bit for bit what a renderer built from scratch on the edited scene draws,
and redoes only what the edit needs (UpdatePlan, pack.rs:758-825). Here a moved instance or a
deformed mesh refits the one LBVH in place (report.bvh_work 1), and a refit that degrades the tree
past the limit is rebuilt (bvh_work 2): the image does not depend on the tree."""
import dataclasses

import numpy as np
import pytest

from verion_amd import scene as sc
from verion_amd import scenes
from verion_amd.backend import INVALID_ARGUMENT, MEGAKERNEL, WAVEFRONT, BackendError, Renderer

pytestmark = pytest.mark.gpu


def cube(half, y=0.0):
    """update.rs:13-30."""
    h = half
    pos = [[-h, y - h, -h], [h, y - h, -h], [h, y + h, -h], [-h, y + h, -h],
           [-h, y - h, h], [h, y - h, h], [h, y + h, h], [-h, y + h, h]]
    idx = [0, 1, 2, 0, 2, 3, 4, 5, 6, 4, 6, 7, 0, 1, 5, 0, 5, 4, 2, 3, 7, 2, 7, 6, 1, 2, 6, 1, 6, 5, 0, 3, 7, 0, 7, 4]
    return sc.Mesh(np.array(pos, np.float32), np.array(idx, np.uint32))


def diffuse(color, emission=(0.0, 0.0, 0.0)):
    return sc.BaseBsdf(base_color=tuple(color), emission=tuple(emission), ior=1.0)


def room():
    """update.rs:52-90: a closed room lit by a small emissive cube and a point light, seen from inside."""
    s = sc.Scene()
    room_mesh = s.add_mesh(cube(1.0))
    lamp_mesh = s.add_mesh(cube(0.1))
    walls = s.add_material(diffuse((0.6, 0.6, 0.6)))
    glow = s.add_material(diffuse((0.0, 0.0, 0.0), (8.0, 8.0, 8.0)))
    s.add_instance(sc.Instance(room_mesh, walls))
    lamp = s.add_instance(sc.Instance(lamp_mesh, glow, ((1.0, 0.0, 0.0, 0.0), (0.0, 1.0, 0.0, 0.7), (0.0, 0.0, 1.0, 0.0))))
    s.add_light(sc.PointLight((0.0, -0.5, 0.0), (1.0, 1.0, 1.0), 0.5))
    s.set_camera(scenes.camera_at_z(0.95))
    s.take_changes()
    return s, walls, lamp_mesh, lamp


def render(r, w=48, h=48, spp=8):
    acc = r.new_accumulation(w, h)
    r.accumulate(acc, spp)
    return acc.mean()


def update_matches_fresh(ctx, r, s, **kw):
    """update.rs:98-111."""
    before = render(r, **kw)
    report = r.update(s, s.take_changes())
    after = render(r, **kw)
    fresh = Renderer(ctx, s)
    fresh.set_integrator(r.integrator())
    want = render(fresh, **kw)
    assert np.array_equal(after.view(np.uint32), want.view(np.uint32)), "the update differs from a rebuild"
    assert not np.array_equal(before, after), "the edit changed nothing"
    return report


@pytest.mark.parametrize("integrator", [MEGAKERNEL, WAVEFRONT])
def test_a_material_edit_reuses_every_acceleration_structure(ctx, integrator):
    s, walls, _, _ = room()
    r = Renderer(ctx, s)
    r.set_integrator(integrator)
    s.set_material(walls, diffuse((0.9, 0.1, 0.1)))
    rep = update_matches_fresh(ctx, r, s)
    assert (rep.blases_built, rep.tlas_built, rep.bvh_work) == (0, 0, 0)
    assert rep.bytes_uploaded > 0


def test_a_light_edit_reuses_every_acceleration_structure(ctx):
    s, _, _, _ = room()
    r = Renderer(ctx, s)
    s.set_light(0, sc.PointLight((0.0, -0.5, 0.0), (1.0, 0.5, 0.2), 2.0))
    rep = update_matches_fresh(ctx, r, s)
    assert (rep.blases_built, rep.tlas_built, rep.bvh_work) == (0, 0, 0)


@pytest.mark.parametrize("integrator", [MEGAKERNEL, WAVEFRONT])
def test_moving_an_instance_refits_the_tree(ctx, integrator):
    s, _, _, lamp = room()
    r = Renderer(ctx, s)
    r.set_integrator(integrator)
    inst = s.instances[lamp]
    t = [list(row) for row in inst.transform]
    t[0][3] = 0.4
    s.set_instance(lamp, dataclasses.replace(inst, transform=tuple(tuple(row) for row in t)))
    rep = update_matches_fresh(ctx, r, s)
    assert (rep.blases_built, rep.tlas_built) == (0, 1)
    assert rep.bvh_work in (1, 2)


def test_deforming_a_mesh_redoes_its_triangles(ctx):
    s, _, lamp_mesh, _ = room()
    r = Renderer(ctx, s)
    s.set_mesh(lamp_mesh, cube(0.3))
    rep = update_matches_fresh(ctx, r, s)
    assert (rep.blases_built, rep.tlas_built) == (1, 1)
    assert rep.bvh_work in (1, 2)


def test_a_camera_edit_rebuilds_nothing(ctx):
    s, _, _, _ = room()
    r = Renderer(ctx, s)
    cam = s.camera
    t = [list(row) for row in cam.transform]
    t[0][3] = 0.3
    s.set_camera(dataclasses.replace(cam, transform=tuple(tuple(row) for row in t)))
    rep = update_matches_fresh(ctx, r, s)
    assert (rep.blases_built, rep.tlas_built, rep.bytes_uploaded, rep.bvh_work) == (0, 0, 0, 0)


def test_a_new_instance_rebuilds_everything_but_keeps_the_settings(ctx):
    s, walls, _, _ = room()
    r = Renderer(ctx, s)
    r.set_integrator(WAVEFRONT)
    mesh = s.add_mesh(cube(0.2))
    s.add_instance(sc.Instance(mesh, walls, ((1.0, 0.0, 0.0, -0.5), (0.0, 1.0, 0.0, 0.0), (0.0, 0.0, 1.0, 0.0))))
    rep = update_matches_fresh(ctx, r, s)
    assert (rep.blases_built, rep.tlas_built, rep.bvh_work) == (3, 1, 2)
    assert r.integrator() == WAVEFRONT


def test_small_moves_refit_and_scattering_everything_rebuilds(ctx):
    """Beyond update.rs: on the 1,024-sphere scene a few small moves keep the tree (refit), moving
    every sphere far degrades it past the limit and the tree is built anew; both draw the fresh image."""
    s = scenes.procedural(spheres=1024)
    s.take_changes()
    r = Renderer(ctx, s)
    r.set_integrator(WAVEFRONT)
    rng = np.random.default_rng(5)
    for k in range(0, 1024, 97):
        inst = s.instances[k]
        t = [list(row) for row in inst.transform]
        t[0][3] += 0.05
        s.set_instance(k, dataclasses.replace(inst, transform=tuple(tuple(row) for row in t)))
    rep = update_matches_fresh(ctx, r, s, w=96, h=64, spp=2)
    assert rep.bvh_work == 1 and rep.tlas_built == 1
    for k, inst in enumerate(s.instances):
        t = [list(row) for row in inst.transform]
        for a in range(3):
            t[a][3] = float(rng.uniform(-10.0, 10.0))
        s.set_instance(k, dataclasses.replace(inst, transform=tuple(tuple(row) for row in t)))
    rep = update_matches_fresh(ctx, r, s, w=96, h=64, spp=2)
    assert rep.bvh_work == 2


def test_a_partial_plan_must_keep_the_geometry_sizes(ctx):
    s, _, lamp_mesh, _ = room()
    r = Renderer(ctx, s)
    s.meshes[lamp_mesh] = sc.Mesh(np.array([[0, 0, 0], [0.1, 0, 0], [0, 0.1, 0]], np.float32), np.array([0, 1, 2], np.uint32))
    with pytest.raises(BackendError) as e:
        r.update(s, sc.Changes(deformed=[lamp_mesh]))
    assert e.value.status == INVALID_ARGUMENT
    r.update(s, sc.Changes(geometry=True))
    fresh = Renderer(ctx, s)
    assert np.array_equal(render(r).view(np.uint32), render(fresh).view(np.uint32))
