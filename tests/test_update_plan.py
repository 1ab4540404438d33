"""Scene change tracking and the update plan, restating crates/verion-scene/src/scene.rs:562-608
(`changes_name_what_to_update`) and crates/verion-gpu/src/pack.rs:948-967
(`plan_follows_the_reuse_table`)."""
import dataclasses

import numpy as np

from verion_amd import scene as sc
from verion_amd.pack import UpdatePlan


def tri():
    return sc.Mesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.array([0, 1, 2], np.uint32))


def test_changes_name_what_to_update():
    s = sc.Scene()
    mesh = s.add_mesh(tri())
    material = s.add_material(sc.BaseBsdf())
    inst = sc.Instance(mesh, material)
    i = s.add_instance(inst)
    assert s.take_changes().geometry
    assert s.take_changes().is_empty()

    moved = dataclasses.replace(inst, transform=((1.0, 0.0, 0.0, 1.0), (0.0, 1.0, 0.0, 0.0), (0.0, 0.0, 1.0, 0.0)))
    s.set_instance(i, moved)
    s.set_instance(i, moved)
    c = s.take_changes()
    assert (c.geometry, c.instances) == (False, [i])

    bent = tri()
    bent.positions[1, 0] = 2.0
    s.set_mesh(mesh, bent)
    c = s.take_changes()
    assert (c.geometry, c.deformed) == (False, [mesh])

    s.set_material(material, sc.BaseBsdf())
    assert s.take_changes().is_empty(), "the same parameters are no change"
    s.set_material(material, sc.BaseBsdf(roughness=0.1))
    assert s.take_changes().materials == [material]

    s.set_camera(sc.Camera())
    light = s.add_light(sc.Environment((1.0, 1.0, 1.0)))
    s.set_light(light, sc.Environment((2.0, 2.0, 2.0)))
    c = s.take_changes()
    assert c.camera and c.lights and not c.geometry


def test_shape_changes_and_masks():
    s = sc.Scene()
    a = s.add_mesh(tri())
    b = s.add_mesh(tri())
    m = s.add_material(sc.BaseBsdf())
    i = s.add_instance(sc.Instance(a, m))
    s.take_changes()
    s.set_instance(i, sc.Instance(b, m))
    assert s.take_changes().geometry, "another mesh is a geometry change"
    quad = sc.Mesh(np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32), np.array([0, 1, 2, 0, 2, 3], np.uint32))
    s.set_mesh(a, quad)
    assert s.take_changes().geometry, "other triangles are a geometry change"
    s.set_instance_light_mask(i, 0xFFFFFFFF)
    assert s.take_changes().is_empty()
    s.set_instance_light_mask(i, 1)
    c = s.take_changes()
    assert c.lights and not c.geometry
    s.set_instance_emit_mask(i, 2)
    assert s.take_changes().geometry, "emitter records carry the emit mask"
    s.set_transparent_background(True)
    assert s.take_changes().lights


def test_plan_follows_the_reuse_table():
    p = UpdatePlan.new(sc.Changes(camera=True))
    assert p.camera and not p.materials and not p.tlas and not p.blases and not p.full
    p = UpdatePlan.new(sc.Changes(lights=True))
    assert p.lights and not p.emitters and not p.tlas and not p.vertices
    p = UpdatePlan.new(sc.Changes(geometry=True))
    assert p.full and p.tlas and p.vertices
    assert UpdatePlan.new(sc.Changes()).is_empty()
    p = UpdatePlan.new(sc.Changes(materials=[3]))
    assert p.materials and p.emitters and p.instances and not p.tlas
    p = UpdatePlan.new(sc.Changes(instances=[1]))
    assert p.instances and p.tlas and p.emitters and not p.vertices
    p = UpdatePlan.new(sc.Changes(deformed=[0, 2]))
    assert p.vertices and p.blases == [0, 2] and p.tlas and p.instances
