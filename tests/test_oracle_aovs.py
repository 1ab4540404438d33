"""Pins of the oracle's primary-hit AOVs (aov.slang:84-188, normals.slang:39-110).

- the reference's own GPU test (crates/verion-gpu-vulkan/tests/gpu.rs:105-126) restated: where the
  centre ray hits, the normal is a unit vector, depth > 0 and the material id > 0; elsewhere both
  are 0;
- closed forms: a plane facing the camera at distance D has depth D at every pixel (forward-axis
  distance, not ray length), normal +z, ids 1/1, coverage 1 and the material's base colour; an
  edge through the centre of the middle pixel column gives coverage 1/2; two stacked layers of
  opacity 1/2 give coverage 3/4 and an opaque layer under them 1; a constant base texture scales
  the albedo; a smooth sphere's normals match the radial direction; a mirrored instance still
  faces the viewer.
"""
import numpy as np

import oracle
from verion_amd import pack, scenes
from verion_amd import scene as sc

from helpers import params


def run(scene, w, h):
    p = pack.pack(scene)
    o = oracle.OracleScene(p)
    par = params(p, w, h)
    return o.aovs(par), o.normals(par), o


def plane_scene(d=4.0, half=50.0, colour=(0.3, 0.6, 0.9), opacity=1.0, z=0.0):
    s = sc.Scene()
    m = s.add_mesh(scenes.plane_mesh(half, z))
    s.add_instance(sc.Instance(m, s.add_material(sc.BaseBsdf(base_color=colour, opacity=opacity))))
    s.camera = scenes.camera_at_z(d)
    return s


def test_reference_first_hit_invariants():
    w = h = 48
    aov, nrm, _ = run(scenes.cornell(), w, h)
    hits = aov["object_id"] > 0
    assert hits.sum() > w * h // 2
    length = np.linalg.norm(nrm, axis=1)
    assert np.all(np.abs(length[hits] - 1.0) < 1e-3)
    assert np.all(aov["depth"][hits] > 0.0) and np.all(aov["material_id"][hits] > 0)
    assert np.all(length[~hits] == 0.0) and np.all(aov["depth"][~hits] == 0.0)


def test_plane_facing_the_camera():
    w, h = 33, 21
    aov, nrm, _ = run(plane_scene(d=4.0), w, h)
    assert np.all(aov["object_id"] == 1) and np.all(aov["material_id"] == 1)
    assert np.allclose(aov["depth"], 4.0, rtol=2e-6)
    assert np.allclose(nrm, [0.0, 0.0, 1.0], atol=1e-6)
    assert np.all(aov["coverage"] == 1.0)
    assert np.allclose(aov["albedo"], [0.3, 0.6, 0.9], atol=0)


def test_miss_everywhere_behind_the_camera():
    aov, nrm, _ = run(plane_scene(d=-4.0), 16, 16)
    assert np.all(aov["object_id"] == 0) and np.all(aov["coverage"] == 0.0) and np.all(nrm == 0.0)


def test_half_covered_pixel_column():
    # a plane over x >= 0 only; odd width: the middle column's centre lies on the edge, and two of
    # the four stratified columns of each pixel there see the plane
    w, h = 15, 9
    s = sc.Scene()
    v = np.array([[0, -50, 0], [50, -50, 0], [50, 50, 0], [0, 50, 0]], np.float32)
    m = s.add_mesh(sc.Mesh(v, np.array([0, 1, 2, 0, 2, 3], np.uint32)))
    s.add_instance(sc.Instance(m, s.add_material(sc.lambert((0.5, 0.5, 0.5)))))
    s.camera = scenes.camera_at_z(3.0)
    aov, _, _ = run(s, w, h)
    cov = aov["coverage"].reshape(h, w)
    assert np.all(cov[:, :w // 2] == 0.0)
    assert np.all(cov[:, w // 2] == 0.5)
    assert np.all(cov[:, w // 2 + 1:] == 1.0)


def test_translucent_layers():
    s = sc.Scene()
    half = scenes.plane_mesh(50.0, 1.0)
    half2 = scenes.plane_mesh(50.0, 0.5)
    m1, m2 = s.add_mesh(half), s.add_mesh(half2)
    glassy = s.add_material(sc.BaseBsdf(base_color=(1.0, 1.0, 1.0), opacity=0.5))
    s.add_instance(sc.Instance(m1, glassy))
    s.add_instance(sc.Instance(m2, glassy))
    s.camera = scenes.camera_at_z(3.0)
    aov, _, _ = run(s, 9, 9)
    assert np.all(aov["coverage"] == 0.75)
    assert np.all(aov["object_id"] == 1)  # the nearer layer
    s.add_instance(sc.Instance(s.add_mesh(scenes.plane_mesh(50.0, 0.0)), s.add_material(sc.lambert((1, 1, 1)))))
    aov, _, _ = run(s, 9, 9)
    assert np.all(aov["coverage"] == 1.0)


def test_constant_base_texture_scales_the_albedo():
    s = sc.Scene()
    tex = s.add_image(sc.Image(4, 4, np.tile(np.array([[0.5, 0.25, 1.0]], np.float32), (16, 1))))
    v = np.array([[-5, -5, 0], [5, -5, 0], [5, 5, 0], [-5, 5, 0]], np.float32)
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    m = s.add_mesh(sc.Mesh(v, np.array([0, 1, 2, 0, 2, 3], np.uint32), uvs=uv))
    s.add_instance(sc.Instance(m, s.add_material(sc.BaseBsdf(base_color=(0.8, 0.8, 0.8), base_color_texture=tex))))
    s.camera = scenes.camera_at_z(3.0)
    aov, _, _ = run(s, 8, 8)
    hit = aov["object_id"] > 0
    assert hit.all()
    assert np.allclose(aov["albedo"][hit], np.float32(0.8) * np.array([0.5, 0.25, 1.0], np.float32), rtol=1e-6)


def test_sphere_normals_are_radial():
    s = sc.Scene()
    m = s.add_mesh(scenes.icosphere(4))
    s.add_instance(sc.Instance(m, s.add_material(sc.lambert((0.5, 0.5, 0.5))), ((2, 0, 0, 0), (0, 2, 0, 0), (0, 0, 2, 0))))
    s.camera = scenes.camera_at_z(8.0)
    w = 40
    aov, nrm, _ = run(s, w, w)
    hit = aov["object_id"] > 0
    assert hit.sum() > 100
    # the hit point from depth and the camera-space direction of each pixel
    ys, xs = np.divmod(np.arange(w * w), w)
    pk = pack.pack(s)
    tanf = float(pk.view.camera[2])
    u = (2.0 * (xs + 0.5) / w - 1.0) * tanf
    v = (1.0 - 2.0 * (ys + 0.5) / w) * tanf
    d = np.stack([u, v, -np.ones_like(u)], 1)
    p = np.array([0.0, 0.0, 8.0]) + d * aov["depth"][:, None]  # depth is the forward (-z) distance
    radial = p / np.linalg.norm(p, axis=1, keepdims=True)
    cosang = np.sum(radial[hit] * nrm[hit], axis=1)
    assert np.all(cosang > 0.999), cosang.min()


def test_mirrored_instance_faces_the_viewer():
    s = sc.Scene()
    m = s.add_mesh(scenes.plane_mesh(50.0, 0.0))
    s.add_instance(sc.Instance(m, s.add_material(sc.lambert((0.5, 0.5, 0.5))), ((-1, 0, 0, 0), (0, 1, 0, 0), (0, 0, 1, 0))))
    s.camera = scenes.camera_at_z(2.0)
    _, nrm, _ = run(s, 8, 8)
    assert np.allclose(nrm, [0.0, 0.0, 1.0], atol=1e-6)


def test_bvh_and_brute_force_agree():
    p = pack.pack(scenes.zoo())
    o = oracle.OracleScene(p)
    par = params(p, 24, 24)
    a, b = o.aovs(par), o.aovs(par, brute=True)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(o.normals(par), o.normals(par, brute=True))
