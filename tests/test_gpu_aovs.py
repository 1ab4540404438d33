"""Primary-hit AOVs on the GPU (vh_render_normals / vh_render_aovs) against the oracle, bit for bit:
normals, albedo, depth, object and material ids, and the 16-ray translucent coverage; the
reference's own invariants (crates/verion-gpu-vulkan/tests/gpu.rs:105-126) on the GPU output; the
zero-size error."""
import numpy as np
import pytest

import oracle
from verion_amd import pack, scenes
from verion_amd import scene as sc
from verion_amd.backend import BackendError, Renderer

from helpers import params

pytestmark = pytest.mark.gpu


def translucent_stack():
    s = sc.Scene()
    glassy = s.add_material(sc.BaseBsdf(base_color=(1.0, 1.0, 1.0), opacity=0.5))
    for z in (1.0, 0.5):
        s.add_instance(sc.Instance(s.add_mesh(scenes.plane_mesh(1.5, z)), glassy))
    s.add_instance(sc.Instance(s.add_mesh(scenes.plane_mesh(0.7, 0.0)), s.add_material(sc.lambert((0.4, 0.4, 0.4)))))
    s.camera = scenes.camera_at_z(3.0)
    return s


CASES = {
    "cornell": (scenes.cornell, 64, 48),
    "zoo": (scenes.zoo, 80, 60),
    "translucent": (translucent_stack, 37, 29),
    "procedural": (scenes.procedural, 96, 54),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_aovs_and_normals_match_the_oracle(ctx, name):
    make, w, h = CASES[name]
    s = make()
    p = pack.pack(s)
    r = Renderer(ctx, p)
    got = r.render_aovs(w, h)
    nrm = r.render_normals(w, h)
    o = oracle.OracleScene(p)
    par = params(p, w, h)
    want = o.aovs(par)
    for k in ("albedo", "depth", "object_id", "material_id", "coverage"):
        assert np.array_equal(got[k], want[k]), (name, k, np.argwhere(got[k] != want[k])[:5])
    assert np.array_equal(nrm, o.normals(par)), name
    if name == "translucent":
        assert set(np.unique(got["coverage"])) >= {0.75, 1.0}


def test_reference_first_hit_invariants(ctx):
    w = h = 64
    r = Renderer(ctx, scenes.cornell())
    aov = r.render_aovs(w, h)
    nrm = r.render_normals(w, h)
    hits = aov["object_id"] > 0
    assert hits.sum() > w * h // 2
    length = np.linalg.norm(nrm, axis=1)
    assert np.all(np.abs(length[hits] - 1.0) < 1e-3)
    assert np.all(aov["depth"][hits] > 0.0) and np.all(aov["material_id"][hits] > 0)
    assert np.all(length[~hits] == 0.0) and np.all(aov["depth"][~hits] == 0.0)


def test_zero_size_is_a_scene_error(ctx):
    r = Renderer(ctx, scenes.cornell())
    with pytest.raises(BackendError):
        r.render_aovs(0, 16)
    with pytest.raises(BackendError):
        r.render_normals(16, 0)
