"""A render of the reference's Cornell box scene written as the reference CLI writes it
(crates/verion-render-cli/src/main.rs:359-441 -> exr_out.rs): beauty from the accumulation, the
lighting split, sample counts, the primary normals and AOVs; read back channel by channel."""
import os

import numpy as np
import pytest

from verion_amd import gltf, output
from verion_amd.backend import WAVEFRONT, Adaptive, Renderer

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_cornell_render_round_trips_through_exr(ctx, tmp_path):
    s, _ = gltf.load_gltf(os.path.join(GOLDEN, "cornell_box.gltf"))
    r = Renderer(ctx, s)
    r.set_integrator(WAVEFRONT)
    r.set_adaptive(Adaptive(0.3, 4))
    w, h = 48, 36
    acc = r.new_accumulation(w, h, lighting=True)
    r.accumulate(acc, 8)
    r.set_adaptive(None)
    info = output.RenderInfo(physical=True, samples=8, adaptive_threshold=0.3)
    f = output.frame_from_accumulation(r, acc, w, h, info)
    p = os.path.join(tmp_path, "cornell.exr")
    output.write_exr(p, f)
    attrs, ch = output.read_exr(p)
    mean = acc.mean().reshape(h, w, 3)
    assert np.array_equal(ch["R"], mean[:, :, 0]) and np.array_equal(ch["B"], mean[:, :, 2])
    aov = r.render_aovs(w, h)
    assert np.array_equal(ch["ObjectId"].ravel(), aov["object_id"])
    assert np.array_equal(ch["Z"].ravel(), aov["depth"].astype(np.float32))
    assert np.array_equal(ch["N.Y"].ravel(), r.render_normals(w, h)[:, 1])
    assert np.array_equal(ch["SampleCount"].ravel(), acc.sample_counts())
    assert "DirectDiffuse.R" in ch and "Visibility.V" in ch
    assert attrs["verion/physical"] == "false" and attrs["verion/adaptiveThreshold"] == np.float32(0.3)
