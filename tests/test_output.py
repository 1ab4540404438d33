"""EXR and metrics output, restating the reference's own tests: crates/verion-render-io/src/
exr_out.rs:330-560 (channels, metadata, primaries, bad frames, lighting parts),
crates/verion-scene/src/color.rs:140-196 (conversions, display transform) and
crates/verion-render-cli/src/main.rs:546-549 (JSON escaping). The reader is pinned on a file
assembled byte by byte from the OpenEXR 2.0 layout (magic, version, attributes, offset table,
scanline chunks), so a write -> read round trip checks the writer against the format, not against
itself alone."""
import json
import os
import struct

import numpy as np
import pytest

from verion_amd import output as out


def frame(beauty, normal=None):
    b = np.asarray(beauty, np.float32)
    return out.ExrFrame(2, 1, b, normal=None if normal is None else np.asarray(normal, np.float32),
                        info=out.RenderInfo(physical=True, samples=64, stylization_revision=3))


def write_read(tmp_path, f, t=out.OutputTransform.linear(), name="a.exr"):
    p = os.path.join(tmp_path, name)
    out.write_exr(p, f, t)
    return out.read_exr(p)


def test_reader_parses_a_hand_assembled_file(tmp_path):
    def attr(name, typ, data):
        return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(data)) + data
    chl = b"G\0" + struct.pack("<iB3xii", 2, 0, 1, 1) + b"R\0" + struct.pack("<iB3xii", 2, 0, 1, 1) + b"\0"
    head = struct.pack("<ii", 20000630, 2) + attr("channels", "chlist", chl) + attr("compression", "compression", b"\0") \
        + attr("dataWindow", "box2i", struct.pack("<4i", 0, 0, 1, 1)) + attr("displayWindow", "box2i", struct.pack("<4i", 0, 0, 1, 1)) \
        + attr("lineOrder", "lineOrder", b"\0") + attr("pixelAspectRatio", "float", struct.pack("<f", 1.0)) \
        + attr("screenWindowCenter", "v2f", struct.pack("<2f", 0, 0)) + attr("screenWindowWidth", "float", struct.pack("<f", 1.0)) + b"\0"
    rows = [struct.pack("<2f", 1.0, 2.0) + struct.pack("<2f", 3.0, 4.0), struct.pack("<2f", 5.0, 6.0) + struct.pack("<2f", 7.0, 8.0)]
    base = len(head) + 16
    offs = [base, base + 8 + len(rows[0])]
    data = head + struct.pack("<2Q", *offs) + b"".join(struct.pack("<ii", y, len(r)) + r for y, r in enumerate(rows))
    p = os.path.join(tmp_path, "hand.exr")
    open(p, "wb").write(data)
    attrs, ch = out.read_exr(p)
    assert attrs["dataWindow"] == (0, 0, 1, 1) and [c[0] for c in attrs["channels"]] == ["G", "R"]
    assert ch["G"].tolist() == [[1.0, 2.0], [5.0, 6.0]] and ch["R"].tolist() == [[3.0, 4.0], [7.0, 8.0]]


def test_writer_follows_the_layout(tmp_path):
    p = os.path.join(tmp_path, "l.exr")
    out.write_exr(p, frame([[1.0, 2.0, 3.0], [0.25, 0.5, 4.0]]))
    b = open(p, "rb").read()
    assert struct.unpack("<ii", b[:8]) == (20000630, 2)
    assert b[8:17] == b"channels\0"
    # the offset table points at scanline chunks that start with their y and size
    attrs, _ = out.read_exr(p)
    chunk_start = len(b) - (8 + 3 * 2 * 4)  # one scanline: y, size, then B, G, R of two pixels
    assert struct.unpack("<ii", b[chunk_start:chunk_start + 8]) == (0, 24)
    assert struct.unpack("<Q", b[chunk_start - 8:chunk_start])[0] == chunk_start
    assert struct.unpack("<6f", b[chunk_start + 8:]) == (3.0, 4.0, 2.0, 0.5, 1.0, 0.25)
    assert [c[0] for c in attrs["channels"]] == ["B", "G", "R"], "channels are sorted by name"


def test_writes_beauty_normal_and_metadata(tmp_path):
    attrs, ch = write_read(tmp_path, frame([[1.0, 2.0, 3.0], [0.25, 0.5, 4.0]], [[0.0, 0.0, 1.0], [1.0, 0.0, 0.0]]))
    assert ch["R"].ravel().tolist() == [1.0, 0.25]
    assert ch["B"].ravel().tolist() == [3.0, 4.0]
    assert ch["N.Z"].ravel().tolist() == [1.0, 0.0]
    assert attrs["verion/physical"] == "true"
    assert attrs["verion/samples"] == 64
    assert attrs["verion/stylizationRevision"] == "3"
    assert attrs["chromaticities"][0] == (np.float32(0.64), np.float32(0.33))
    assert attrs["verion/workingSpace"] == "LinearRec709"
    assert attrs["verion/outputTransform"] == "OutputTransform { target: LinearRec709, tone_map: None, oetf: Linear }"


def test_records_the_color_grade_only_when_one_was_applied(tmp_path):
    beauty = [[0.0] * 3] * 2
    attrs, _ = write_read(tmp_path, frame(beauty))
    assert "verion/exposure" not in attrs and "verion/denoiser" not in attrs
    f = frame(beauty)
    f.info.denoiser = "oidn"
    attrs, _ = write_read(tmp_path, f)
    assert attrs["verion/denoiser"] == "oidn" and attrs["verion/physical"] == "false"
    f = frame(beauty)
    f.info.adaptive_threshold = 0.02
    attrs, _ = write_read(tmp_path, f)
    assert attrs["verion/adaptiveThreshold"] == np.float32(0.02) and attrs["verion/physical"] == "false"
    f = frame(beauty)
    f.info.restir = "di+gi"
    attrs, _ = write_read(tmp_path, f)
    assert attrs["verion/restir"] == "di+gi" and attrs["verion/physical"] == "false"
    f = frame(beauty)
    f.info.grade = out.GradeInfo(-1.0, 0.8, (1.1, 1.0, 0.9))
    attrs, _ = write_read(tmp_path, f)
    assert attrs["verion/exposure"] == -1.0 and attrs["verion/saturation"] == np.float32(0.8)
    assert attrs["verion/whiteBalance"] == tuple(np.float32(v) for v in (1.1, 1.0, 0.9))


def test_writes_albedo_depth_and_object_id(tmp_path):
    f = frame([[0.0] * 3] * 2)
    f.aovs = out.ExrAovs(albedo=np.array([[1.0, 0.0, 0.0], [0.5, 0.5, 0.5]], np.float32), depth=np.array([3.0, 0.0]),
                         object_id=np.array([7, 0]), material_id=np.array([3, 0]), sample_count=np.array([16, 64]))
    attrs, ch = write_read(tmp_path, f)
    assert ch["Albedo.R"].ravel().tolist() == [1.0, 0.5]
    assert ch["Z"].ravel().tolist() == [3.0, 0.0]
    assert ch["ObjectId"].ravel().tolist() == [7, 0] and ch["ObjectId"].dtype == np.uint32
    assert ch["MaterialId"].ravel().tolist() == [3, 0]
    assert ch["SampleCount"].ravel().tolist() == [16, 64]
    types = {c[0]: c[1] for c in attrs["channels"]}
    assert types["ObjectId"] == 0 and types["Z"] == 2, "ids are 32-bit unsigned integer channels"
    f.aovs.depth = np.array([1.0])
    with pytest.raises(out.BadFrame):
        out.write_exr(os.path.join(tmp_path, "bad.exr"), f)


def test_converts_primaries_and_marks_display_output_non_physical(tmp_path):
    f = frame([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    attrs, ch = write_read(tmp_path, f, out.OutputTransform.linear(out.ACES_CG))
    assert abs(ch["R"].ravel()[0] - 0.613097) < 1e-6
    assert attrs["chromaticities"][3] == (np.float32(0.32168), np.float32(0.33767))
    attrs, _ = write_read(tmp_path, f, out.OutputTransform.display())
    assert attrs["verion/physical"] == "false"


def test_rejects_mismatched_frames(tmp_path):
    p = os.path.join(tmp_path, "bad.exr")
    one = np.zeros((1, 3), np.float32)
    with pytest.raises(out.BadFrame):
        out.write_exr(p, out.ExrFrame(2, 1, one))
    with pytest.raises(out.BadFrame):
        out.write_exr(p, out.ExrFrame(2, 1, np.zeros((2, 3), np.float32), normal=one))
    with pytest.raises(out.BadFrame):
        out.write_exr(p, out.ExrFrame(0, 1, np.zeros((0, 3), np.float32)))


def test_writes_lighting_parts(tmp_path):
    part = lambda v: np.array([[v, 2 * v, 3 * v]] * 2, np.float32)
    f = frame([[0.0] * 3] * 2)
    f.aovs.lighting = out.ExrLighting(part(1.0), part(2.0), part(3.0), part(4.0), part(5.0), np.array([0.25, 1.0], np.float32))
    _, ch = write_read(tmp_path, f)
    assert ch["DirectDiffuse.G"].ravel().tolist() == [2.0, 2.0]
    assert ch["IndirectDiffuse.R"].ravel().tolist() == [2.0, 2.0]
    assert ch["Specular.B"].ravel().tolist() == [9.0, 9.0]
    assert ch["Transmission.R"].ravel().tolist() == [4.0, 4.0]
    assert ch["Emission.R"].ravel().tolist() == [5.0, 5.0]
    assert ch["Visibility.V"].ravel().tolist() == [0.25, 1.0]
    f.aovs.lighting.visibility = np.array([1.0], np.float32)
    with pytest.raises(out.BadFrame):
        out.write_exr(os.path.join(tmp_path, "bad.exr"), f)


def test_color_spaces():
    # color.rs:148-196
    for a in (out.LINEAR_REC709, out.ACES_CG):
        for b in (out.LINEAR_REC709, out.ACES_CG):
            assert np.allclose(out.convert([1.0, 1.0, 1.0], a, b), 1.0, atol=1e-4)
            c = np.array([0.2, 0.5, 0.9], np.float32)
            assert np.allclose(out.convert(out.convert(c, a, b), b, a), c, atol=1e-4)
        assert abs(out.luminance(np.ones(3, np.float32), a) - 1.0) < 1e-4
    red = out.convert([1.0, 0.0, 0.0], out.LINEAR_REC709, out.ACES_CG)
    assert np.allclose(red, [0.613097, 0.070194, 0.020616], atol=1e-6)
    d = out.OutputTransform.display()
    assert np.allclose(d.apply(out.LINEAR_REC709, np.ones(3, np.float32)), 0.7354, atol=1e-3)
    assert d.apply(out.LINEAR_REC709, np.zeros(3, np.float32)).tolist() == [0.0, 0.0, 0.0]
    c = np.array([0.3, 0.6, 0.1], np.float32)
    assert out.OutputTransform.linear().apply(out.LINEAR_REC709, c).tolist() == c.tolist()


def test_metrics_line_has_the_reference_keys_and_escaping():
    assert out.json_str('a"b\\c\n') == '"a\\"b\\\\c\\u000a"'
    m = dict(backend="hip", scene='s"1.gltf', output="o.exr", width=2, height=1, spp=4, mode="reference",
             integrator="Wavefront", adaptive=None, restir="di+gi", mean_spp=4.0, wavefront_bytes=10, restir_bytes=0,
             gpu_budget_bytes=1, gpu_peak_bytes=2, upload_bytes=3, load_s=0.5, build_s=0.1, render_s=0.25, gpu_s=0.125,
             extend_s=0.05, shade_s=0.07, aov_s=0.0, denoise_s=0.0, stylize_s=0.0, write_s=0.01, total_s=1.0)
    line = out.metrics_json(m)
    d = json.loads(line)
    assert list(d) == list(out.METRIC_KEYS)
    assert d["scene"] == 's"1.gltf' and d["adaptive"] is None and d["restir"] == "di+gi" and d["gpu_s"] == 0.125
    m["extend_s"] = m["shade_s"] = None
    assert "extend_s" not in json.loads(out.metrics_json(m))
