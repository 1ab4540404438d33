"""Outputs around the path: the multi-channel EXR file and the metrics JSON of a render.

Restates crates/verion-render-io/src/exr_out.rs (channels, their names and types, the
`verion/*` header attributes, chromaticities, the BadFrame checks), the linear output transforms
of crates/verion-scene/src/color.rs (Rec.709 <-> ACEScg primaries, Reinhard, sRGB OETF), and the
metrics line of crates/verion-render-cli/src/main.rs:458-461 (its keys and its JSON string
escaping, :211-223). The reference writes EXR through the `exr` crate; this module writes the
OpenEXR 2.0 single-part scanline layout itself (uncompressed, one scanline per chunk), and
`read_exr` reads that layout back (no OpenEXR library is installed here).

A frame from the HIP backend: `frame_from_accumulation(renderer, acc, ...)` gathers the beauty
(`Accumulation.mean`), sample counts, lighting split and the primary AOVs the way the reference CLI
does before it writes.
"""
from __future__ import annotations

import math
import struct
from dataclasses import dataclass, field
from typing import Dict, Optional, Sequence, Tuple

import numpy as np

from . import readback

# ---------------------------------------------------------------- color (color.rs)

LINEAR_REC709 = "LinearRec709"
ACES_CG = "AcesCg"

# color.rs:18-28: Rec.709 -> ACEScg and its inverse (Bradford-adapted from D65 to D60)
_TO_ACES = np.array([[0.613097, 0.339523, 0.047379], [0.070194, 0.916354, 0.013452], [0.020616, 0.109570, 0.869815]],
                    np.float32)
_TO_709 = np.array([[1.705052, -0.621792, -0.083258], [-0.130257, 1.140805, -0.010548], [-0.024004, -0.128969, 1.152972]],
                   np.float32)
_CHROMA = {LINEAR_REC709: ((0.64, 0.33), (0.30, 0.60), (0.15, 0.06), (0.3127, 0.3290)),
           ACES_CG: ((0.713, 0.293), (0.165, 0.830), (0.128, 0.044), (0.32168, 0.33767))}
_LUMA = {LINEAR_REC709: (0.2126, 0.7152, 0.0722), ACES_CG: (0.272229, 0.674082, 0.053689)}


def convert(rgb: np.ndarray, frm: str, to: str) -> np.ndarray:
    """ColorSpace::convert (color.rs:35-42): linear RGB between spaces, negatives kept."""
    rgb = np.asarray(rgb, np.float32)
    if frm == LINEAR_REC709 and to == ACES_CG:
        m = _TO_ACES
    elif frm == ACES_CG and to == LINEAR_REC709:
        m = _TO_709
    else:
        return rgb.copy()
    # row-wise in float32, summed in the order row[0] c0 + row[1] c1 + row[2] c2
    return (rgb[..., 0:1] * m[:, 0] + rgb[..., 1:2] * m[:, 1] + rgb[..., 2:3] * m[:, 2]).astype(np.float32)


def chromaticities(space: str):
    return _CHROMA[space]


def luminance(rgb: np.ndarray, space: str) -> np.ndarray:
    w = _LUMA[space]
    rgb = np.asarray(rgb, np.float32)
    return rgb[..., 0] * np.float32(w[0]) + rgb[..., 1] * np.float32(w[1]) + rgb[..., 2] * np.float32(w[2])


@dataclass(frozen=True)
class OutputTransform:
    """color.rs:84-125: a target space, then tone map ("None" | "Reinhard"), then OETF ("Linear" | "Srgb")."""

    target: str = LINEAR_REC709
    tone_map: str = "None"
    oetf: str = "Linear"

    @staticmethod
    def linear(target: str = LINEAR_REC709) -> "OutputTransform":
        return OutputTransform(target, "None", "Linear")

    @staticmethod
    def display() -> "OutputTransform":
        return OutputTransform(LINEAR_REC709, "Reinhard", "Srgb")

    def apply(self, working: str, rgb: np.ndarray) -> np.ndarray:
        c = convert(rgb, working, self.target)
        if self.tone_map == "Reinhard":
            scale = np.float32(1.0) / (np.float32(1.0) + np.maximum(luminance(c, self.target), np.float32(0.0)))
            c = (c * scale[..., None]).astype(np.float32)
        if self.oetf == "Srgb":
            v = np.clip(c, 0.0, 1.0).astype(np.float32)
            c = np.where(v <= 0.0031308, v * np.float32(12.92),
                         np.float32(1.055) * np.power(v, np.float32(1.0 / 2.4)) - np.float32(0.055)).astype(np.float32)
        return c

    def debug(self) -> str:
        """The Rust Debug form the reference records as `verion/outputTransform`."""
        return f"OutputTransform {{ target: {self.target}, tone_map: {self.tone_map}, oetf: {self.oetf} }}"


# ---------------------------------------------------------------- frame (exr_out.rs:9-78)

@dataclass
class GradeInfo:
    exposure: float
    saturation: float
    white_balance: Tuple[float, float, float]


@dataclass
class RenderInfo:
    physical: bool = True
    samples: int = 0
    stylization_revision: int = 0
    grade: Optional[GradeInfo] = None
    denoiser: Optional[str] = None
    adaptive_threshold: Optional[float] = None
    restir: Optional[str] = None  # "di" or "di+gi"


@dataclass
class ExrLighting:
    direct_diffuse: np.ndarray
    indirect_diffuse: np.ndarray
    specular: np.ndarray
    transmission: np.ndarray
    emission: np.ndarray
    visibility: np.ndarray


@dataclass
class ExrAovs:
    albedo: Optional[np.ndarray] = None
    depth: Optional[np.ndarray] = None
    object_id: Optional[np.ndarray] = None
    material_id: Optional[np.ndarray] = None
    sample_count: Optional[np.ndarray] = None
    lighting: Optional[ExrLighting] = None


@dataclass
class ExrFrame:
    width: int
    height: int
    beauty: np.ndarray  # (w h, 3), row-major from the top-left, radiance in working_space
    working_space: str = LINEAR_REC709
    normal: Optional[np.ndarray] = None
    aovs: ExrAovs = field(default_factory=ExrAovs)
    info: RenderInfo = field(default_factory=RenderInfo)


class BadFrame(ValueError):
    """ExrError::BadFrame: zero size, or a pixel buffer that does not match it."""


# ---------------------------------------------------------------- OpenEXR 2.0 layout

_UINT, _HALF, _FLOAT = 0, 1, 2
_MAGIC = 20000630


def _attr(name: str, typ: str, data: bytes) -> bytes:
    return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(data)) + data


def _chlist(channels: Sequence[Tuple[str, int]]) -> bytes:
    out = b""
    for name, ptype in channels:
        out += name.encode() + b"\0" + struct.pack("<iB3xii", ptype, 0, 1, 1)  # pLinear 0, reserved, x/y sampling 1
    return out + b"\0"


def write_exr(path: str, frame: ExrFrame, transform: OutputTransform = OutputTransform.linear()) -> None:
    """write_exr (exr_out.rs:104-275)."""
    w, h = int(frame.width), int(frame.height)
    size = w * h
    if size == 0:
        raise BadFrame("size must be non-zero")
    a = frame.aovs

    def n(x):
        return len(np.asarray(x))

    parts = []
    if a.lighting is not None:
        l = a.lighting
        parts = [l.direct_diffuse, l.indirect_diffuse, l.specular, l.transmission, l.emission]
    if (n(frame.beauty) != size or (frame.normal is not None and n(frame.normal) != size)
            or any(x is not None and n(x) != size for x in (a.albedo, a.depth, a.object_id, a.material_id, a.sample_count))
            or any(n(p) != size for p in parts) or (a.lighting is not None and n(a.lighting.visibility) != size)):
        raise BadFrame("pixel count does not match width x height")

    rgb3 = lambda x: np.asarray(x, np.float32).reshape(size, 3)
    chans: Dict[str, np.ndarray] = {}
    beauty = transform.apply(frame.working_space, rgb3(frame.beauty))
    for i, c in enumerate("RGB"):
        chans[c] = beauty[:, i]
    if frame.normal is not None:
        nrm = rgb3(frame.normal)
        for i, c in enumerate(("N.X", "N.Y", "N.Z")):
            chans[c] = nrm[:, i]
    primaries = OutputTransform.linear(transform.target)
    if a.albedo is not None:
        alb = primaries.apply(frame.working_space, rgb3(a.albedo))
        for i, c in enumerate("RGB"):
            chans[f"Albedo.{c}"] = alb[:, i]
    if a.lighting is not None:
        for name, part in zip(("DirectDiffuse", "IndirectDiffuse", "Specular", "Transmission", "Emission"), parts):
            p = primaries.apply(frame.working_space, rgb3(part))
            for i, c in enumerate("RGB"):
                chans[f"{name}.{c}"] = p[:, i]
        chans["Visibility.V"] = np.asarray(a.lighting.visibility, np.float32).reshape(size)
    if a.depth is not None:
        chans["Z"] = np.asarray(a.depth, np.float32).reshape(size)
    for name, ids in (("ObjectId", a.object_id), ("MaterialId", a.material_id), ("SampleCount", a.sample_count)):
        if ids is not None:
            chans[name] = np.asarray(ids, np.uint32).reshape(size)

    info = frame.info
    untouched = transform.tone_map == "None" and transform.oetf == "Linear"
    physical = info.physical and untouched and info.denoiser is None and info.adaptive_threshold is None and info.restir is None
    other = [("verion/physical", "string", ("true" if physical else "false").encode()),
             ("verion/samples", "int", struct.pack("<i", min(int(info.samples), 2 ** 31 - 1))),
             ("verion/stylizationRevision", "string", str(int(info.stylization_revision)).encode())]
    if info.denoiser is not None:
        other.append(("verion/denoiser", "string", info.denoiser.encode()))
    if info.adaptive_threshold is not None:
        other.append(("verion/adaptiveThreshold", "float", struct.pack("<f", info.adaptive_threshold)))
    if info.restir is not None:
        other.append(("verion/restir", "string", info.restir.encode()))
    if info.grade is not None:
        g = info.grade
        other.append(("verion/exposure", "float", struct.pack("<f", g.exposure)))
        other.append(("verion/saturation", "float", struct.pack("<f", g.saturation)))
        other.append(("verion/whiteBalance", "v3f", struct.pack("<3f", *g.white_balance)))
    other.append(("verion/workingSpace", "string", frame.working_space.encode()))
    other.append(("verion/outputTransform", "string", transform.debug().encode()))

    names = sorted(chans)  # channel lists are sorted by name (AnyChannels::sort)
    types = [(nm, _UINT if chans[nm].dtype == np.uint32 else _FLOAT) for nm in names]
    chroma = chromaticities(transform.target)
    header = [
        ("channels", "chlist", _chlist(types)),
        ("chromaticities", "chromaticities", struct.pack("<8f", *[v for xy in chroma for v in xy])),
        ("compression", "compression", bytes([0])),  # NO_COMPRESSION
        ("dataWindow", "box2i", struct.pack("<4i", 0, 0, w - 1, h - 1)),
        ("displayWindow", "box2i", struct.pack("<4i", 0, 0, w - 1, h - 1)),
        ("lineOrder", "lineOrder", bytes([0])),  # INCREASING_Y
        ("pixelAspectRatio", "float", struct.pack("<f", 1.0)),
        ("screenWindowCenter", "v2f", struct.pack("<2f", 0.0, 0.0)),
        ("screenWindowWidth", "float", struct.pack("<f", 1.0)),
    ] + other
    head = struct.pack("<ii", _MAGIC, 2) + b"".join(_attr(nm, t, d) for nm, t, d in header) + b"\0"
    planes = [np.ascontiguousarray(chans[nm]).astype(chans[nm].dtype.newbyteorder("<")).reshape(h, w) for nm in names]
    rows = [b"".join(pl[y].tobytes() for pl in planes) for y in range(h)]
    table_end = len(head) + 8 * h
    offsets, pos = [], table_end
    for row in rows:
        offsets.append(pos)
        pos += 8 + len(row)
    with open(path, "wb") as f:
        f.write(head)
        f.write(struct.pack(f"<{h}Q", *offsets))
        for y, row in enumerate(rows):
            f.write(struct.pack("<ii", y, len(row)))
            f.write(row)


def _parse_attr_value(typ: str, data: bytes):
    if typ == "string":
        return data.decode()
    if typ == "int":
        return struct.unpack("<i", data)[0]
    if typ == "float":
        return struct.unpack("<f", data)[0]
    if typ == "v2f":
        return struct.unpack("<2f", data)
    if typ == "v3f":
        return struct.unpack("<3f", data)
    if typ == "box2i":
        return struct.unpack("<4i", data)
    if typ in ("compression", "lineOrder"):
        return data[0]
    if typ == "chromaticities":
        v = struct.unpack("<8f", data)
        return tuple((v[2 * k], v[2 * k + 1]) for k in range(4))
    if typ == "chlist":
        out, i = [], 0
        while data[i] != 0:
            j = data.index(b"\0", i)
            name = data[i:j].decode()
            ptype, plin, xs, ys = struct.unpack("<iB3xii", data[j + 1:j + 17])
            out.append((name, ptype, xs, ys))
            i = j + 17
        return out
    return data


def read_exr(path: str):
    """Reads a single-part scanline OpenEXR file without compression (what write_exr writes):
    returns (attributes, channels), channels by name as float32 or uint32 arrays of h x w."""
    with open(path, "rb") as f:
        buf = f.read()
    magic, version = struct.unpack("<ii", buf[:8])
    if magic != _MAGIC:
        raise ValueError("not an OpenEXR file")
    if version & 0xFF != 2 or version & 0x1E00:
        raise ValueError("only single-part scanline files are read")
    attrs, i = {}, 8
    while buf[i] != 0:
        j = buf.index(b"\0", i)
        name = buf[i:j].decode()
        k = buf.index(b"\0", j + 1)
        typ = buf[j + 1:k].decode()
        size = struct.unpack("<i", buf[k + 1:k + 5])[0]
        attrs[name] = (typ, _parse_attr_value(typ, buf[k + 5:k + 5 + size]))
        i = k + 5 + size
    i += 1
    if attrs["compression"][1] != 0:
        raise ValueError("only uncompressed files are read")
    x0, y0, x1, y1 = attrs["dataWindow"][1]
    w, h = x1 - x0 + 1, y1 - y0 + 1
    chl = attrs["channels"][1]
    offsets = struct.unpack(f"<{h}Q", buf[i:i + 8 * h])
    chans = {nm: np.zeros((h, w), np.uint32 if pt == _UINT else (np.float16 if pt == _HALF else np.float32)) for nm, pt, _, _ in chl}
    for off in offsets:
        y, size = struct.unpack("<ii", buf[off:off + 8])
        row = buf[off + 8:off + 8 + size]
        p = 0
        for nm, pt, _, _ in chl:
            nb = (2 if pt == _HALF else 4) * w
            chans[nm][y - y0] = np.frombuffer(row[p:p + nb], chans[nm].dtype)
            p += nb
    return {k: v for k, (t, v) in attrs.items()}, chans


# ---------------------------------------------------------------- metrics (main.rs:440-470)

def json_str(s: str) -> str:
    """main.rs:211-223: quotes and backslashes escaped, control characters as \\u00XX."""
    out = ['"']
    for c in s:
        if c == '"':
            out.append('\\"')
        elif c == "\\":
            out.append("\\\\")
        elif ord(c) < 0x20:
            out.append(f"\\u{ord(c):04x}")
        else:
            out.append(c)
    out.append('"')
    return "".join(out)


METRIC_KEYS = ("backend", "scene", "output", "width", "height", "spp", "mode", "integrator", "adaptive", "restir",
               "mean_spp", "wavefront_bytes", "restir_bytes", "gpu_budget_bytes", "gpu_peak_bytes", "upload_bytes",
               "load_s", "build_s", "render_s", "gpu_s", "extend_s", "shade_s", "aov_s", "denoise_s", "stylize_s",
               "write_s", "total_s")


def _num(v) -> str:
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, int):
        return str(v)
    v = float(v)
    if math.isfinite(v) and v == int(v) and abs(v) < 1e15:
        return str(int(v))  # Rust prints 2.0_f64 as "2"
    return repr(v)


def metrics_json(m: dict) -> str:
    """One metrics line with the reference's keys in its order (main.rs:458-461): `backend`,
    `scene`, `output`, `mode` strings, `integrator` the Rust Debug name ("Megakernel" /
    "Wavefront"), `adaptive` the threshold or null, `restir` a JSON string or null, `extend_s` and
    `shade_s` only when the stages were timed."""
    parts = []
    for k in METRIC_KEYS:
        if k in ("extend_s", "shade_s") and m.get(k) is None:
            continue
        v = m.get(k)
        if k in ("backend", "scene", "output", "mode"):
            parts.append(f'"{k}":{json_str(str(v))}')
        elif k == "integrator":
            parts.append(f'"integrator":"{v}"')
        elif k == "adaptive":
            parts.append(f'"adaptive":{"null" if v is None else _num(v)}')
        elif k == "restir":
            parts.append(f'"restir":{"null" if v is None else json_str(str(v))}')
        else:
            parts.append(f'"{k}":{_num(v if v is not None else 0)}')
    return "{" + ",".join(parts) + "}\n"


# ---------------------------------------------------------------- from the HIP backend

def frame_from_accumulation(renderer, acc, width: int, height: int, info: RenderInfo, normals: bool = True,
                            aovs: bool = True) -> ExrFrame:
    """The frame the reference CLI writes (main.rs:359-441): the accumulation's mean as beauty,
    its per-pixel sample counts under adaptive sampling, its lighting split when it has one, and
    the primary-ray normals and AOVs rendered by the same renderer."""
    f = ExrFrame(width, height, acc.mean().astype(np.float32), info=info)
    if normals:
        f.normal = renderer.render_normals(width, height)
    if aovs:
        a = renderer.render_aovs(width, height)
        f.aovs.albedo = a["albedo"]
        f.aovs.depth = a["depth"]
        f.aovs.object_id = a["object_id"]
        f.aovs.material_id = a["material_id"]
    counts = acc.sample_counts()
    if counts is not None:
        f.aovs.sample_count = np.asarray(counts, np.uint32)
    sums = acc.lighting_sums()
    if sums is not None:
        st = acc.stats()
        l = readback.lighting(sums, readback.inverse_samples(st, width * height, acc.samples()))
        f.aovs.lighting = ExrLighting(l.direct_diffuse, l.indirect_diffuse, l.specular, l.transmission, l.emission,
                                      l.visibility)
    return f
