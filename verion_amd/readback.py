"""Running sums -> the images `Accumulation` returns: a restatement of crates/verion-gpu/src/readback.rs
(`inverse_samples` :18-29, `mean` :31-38, `sample_counts` :41-48, `active_pixels` :51-61,
`lighting` :64-137). The sums arrive as arrays read through the C ABI (FP64 film and statistics,
FP32 lighting split)."""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np

LIGHTING_STRIDE = 65  # readback.rs:7
SLOT_WORDS = 12       # readback.rs:9
LIGHT_SLOTS = 4


def inverse_samples(stats: Optional[np.ndarray], pixels: int, samples: int) -> np.ndarray:
    if stats is not None:
        return 1.0 / np.maximum(np.asarray(stats).reshape(-1, 3)[:, 0], 1.0)
    return np.full(pixels, 1.0 / max(samples, 1))


def mean(sums: np.ndarray, inverse: np.ndarray) -> np.ndarray:
    return np.asarray(sums).reshape(-1, 3) * np.asarray(inverse)[:, None]


def sample_counts(stats: np.ndarray) -> np.ndarray:
    return np.asarray(stats).reshape(-1, 3)[:, 0].astype(np.uint32)


def active_pixels(stats: Optional[np.ndarray], adaptive, pixels: int) -> int:
    if stats is None or adaptive is None:
        return pixels
    s = np.asarray(stats).reshape(-1, 3)
    return int(sum(adaptive.active(float(a), float(b), float(c)) for a, b, c in s))


@dataclass
class LightingAovs:
    """lib.rs:182-238."""

    direct_diffuse: np.ndarray
    indirect_diffuse: np.ndarray
    specular: np.ndarray
    transmission: np.ndarray
    emission: np.ndarray
    visibility: np.ndarray
    light_u: np.ndarray         # (pixels, 4, 3)
    light_response: np.ndarray  # (pixels, 4)
    light_s: np.ndarray         # (pixels, 4, 3)
    light_dir: np.ndarray       # (pixels, 4, 3)
    light_lit: np.ndarray       # (pixels, 4)


def lighting(data: np.ndarray, inverse: np.ndarray) -> LightingAovs:
    r = np.asarray(data, np.float32).reshape(-1, LIGHTING_STRIDE)
    inv = np.asarray(inverse, np.float32)[:, None]
    part = lambda k: r[:, 3 * k:3 * k + 3] * inv
    slots = r[:, 17:17 + SLOT_WORDS * LIGHT_SLOTS].reshape(-1, LIGHT_SLOTS, SLOT_WORDS)
    u = slots[:, :, 0:3] * inv[:, :, None]
    s = slots[:, :, 5:8] * inv[:, :, None]
    w = slots[:, :, 3] + np.float32(0.2126) * slots[:, :, 5] + np.float32(0.7152) * slots[:, :, 6] + np.float32(0.0722) * slots[:, :, 7]
    with np.errstate(divide="ignore", invalid="ignore"):
        response = np.where(slots[:, :, 3] > 0, slots[:, :, 4] / slots[:, :, 3], 0.0).astype(np.float32)
        direction = np.where((w > 0)[:, :, None], slots[:, :, 8:11] / w[:, :, None], 0.0).astype(np.float32)
        lit = np.where(w > 0, slots[:, :, 11] / w, 0.0).astype(np.float32)
        vis = np.where(r[:, 16] > 0, r[:, 15] / r[:, 16], 1.0).astype(np.float32)
    return LightingAovs(part(0), part(1), part(2), part(3), part(4), vis, u, response, s, direction, lit)
