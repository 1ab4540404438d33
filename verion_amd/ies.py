"""IES (IESNA LM-63) photometric profiles of point and spot lights (crates/verion-scene/src/ies.rs).

A profile is the file's candela distribution resampled to a one-degree grid: HORIZONTAL = 361 rows
(horizontal angle 0..360) of VERTICAL = 181 samples (vertical angle 0..180, 0 along the light's
-z axis), normalised so the peak is 1 (ies.rs:5-19). Only photometric type C is read (ies.rs:83-85);
vertical samples are interpolated linearly and are zero outside the measured range (ies.rs:93-106),
horizontal planes are folded by the symmetry the last plane angle declares (one plane: round; up to
90: quadrant; up to 180: bilateral; else full, ies.rs:108-121) and interpolated linearly
(ies.rs:122-133). `at` is the bilinear lookup the shader makes (pt_core.slang:199-216).

The packer appends a profile after the light table's link masks: the light's frame (three rows
x, y, z in world space) and the table, with word 15 of the light pointing one past its offset
(pack.rs:628-629, 679-686).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Sequence

import numpy as np

VERTICAL = 181
HORIZONTAL = 361


class IesError(ValueError):
    """A file the parser refuses: `kind` is one of truncated, number, photometric, empty (ies.rs:22-29)."""

    def __init__(self, kind: str, detail: str = ""):
        super().__init__(f"{kind}: {detail}" if detail else kind)
        self.kind = kind
        self.detail = detail


@dataclass(eq=False)
class IesProfile:
    table: np.ndarray  # float32, HORIZONTAL x VERTICAL, row = horizontal angle in degrees
    peak: float        # the file's largest candela value times its multiplier

    def scaled(self, factor: float) -> "IesProfile":
        return IesProfile((self.table * np.float32(factor)).astype(np.float32), self.peak)

    def at(self, theta: float, phi: float) -> float:
        """Relative intensity at vertical angle theta and horizontal angle phi (degrees), with the
        bilinear filter of the shader (ies.rs:147-159)."""
        th = min(max(np.float32(theta), np.float32(0.0)), np.float32(180.0))
        ph = np.float32(phi) % np.float32(360.0)
        i, j = int(math.floor(th)), int(math.floor(ph))
        i1, j1 = min(i + 1, VERTICAL - 1), min(j + 1, HORIZONTAL - 1)
        ft, fp = np.float32(th - np.float32(i)), np.float32(ph - np.float32(j))
        t = self.table
        a = t[j, i] * (np.float32(1) - ft) + t[j, i1] * ft
        b = t[j1, i] * (np.float32(1) - ft) + t[j1, i1] * ft
        return float(a * (np.float32(1) - fp) + b * fp)

    @staticmethod
    def parse(text: str) -> "IesProfile":
        lines = text.splitlines()
        tilt = None
        k = 0
        while k < len(lines):  # keywords up to the TILT line
            s = lines[k].strip()
            k += 1
            if s.startswith("TILT="):
                tilt = s[len("TILT="):]
                break
        if tilt is None:
            raise IesError("truncated", "no TILT line")
        tokens = [t for line in lines[k:] for t in line.replace(",", " ").split()]
        pos = 0

        def num() -> float:
            nonlocal pos
            if pos >= len(tokens):
                raise IesError("truncated", "the IES file ends early")
            t = tokens[pos]
            pos += 1
            try:
                return float(t)
            except ValueError:
                raise IesError("number", t) from None

        if tilt.strip() == "INCLUDE":
            num()  # lamp-to-luminaire geometry
            n = int(num())
            for _ in range(2 * n):
                num()  # tilt angles and multipliers
        num()  # lamps
        num()  # lumens per lamp
        multiplier = num()
        nv, nh = int(num()), int(num())
        photometric = int(num())
        for _ in range(7):  # units, width, length, height, ballast, version, watts
            num()
        if photometric != 1:
            raise IesError("photometric", str(photometric))
        if nv == 0 or nh == 0:
            raise IesError("empty")
        v = np.array([num() for _ in range(nv)], np.float64)
        h = np.array([num() for _ in range(nh)], np.float64)
        cd = np.array([num() * multiplier for _ in range(nv * nh)], np.float64).reshape(nh, nv)
        peak = float(cd.max(initial=0.0))
        if peak <= 0.0:
            raise IesError("empty")
        return IesProfile((_resample(v, h, cd) / peak).astype(np.float32), float(np.float32(peak)))


def _segment(xs: np.ndarray, x: float):
    """Index i in [1, n-1] of the segment [xs[i-1], xs[i]] holding x (the first sample above x,
    clamped) and the weight of xs[i]; 0 for a zero-length segment."""
    n = len(xs)
    i = int(np.searchsorted(xs, x, side="right"))
    i = min(max(i, 1), n - 1)
    w = (x - xs[i - 1]) / (xs[i] - xs[i - 1]) if xs[i] > xs[i - 1] else 0.0
    return i, w


def _resample(v: np.ndarray, h: np.ndarray, cd: np.ndarray) -> np.ndarray:
    nv, nh = len(v), len(h)
    last = h[-1]

    def vertical(plane: int, theta: float) -> float:
        if theta < v[0] or theta > v[-1]:
            return 0.0
        if nv == 1:
            return float(cd[plane, 0])
        i, w = _segment(v, theta)
        return float(cd[plane, i - 1] + (cd[plane, i] - cd[plane, i - 1]) * w)

    def fold(phi: float) -> float:
        if nh == 1:
            return 0.0
        if last <= 90.0:
            p = 360.0 - phi if phi > 180.0 else phi
            return 180.0 - p if p > 90.0 else p
        if last <= 180.0:
            return 360.0 - phi if phi > 180.0 else phi
        return phi

    out = np.zeros((HORIZONTAL, VERTICAL), np.float64)
    for ph in range(HORIZONTAL):
        for th in range(VERTICAL):
            if nh == 1:
                out[ph, th] = vertical(0, float(th))
                continue
            p = min(max(fold(float(ph)), h[0]), last)
            i, w = _segment(h, p)
            a, b = vertical(i - 1, float(th)), vertical(i, float(th))
            out[ph, th] = a + (b - a) * w
    return out


def isotropic() -> IesProfile:
    """A profile of 1 in every direction (its lights shine as without a profile)."""
    return IesProfile(np.ones((HORIZONTAL, VERTICAL), np.float32), 1.0)


def frame_rows(frame: Sequence[Sequence[float]]) -> List[float]:
    """The light's world frame as the nine floats the packer writes (rows x, y, z)."""
    f = np.asarray(frame, np.float32).reshape(3, 3)
    return [float(x) for x in f.reshape(-1)]
