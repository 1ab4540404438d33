"""Image error and time-to-quality: relMSE and the log-log interpolation of scripts/eval_restir.sh:37-59.

relMSE = mean over pixels and channels of (x - r)^2 / (r^2 + eps), eps = 0.01 (the reference
reports mae/mean and PSNR only; this definition and its eps are this backend's).
"""
from __future__ import annotations

import math
from typing import Callable, List, Optional, Tuple

import numpy as np

EPS = 0.01


def relmse(x: np.ndarray, ref: np.ndarray, eps: float = EPS) -> float:
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.mean((x - ref) ** 2 / (ref * ref + eps)))


def time_to(points: List[Tuple[float, float]], target: float) -> Optional[float]:
    """First time the error reaches `target`, by log-log interpolation between the two ladder points
    around it (eval_restir.sh:48-55); the first point itself when it is already below; None if never."""
    if not points:
        return None
    if points[0][1] <= target:
        return points[0][0]
    for (t0, e0), (t1, e1) in zip(points, points[1:]):
        if e0 >= target >= e1 and e0 > 0 and e1 > 0 and t0 > 0 and t1 > 0:
            f = (math.log(e0) - math.log(target)) / (math.log(e0) - math.log(e1)) if e0 != e1 else 0.0
            return math.exp(math.log(t0) + f * (math.log(t1) - math.log(t0)))
    return None


def ladder(step: Callable[[int], float], read: Callable[[], np.ndarray], ref: np.ndarray, max_frames: int,
           stop_below: Optional[float] = None) -> List[Tuple[int, float, float]]:
    """Advances a render along a doubling frame ladder (1, 2, 4, ...): `step(k)` renders k more
    frames and returns their GPU seconds; after each rung the image `read()` is scored.
    Returns (frames, cumulative GPU seconds, relMSE) per rung."""
    out = []
    frames, secs, rung = 0, 0.0, 1
    while frames < max_frames:
        k = min(rung, max_frames) - frames
        secs += step(k)
        frames += k
        e = relmse(read(), ref)
        out.append((frames, secs, e))
        if stop_below is not None and e <= stop_below:
            break
        rung *= 2
    return out
