"""Launch-shape autotuner of the single-accumulation wavefront: grid size of the grid-stride
extend/shade kernels and the depth-0 pixel order (tiles or scanlines).

Each candidate renders real sample passes of the caller's workload into a scratch accumulation;
the GPU seconds the library measures per pass pick the fastest. Images do not depend on the launch
shape (one path per pixel and pass, random streams keyed by pixel and sample), so the choice only
changes speed.
"""
from __future__ import annotations

from typing import List, Sequence, Tuple

CANDIDATES: Sequence[Tuple[int, int, int]] = (
    (2048, 8, 8), (4096, 8, 8), (1024, 8, 8), (2048, 0, 0), (4096, 16, 4), (2048, 32, 2), (8192, 8, 8), (4096, 0, 0))


def tune_wavefront(renderer, width: int, height: int, passes: int = 2,
                   candidates: Sequence[Tuple[int, int, int]] = CANDIDATES) -> Tuple[Tuple[int, int, int], List]:
    """Returns the fastest (grid_blocks, tile_width, tile_height) and the table of
    (candidate, ms per pass); leaves the renderer set to the fastest."""
    table = []
    for blocks, tw, th in candidates:
        renderer.set_wavefront_blocks(blocks)
        renderer.set_wavefront_tiles(tw, th)
        acc = renderer.new_accumulation(width, height)
        renderer.accumulate(acc, 1)  # first touch of the buffers
        before = acc.gpu_seconds()
        renderer.accumulate(acc, passes)
        table.append(((blocks, tw, th), (acc.gpu_seconds() - before) / passes * 1e3))
        acc.close()
    best = min(table, key=lambda e: e[1])[0]
    renderer.set_wavefront_blocks(best[0])
    renderer.set_wavefront_tiles(best[1], best[2])
    return best, table
