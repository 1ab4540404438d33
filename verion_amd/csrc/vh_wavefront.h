// vh_wavefront.h - slot-indexed wavefront path state (wf_state.slang:35-92), shared by the
// per-accumulation wavefront (vh_render.hip) and the batched one (vh_batch.hip).
//
// State word f of slot i lives at s[f * n + i] (structure of arrays over n slots): every stage
// reads and writes contiguous runs of each word.
#pragma once

#include "vh_shade.h"

namespace vh {

constexpr uint32_t kBaseWords = 17; // 16 words of wf_state.slang:75 + the pixel id; WF_SPLIT (17) more follow

// Rays per wave of a trace kernel. A wave's traversal lasts as long as its slowest ray, so a queue
// too short to fill `waves` waves of 64 rays is spread over that many waves, fewer rays to each;
// waves == 0: 64. Wave-uniform.
__device__ __forceinline__ uint32_t raysPerWave(uint32_t count, uint32_t waves)
{
    if (waves == 0)
        return 64u;
    uint32_t per = (count + waves - 1) / waves;
    return per < 1u ? 1u : (per > 64u ? 64u : per);
}

// Depth 0 takes pixels in tile order rather than scanline order, so that a wave's 64 camera rays
// (and their first bounces) come from one compact patch and share BVH nodes. The image does not
// depend on the order: a pixel takes one path per sample pass.
// Local index i (0 .. w h - 1) of the depth-0 order of a w x h image -> pixel. Bands of tileH rows, each cut
// into tiles of tileW columns (the last band and the last tile of a band may be smaller); tiles
// in row-major order, pixels row-major inside a tile.
__device__ __forceinline__ uint32_t tilePixel(uint32_t i, uint32_t w, uint32_t h, uint32_t tileW, uint32_t tileH)
{
    if (tileW == 0)
        return i;
    uint32_t band = i / (tileH * w);
    uint32_t bh = min(tileH, h - band * tileH);
    uint32_t j = i - band * tileH * w;
    uint32_t t = j / (tileW * bh);
    uint32_t k = j - t * tileW * bh;
    uint32_t tcw = min(tileW, w - t * tileW);
    uint32_t r = k / tcw, c = k - r * tcw;
    return (band * tileH + r) * w + t * tileW + c;
}


__device__ __forceinline__ void put(uint32_t* s, uint32_t n, uint32_t f, uint32_t i, float v) { s[(size_t)f * n + i] = asu(v); }
__device__ __forceinline__ float getf(const uint32_t* s, uint32_t n, uint32_t f, uint32_t i) { return asf(s[(size_t)f * n + i]); }
__device__ __forceinline__ void put3(uint32_t* s, uint32_t n, uint32_t f, uint32_t i, V3 v)
{
    put(s, n, f, i, v.x);
    put(s, n, f + 1, i, v.y);
    put(s, n, f + 2, i, v.z);
}
__device__ __forceinline__ V3 get3(const uint32_t* s, uint32_t n, uint32_t f, uint32_t i)
{
    return v3(getf(s, n, f, i), getf(s, n, f + 1, i), getf(s, n, f + 2, i));
}

// storePath / loadPath (wf_state.slang:35-92), slot-indexed
template <bool SPLIT>
__device__ __forceinline__ void storePath(uint32_t* s, uint32_t n, uint32_t i, const PathState& p, uint32_t pixel)
{
    put3(s, n, 0, i, p.org);
    put3(s, n, 3, i, p.dir);
    put3(s, n, 6, i, p.thr);
    s[(size_t)9 * n + i] = p.rng;
    s[(size_t)10 * n + i] = p.cls | ((uint32_t)(p.chan + 1) << 2) | (p.misPrev ? 16u : 0u) | (p.restir ? 32u : 0u) |
                            (p.gi ? 64u : 0u) | (p.prevInst << 8);
    put(s, n, 11, i, p.pathLen);
    put(s, n, 12, i, p.qPrev);
    put3(s, n, 13, i, p.parts.total);
    s[(size_t)16 * n + i] = pixel;
    if (SPLIT) {
        put3(s, n, kBaseWords, i, p.parts.directDiffuse);
        put3(s, n, kBaseWords + 3, i, p.parts.indirectDiffuse);
        put3(s, n, kBaseWords + 6, i, p.parts.specular);
        put3(s, n, kBaseWords + 9, i, p.parts.transmission);
        put3(s, n, kBaseWords + 12, i, p.parts.emission);
        put(s, n, kBaseWords + 15, i, p.visLit);
        put(s, n, kBaseWords + 16, i, p.visTotal);
    }
}

template <bool SPLIT>
__device__ __forceinline__ void loadPath(const uint32_t* s, uint32_t n, uint32_t i, PathState& p, uint32_t& pixel)
{
    p.org = get3(s, n, 0, i);
    p.dir = get3(s, n, 3, i);
    p.thr = get3(s, n, 6, i);
    p.rng = s[(size_t)9 * n + i];
    uint32_t flags = s[(size_t)10 * n + i];
    p.cls = flags & 3u;
    p.chan = (int)((flags >> 2) & 3u) - 1;
    p.misPrev = (flags & 16u) != 0;
    p.restir = (flags & 32u) != 0;
    p.gi = (flags & 64u) != 0;
    p.prevInst = flags >> 8;
    p.pathLen = getf(s, n, 11, i);
    p.qPrev = getf(s, n, 12, i);
    p.parts = Parts{get3(s, n, 13, i), v3(0.0f), v3(0.0f), v3(0.0f), v3(0.0f), v3(0.0f)};
    pixel = s[(size_t)16 * n + i];
    p.visLit = 0.0f;
    p.visTotal = 0.0f;
    if (SPLIT) {
        p.parts.directDiffuse = get3(s, n, kBaseWords, i);
        p.parts.indirectDiffuse = get3(s, n, kBaseWords + 3, i);
        p.parts.specular = get3(s, n, kBaseWords + 6, i);
        p.parts.transmission = get3(s, n, kBaseWords + 9, i);
        p.parts.emission = get3(s, n, kBaseWords + 12, i);
        p.visLit = getf(s, n, kBaseWords + 15, i);
        p.visTotal = getf(s, n, kBaseWords + 16, i);
    }
}

} // namespace vh
