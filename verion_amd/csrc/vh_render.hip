// vh_render.hip - the path-tracing kernels: megakernel, wavefront stages, debug ray queries.
//
// Megakernel: shaders/pathtrace.slang:9-25, one thread per pixel, params[35] samples in a row.
// Wavefront: shaders/wf_select.slang, wf_args.slang, wf_extend.slang, wf_shade.slang with
// include/wf_state.slang, re-shaped for CDNA:
//   - path state is indexed by queue slot, not by pixel: every stage reads and writes whole
//     contiguous runs of each state word (the reference indexes state by pixel and scatters);
//     two copies ping-pong between depths and the pixel id travels as state word 16;
//   - compaction of the survivors is one ballot per wave, an LDS prefix over the block's waves and
//     one global atomic per block (the reference appends with one atomic per path, wf_shade.slang:24);
//   - instead of an indirect dispatch sized by a one-thread kernel (wf_args.slang), the extend and
//     shade kernels are grid-stride loops over the live count of their depth, read on the device,
//     so the host never waits between bounces and no launch depends on a count;
//   - the extend and shadow kernels trace with the wave-cooperative traversal (traceWave in
//     vh_trace.h): a wave's leaf-triangle tests are shared by its 64 lanes through LDS
//     (1080p procedural scene: wf_extend 0.366 -> 0.281 ms, wf_shadow 0.315 -> 0.204 ms per
//     launch), and run as one-wave blocks (0.279 -> 0.252 ms and 0.192 -> 0.179 ms);
//   - the shading records its NEE shadow rays (and ReSTIR DI's winner) for the shadow kernel and a
//     resolve kernel adds the visible contributions in the order the fused shading adds them;
//     ReSTIR GI's reuse runs ahead of the shading in k_gi_prepare.
// Both integrators call the same device functions (vh_shade.h), so their images are identical.
#include "vh_internal.h"
#include "vh_wavefront.h"

namespace vh {

namespace {

constexpr uint32_t kWfBlock = 256;
// The traversal kernels (extend, shadow) synchronise within a wave only: one wave per block lets
// each wave retire on its own instead of holding its block's slot until the slowest of its waves
// ends (their grid is given in kWfBlock-sized blocks like the others' and scaled at launch).
#ifndef VH_TRACE_BLOCK
#define VH_TRACE_BLOCK 64
#endif
constexpr uint32_t kTraceBlock = VH_TRACE_BLOCK;
constexpr uint32_t kTraceScale = kWfBlock / kTraceBlock;


// Minimum waves per SIMD the compiler must fit the extend and shade kernels into (their register
// budget). Measured on the 1080p procedural scene (8192 blocks, 8x8 tiles): free / free 167,
// 5 / 3 177, 6 / free 154 (spills) Msamples/s.
#ifndef VH_EXTEND_WAVES
#define VH_EXTEND_WAVES 5
#endif
#ifndef VH_SHADE_WAVES
#define VH_SHADE_WAVES 3
#endif


template <bool RESTIR, bool SPLIT>
__global__ __launch_bounds__(64) void k_pathtrace(Scene S, Frame F, Params P, uint32_t samples)
{
    uint32_t x = blockIdx.x * 8 + (threadIdx.x & 7u);
    uint32_t y = blockIdx.y * 8 + (threadIdx.x >> 3);
    if (x >= P.w[0] || y >= P.w[1])
        return;
    Ctx c{S, F, P};
    uint32_t pixel = y * P.w[0] + x;
    for (uint32_t k = 0; k < samples; k++) {
        if (P.w[32] != 0 && !pixelActive(c, pixel))
            return;
        PathState st;
        generatePath(P, x, y, P.w[6] + k, st);
        for (uint32_t depth = 0; depth < P.w[7]; depth++) {
            Hit h = trace<false>(S.bvh, st.org, st.dir, 1.0e30f);
            if (!shadePath<RESTIR, SPLIT>(c, st, h, depth, pixel))
                break;
        }
        finishPath<SPLIT, RESTIR>(c, st, pixel);
    }
}

// The plain megakernel with a dispatch's samples in parallel and persistent waves. A wave takes
// (pixel, sample) items from `next` (one atomic per wave, items in sample order, 8 x 8 tiles
// within a sample), each lane carries its path one depth per step, and a lane whose path ended
// stages its radiance (zero when not finite, as finishPath does) and takes the next item at the
// following step, so the wave's lanes stay busy instead of waiting for its longest path. The
// loop is wave-uniform, so every ray is traced by the whole wave: the closest hits with the
// wave-cooperative traversal and, with DEFER, the NEE shadow rays the shading records (one
// record slot of the wave at a time, the visible contributions added in record order, as
// k_wf_shadow and k_wf_resolve do). k_film_add_samples then adds the staged samples to the FP64
// film in sample order, the additions finishPath makes one sample at a time, so the film is the
// same bit for bit as k_pathtrace's. Plain integrator only (no adaptive statistics, lighting
// split or ReSTIR, which need the samples in order). A wave ends when the items are drained and
// its lanes' paths are done (at most P.w[7] steps after its last fetch). Shadow records of lane l
// of the grid: shRec[(word K + record) lanes + l]. Cornell box 256 x 256 x 16 spp: 411
// Msamples/s with one thread per (pixel, sample), 677 with persistent waves, 857 with every ray
// traced wave-cooperatively, ~1000 at 4 waves per SIMD.
// Occupancy target 4 waves per SIMD (128 VGPRs, 80 B of spills), 16 resident waves per CU.
// Measured (Cornell 256 x 256 x 16 spp / 1080p procedural, 1 spp per call, Msamples/s): 2 per
// SIMD (177 VGPRs) 848 / 471, 3 (168) 1001 / 547, 4 (128) 997-1039 / 566-568, 5 (96) 840 / 540,
// 6 (80) 794 / 508.
template <bool DEFER>
__global__ __launch_bounds__(64, 4) void k_pathtrace_persist(Scene S, Params P, float* __restrict__ staged, uint32_t* __restrict__ next,
                                                          uint32_t items, float* __restrict__ shRec, uint32_t shK)
{
    Frame F{};
    Ctx c{S, F, P};
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t w = P.w[0], npix = P.w[0] * P.w[1];
    const uint32_t slot = blockIdx.x * 64 + lane, lanes = gridDim.x * 64;
    const size_t stride = (size_t)shK * lanes;
    __shared__ WaveTris sh;
    PathState st;
    uint32_t pixel = 0, k = 0, depth = 0;
    bool have = false, drained = false;
    for (;;) {
        const uint64_t need = __ballot(!have);
        if (need != 0 && !drained) { // wave-uniform
            uint32_t base = 0;
            if (lane == 0)
                base = atomicAdd(next, (uint32_t)__popcll(need));
            base = (uint32_t)__shfl((int)base, 0, 64);
            drained = base + (uint32_t)__popcll(need) >= items;
            if (!have) {
                uint32_t item = base + __builtin_amdgcn_mbcnt_hi((uint32_t)(need >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)need, 0u));
                if (item < items) {
                    k = item / npix;
                    pixel = tilePixel(item - k * npix, w, P.w[1], 8u, 8u);
                    generatePath(P, pixel % w, pixel / w, P.w[6] + k, st);
                    depth = 0;
                    have = true;
                }
            }
        }
        if (__ballot(have) == 0)
            break;
        Hit h = traceWave<false>(S.bvh, have ? st.org : v3(0.0f), have ? st.dir : v3(0.0f, 0.0f, 1.0f), 1.0e30f, have, sh);
        bool alive = false;
        uint32_t nrec = 0;
        if (have) {
            ShadowSink sink{shRec, lanes, shK, slot, 0u};
            alive = shadePath<false, false, 0, DEFER>(c, st, h, depth, pixel, nullptr, DEFER ? &sink : nullptr) && depth + 1 < P.w[7];
            nrec = sink.count;
            depth++;
        }
        if constexpr (DEFER) {
            for (uint32_t r = 0; r < shK; r++) {
                const bool act = r < nrec;
                if (__ballot(act) == 0) // wave-uniform
                    break;
                const float* rec = shRec + (size_t)r * lanes + slot;
                V3 o = v3(0.0f), dir = v3(0.0f, 0.0f, 1.0f);
                float tmax = 1.0e30f;
                if (act) {
                    o = v3(rec[0], rec[stride], rec[2 * stride]);
                    dir = v3(rec[3 * stride], rec[4 * stride], rec[5 * stride]);
                    tmax = rec[6 * stride];
                }
                bool blocked;
                if (S.translucent) {
                    blocked = act && occluded(c, o, dir, tmax);
                } else {
                    Hit hs = traceWave<true>(S.bvh, o, dir, tmax, act, sh);
                    blocked = act && hs.hit;
                }
                if (act && !blocked)
                    st.parts.total = st.parts.total + v3(rec[7 * stride], rec[8 * stride], rec[9 * stride]);
            }
        }
        if (have && !alive) {
            V3 radiance = bad(st.parts.total) ? v3(0.0f) : st.parts.total;
            float* o = staged + 3 * ((size_t)k * npix + pixel);
            o[0] = radiance.x;
            o[1] = radiance.y;
            o[2] = radiance.z;
            have = false;
        }
    }
}

__global__ void k_zero_u32(uint32_t* p)
{
    if (threadIdx.x == 0)
        *p = 0u;
}

__global__ __launch_bounds__(256) void k_film_add_samples(double* __restrict__ film, uint32_t npix, uint32_t samples,
                                                          const float* __restrict__ staged)
{
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += gridDim.x * blockDim.x) {
        double* f = film + 3 * (size_t)p;
        double r = f[0], g = f[1], b = f[2];
        for (uint32_t k = 0; k < samples; k++) {
            const float* v = staged + 3 * ((size_t)k * npix + p);
            r += (double)v[0];
            g += (double)v[1];
            b += (double)v[2];
        }
        f[0] = r;
        f[1] = g;
        f[2] = b;
    }
}

// film += pass, pass = 0, two doubles per thread and step (both buffers are 256-byte aligned)
__global__ __launch_bounds__(256) void k_film_add_pass(double* __restrict__ film, double* __restrict__ pass, uint64_t n)
{
    const uint64_t pairs = n / 2;
    double2* f2 = reinterpret_cast<double2*>(film);
    double2* p2 = reinterpret_cast<double2*>(pass);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < pairs; i += (uint64_t)gridDim.x * blockDim.x) {
        double2 a = f2[i], b = p2[i];
        a.x += b.x;
        a.y += b.y;
        f2[i] = a;
        p2[i] = make_double2(0.0, 0.0);
    }
    if ((n & 1u) != 0 && blockIdx.x == 0 && threadIdx.x == 0) {
        film[n - 1] += pass[n - 1];
        pass[n - 1] = 0.0;
    }
}

// wf_args at depth 0 (wf_args.slang:12-17, and its clear of counts[0] at :21-22): the depth-0
// count (every pixel, or none yet when wf_select fills the queue) and empty later queues
__global__ void k_wf_reset(uint32_t* counts, uint32_t n, uint32_t adaptive, uint32_t maxDepth)
{
    if (threadIdx.x == 0) {
        counts[0] = adaptive ? 0u : n;
        for (uint32_t k = 1; k <= maxDepth; k++)
            counts[k] = 0;
        for (uint32_t k = 0; k < maxDepth; k++)
            counts[kShadowCounts + k] = 0;

    }
}

// Block-wide compaction: returns this thread's slot in the queue `counter` counts, for `keep` threads.
__device__ __forceinline__ uint32_t compact(bool keep, uint32_t* counter)
{
    __shared__ uint32_t s_wave[kWfBlock / 64];
    __shared__ uint32_t s_base;
    uint64_t m = __ballot(keep);
    uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint64_t lt = lane == 0 ? 0ull : ((~0ull) >> (64u - lane));
    uint32_t rank = (uint32_t)__popcll(m & lt);
    if (lane == 0)
        s_wave[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (uint32_t w = 0; w < kWfBlock / 64; w++) {
            uint32_t c = s_wave[w];
            s_wave[w] = run;
            run += c;
        }
        s_base = run ? atomicAdd(counter, run) : 0u;
    }
    __syncthreads();
    uint32_t slot = s_base + s_wave[wave] + rank;
    __syncthreads(); // s_wave and s_base are reused by the next call
    return slot;
}

// Block-wide exclusive prefix of per-thread counts k, one atomic per block on `counter`; returns
// this thread's first index.
__device__ __forceinline__ uint32_t compactN(uint32_t k, uint32_t* counter)
{
    __shared__ uint32_t s_wave[kWfBlock / 64];
    __shared__ uint32_t s_base;
    uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t x = k;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off)
            x += y;
    }
    if (lane == 63)
        s_wave[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (uint32_t w = 0; w < kWfBlock / 64; w++) {
            uint32_t c = s_wave[w];
            s_wave[w] = run;
            run += c;
        }
        s_base = run ? atomicAdd(counter, run) : 0u;
    }
    __syncthreads();
    uint32_t r = s_base + s_wave[wave] + x - k;
    __syncthreads();
    return r;
}

// wf_select (wf_select.slang:7-17): pixels that still take samples -> depth-0 queue
__global__ __launch_bounds__(kWfBlock) void k_wf_select(Frame F, Params P, Wavefront W)
{
    Scene S{};
    Ctx c{S, F, P};
    uint32_t n = W.n;
    for (uint32_t base = blockIdx.x * kWfBlock; base < n; base += gridDim.x * kWfBlock) {
        uint32_t i = base + threadIdx.x;
        uint32_t p = i < n ? tilePixel(i, P.w[0], P.w[1], W.tileW, W.tileH) : 0u;
        bool active = i < n && pixelActive(c, p);
        uint32_t slot = compact(active, &W.counts[0]);
        if (active && slot < n) // slots stay below n: counts[0] starts the sample at 0 (k_wf_reset)
            W.queue0[slot] = p;
    }
}

// wf_extend (wf_extend.slang:6-29): closest hit of every live path at depth d
__global__ __launch_bounds__(kTraceBlock, VH_EXTEND_WAVES) void k_wf_extend(Scene S, Params P, Wavefront W, uint32_t d)
{
    const uint32_t count = W.counts[d];
    const uint32_t* st = W.state[d & 1u];
    auto load = [&](uint32_t i, V3& org, V3& dir, float& tmax) {
        if (d == 0) {
            uint32_t p = P.w[32] != 0 ? W.queue0[i] : tilePixel(i, P.w[0], P.w[1], W.tileW, W.tileH);
            PathState s;
            generatePath(P, p % P.w[0], p / P.w[0], P.w[6], s);
            org = s.org;
            dir = s.dir;
        } else {
            org = get3(st, W.n, 0, i);
            dir = get3(st, W.n, 3, i);
        }
        tmax = 1.0e30f;
    };
    auto done = [&](uint32_t i, const Hit& h) { W.hits[i] = make_uint4(h.hit ? 1u : 0u, h.inst, h.prim, asu(h.t)); };
    __shared__ WaveTris waveTris[kTraceBlock / 64];
    WaveTris& sh = waveTris[threadIdx.x >> 6];
    // wave-uniform grid stride: the 64 lanes of a wave trace together (traceWave), `per` of them
    // with rays of their own
    static_assert(kTraceBlock == 64, "one wave per block");
    const uint32_t per = raysPerWave(count, W.traceWaves);
    for (uint32_t base = blockIdx.x * per; base < count; base += gridDim.x * per) {
        uint32_t i = base + threadIdx.x;
        bool act = threadIdx.x < per && i < count;
        V3 org = v3(0.0f), dir = v3(0.0f, 0.0f, 1.0f);
        float tmax = 1.0e30f;
        if (act)
            load(i, org, dir, tmax);
        Hit h = traceWave<false>(S.bvh, org, dir, tmax, act, sh);
        if (act)
            done(i, h);
    }
}

// wf_shade (wf_shade.slang:7-28): shade, then append the survivor to the next depth or finish it
template <bool RESTIR, bool SPLIT>
__global__ __launch_bounds__(kWfBlock, VH_SHADE_WAVES) void k_wf_shade(Scene S, Frame F, Params P, Wavefront W, uint32_t d)
{
    Ctx c{S, F, P};
    uint32_t count = W.counts[d];
    const uint32_t* in = W.state[d & 1u];
    uint32_t* out = W.state[(d + 1) & 1u];
    for (uint32_t base = blockIdx.x * kWfBlock; base < count; base += gridDim.x * kWfBlock) {
        uint32_t i = base + threadIdx.x;
        bool keep = false;
        PathState st;
        uint32_t pixel = 0;
        if (i < count) {
            if (d == 0) {
                pixel = P.w[32] != 0 ? W.queue0[i] : tilePixel(i, P.w[0], P.w[1], W.tileW, W.tileH);
                generatePath(P, pixel % P.w[0], pixel / P.w[0], P.w[6], st);
            } else {
                loadPath<SPLIT>(in, W.n, i, st, pixel);
            }
            uint4 hw = W.hits[i];
            Hit h;
            h.hit = hw.x != 0;
            h.inst = hw.y;
            h.prim = hw.z;
            h.t = asf(hw.w);
            h.b1 = h.b2 = 0.0f;
            h.gid = 0;
            keep = shadePath<RESTIR, SPLIT>(c, st, h, d, pixel) && d + 1 < P.w[7];
            if (!keep)
                finishPath<SPLIT, RESTIR>(c, st, pixel);
        }
        uint32_t slot = compact(keep, &W.counts[d + 1]);
        if (keep && slot < W.n) // survivors of depth d never outnumber its count (<= n)
            storePath<SPLIT>(out, W.n, slot, st, pixel);
    }
}


// wf_shade of the plain integrator with deferred shadow rays: shade, record the NEE shadow rays
// (with the contribution each adds when unoccluded), append survivors to the next depth. A path
// with records stays pending (its total without them, its next slot or its end) until
// k_wf_resolve; a path without records is stored or finished here. RESTIR: ReSTIR's first hit,
// where DI's winner's shadow ray is one more record (k_wf_resolve zeroes the stored reservoir
// weight when it is blocked). GI: its reuse at the first hit was made by k_gi_prepare; the first
// hit's own light (the path total after the first hit's NEE, which GI keeps apart) is written by
// k_wf_resolve once its shadow rays are known; a path that carries a second-hit sample finishes
// through giFinish, here or in k_wf_resolve.
template <bool GI, bool RESTIR>
__global__ __launch_bounds__(kWfBlock, VH_SHADE_WAVES) void k_wf_shade_deferred(Scene S, Frame F, Params P, Wavefront W, uint32_t d)
{
    Ctx c{S, F, P};
    uint32_t count = W.counts[d];
    const uint32_t* in = W.state[d & 1u];
    uint32_t* out = W.state[(d + 1) & 1u];
    for (uint32_t base = blockIdx.x * kWfBlock; base < count; base += gridDim.x * kWfBlock) {
        uint32_t i = base + threadIdx.x;
        bool keep = false;
        PathState st;
        uint32_t pixel = 0, nrec = 0, tag = 0;
        if (i < count) {
            if (d == 0) {
                pixel = P.w[32] != 0 ? W.queue0[i] : tilePixel(i, P.w[0], P.w[1], W.tileW, W.tileH);
                generatePath(P, pixel % P.w[0], pixel / P.w[0], P.w[6], st);
            } else {
                loadPath<false>(in, W.n, i, st, pixel);
            }
            uint4 hw = W.hits[i];
            Hit h;
            h.hit = hw.x != 0;
            h.inst = hw.y;
            h.prim = hw.z;
            h.t = asf(hw.w);
            h.b1 = h.b2 = 0.0f;
            h.gid = 0;
            ShadowSink sink{W.shRec, W.n, W.shK, i, 0u};
            keep = shadePath<RESTIR, false, 0, true, GI ? 2 : 0>(c, st, h, d, pixel, nullptr, &sink) && d + 1 < P.w[7];
            nrec = sink.count;
            W.shNRec[i] = nrec | (sink.restirRec << 8);
            tag = pixel;
            if (!keep && nrec == 0) {
                finishPath<false, GI>(c, st, pixel);
            } else if (GI && !keep && st.gi) {
                // finished in k_wf_resolve: finishPath's dispersion check now, its pixel bit there
                if (st.chan >= 0)
                    c.F.restir[restirScratch(pixel) + 22] = asf(0u);
                tag |= kPendingGI;
            }
            if (GI && d == 0 && st.gi)
                tag |= kPendingFirst;
        }
        uint32_t slot = compact(keep, &W.counts[d + 1]);
        if (keep && slot < W.n)
            storePath<false>(out, W.n, slot, st, pixel);
        uint32_t q = compactN(nrec, &W.counts[kShadowCounts + d]);
        if (nrec > 0) {
            W.shOut[i] = keep ? slot : 0xffffffffu;
            W.shPixel[i] = tag;
            put3(reinterpret_cast<uint32_t*>(W.shTotal), W.n, 0, i, st.parts.total);
            for (uint32_t k = 0; k < nrec; k++)
                if (q + k < W.shK * W.n)
                    W.shQueue[q + k] = k * W.n + i;
        }
    }
}

// ReSTIR GI's reuse at the first hit (restirGIPrepare, pt_core.slang:901-972) for the depth-0 queue,
// ahead of k_wf_shade_deferred: each lane replays its first hit's shading up to the reuse
// (shadePath's probe), then the wave resamples its lanes' neighbours together, tracing their
// visibility rays with the wave-cooperative traversal (restirGIPrepareWave). The grid stride is
// wave-uniform: every lane of a wave reaches the cooperative traversal.
__global__ __launch_bounds__(kTraceBlock, VH_SHADE_WAVES) void k_gi_prepare(Scene S, Frame F, Params P, Wavefront W)
{
    Ctx c{S, F, P};
    const uint32_t count = W.counts[0];
    __shared__ WaveTris waveTris[kTraceBlock / 64];
    WaveTris& sh = waveTris[threadIdx.x >> 6];
    for (uint32_t base = blockIdx.x * kTraceBlock; base < count; base += gridDim.x * kTraceBlock) {
        uint32_t i = base + threadIdx.x;
        GiSite site;
        site.active = false;
        uint32_t pixel = 0;
        if (i < count) {
            pixel = P.w[32] != 0 ? W.queue0[i] : tilePixel(i, P.w[0], P.w[1], W.tileW, W.tileH);
            PathState st;
            generatePath(P, pixel % P.w[0], pixel / P.w[0], P.w[6], st);
            uint4 hw = W.hits[i];
            Hit h;
            h.hit = hw.x != 0;
            h.inst = hw.y;
            h.prim = hw.z;
            h.t = asf(hw.w);
            h.b1 = h.b2 = 0.0f;
            h.gid = 0;
            (void)shadePath<true, false, 0, false, 1>(c, st, h, 0, pixel, nullptr, nullptr, &site);
        }
        restirGIPrepareWave(c, site, pixel, sh);
    }
}

// The shadow rays recorded at depth d (occluded, pt_core.slang:580-613), one per thread.
__global__ __launch_bounds__(kTraceBlock, VH_EXTEND_WAVES) void k_wf_shadow(Scene S, Params P, Wavefront W, uint32_t d)
{
    Frame F{};
    Ctx c{S, F, P};
    const uint32_t count = W.counts[kShadowCounts + d];
    const size_t stride = (size_t)W.shK * W.n;
    auto load = [&](uint32_t q, V3& o, V3& dir, float& tmax) {
        const float* r = W.shRec + W.shQueue[q];
        o = v3(r[0], r[stride], r[2 * stride]);
        dir = v3(r[3 * stride], r[4 * stride], r[5 * stride]);
        tmax = r[6 * stride];
    };
    if (!S.translucent) { // any hit ends an opaque shadow ray: the wave-cooperative traversal
        __shared__ WaveTris waveTris[kTraceBlock / 64];
        WaveTris& sh = waveTris[threadIdx.x >> 6];
        const uint32_t per = raysPerWave(count, W.traceWaves);
        for (uint32_t base = blockIdx.x * per; base < count; base += gridDim.x * per) {
            uint32_t q = base + threadIdx.x;
            bool act = threadIdx.x < per && q < count;
            V3 o = v3(0.0f), dir = v3(0.0f, 0.0f, 1.0f);
            float tmax = 1.0e30f;
            if (act)
                load(q, o, dir, tmax);
            Hit h = traceWave<true>(S.bvh, o, dir, tmax, act, sh);
            if (act)
                W.shVis[W.shQueue[q]] = h.hit ? 1u : 0u;
        }
        return;
    }
    // translucent scenes: closest hits per layer (occluded)
    for (uint32_t q = blockIdx.x * kTraceBlock + threadIdx.x; q < count; q += gridDim.x * kTraceBlock) {
        V3 o, dir;
        float tmax;
        load(q, o, dir, tmax);
        W.shVis[W.shQueue[q]] = occluded(c, o, dir, tmax) ? 1u : 0u;
    }
}

// k_wf_extend for depth d and the opaque scene's k_wf_shadow for depth d - 1 in one launch: waves
// [0, wc) take closest rays, the rest shadow rays (kind wave-uniform). Both queues are final once
// depth d - 1 is shaded, and neither reads what the other writes; one launch's tail instead of
// two, which at the deep depths is most of their time.
__global__ __launch_bounds__(kTraceBlock, VH_EXTEND_WAVES) void k_wf_trace(Scene S, Params P, Wavefront W, uint32_t d)
{
    static_assert(kTraceBlock == 64, "one wave per block");
    const uint32_t nc = W.counts[d];
    const uint32_t ns = d > 0 ? W.counts[kShadowCounts + d - 1] : 0u;
    const uint32_t per = raysPerWave(nc + ns, W.traceWaves);
    const uint32_t wc = (nc + per - 1) / per, waves = wc + (ns + per - 1) / per;
    const uint32_t* st = W.state[d & 1u];
    const size_t stride = (size_t)W.shK * W.n;
    __shared__ WaveTris sh;
    const uint32_t lane = threadIdx.x;
    for (uint32_t w = blockIdx.x; w < waves; w += gridDim.x) {
        if (w < wc) {
            const uint32_t i = w * per + lane;
            const bool act = lane < per && i < nc;
            V3 org = v3(0.0f), dir = v3(0.0f, 0.0f, 1.0f);
            if (act && d == 0) {
                uint32_t p = P.w[32] != 0 ? W.queue0[i] : tilePixel(i, P.w[0], P.w[1], W.tileW, W.tileH);
                PathState s;
                generatePath(P, p % P.w[0], p / P.w[0], P.w[6], s);
                org = s.org;
                dir = s.dir;
            } else if (act) {
                org = get3(st, W.n, 0, i);
                dir = get3(st, W.n, 3, i);
            }
            Hit h = traceWave<false>(S.bvh, org, dir, 1.0e30f, act, sh);
            if (act)
                W.hits[i] = make_uint4(h.hit ? 1u : 0u, h.inst, h.prim, asu(h.t));
        } else {
            const uint32_t q = (w - wc) * per + lane;
            const bool act = lane < per && q < ns;
            V3 o = v3(0.0f), dir = v3(0.0f, 0.0f, 1.0f);
            float tmax = 1.0e30f;
            uint32_t rid = 0;
            if (act) {
                rid = W.shQueue[q];
                const float* r = W.shRec + rid;
                o = v3(r[0], r[stride], r[2 * stride]);
                dir = v3(r[3 * stride], r[4 * stride], r[5 * stride]);
                tmax = r[6 * stride];
            }
            Hit h = traceWave<true>(S.bvh, o, dir, tmax, act, sh);
            if (act)
                W.shVis[rid] = h.hit ? 1u : 0u;
        }
    }
}

// Adds the visible shadow contributions of every pending path at depth d in record order, then
// stores the total in the path's next slot or finishes the path.
template <bool GI>
__global__ __launch_bounds__(kWfBlock) void k_wf_resolve(Scene S, Frame F, Params P, Wavefront W, uint32_t d)
{
    Ctx c{S, F, P};
    const uint32_t count = W.counts[d];
    uint32_t* out = W.state[(d + 1) & 1u];
    const size_t stride = (size_t)W.shK * W.n;
    for (uint32_t i = blockIdx.x * kWfBlock + threadIdx.x; i < count; i += gridDim.x * kWfBlock) {
        uint32_t nw = W.shNRec[i];
        uint32_t nrec = nw & 0xffu, restirRec = nw >> 8;
        if (nrec == 0)
            continue;
        V3 total = get3(reinterpret_cast<const uint32_t*>(W.shTotal), W.n, 0, i);
        for (uint32_t k = 0; k < nrec; k++) {
            uint32_t rid = k * W.n + i;
            if (W.shVis[rid] == 0) {
                const float* r = W.shRec + rid;
                total = total + v3(r[7 * stride], r[8 * stride], r[9 * stride]);
            } else if (k + 1 == restirRec) {
                // the ReSTIR DI winner is blocked: its reservoir is stored with W = 0 (pt_core.slang:872-879)
                F.restir[restirBase(W.shPixel[i] & ~(kPendingGI | kPendingFirst), P.w[6] & 1u) + 4] = 0.0f;
            }
        }
        if (GI && (W.shPixel[i] & kPendingFirst) != 0) {
            // ReSTIR GI's first-hit light: the total after the first hit (pt_core.slang:1531-1540)
            float* Sc = F.restir + restirScratch(W.shPixel[i] & ~(kPendingGI | kPendingFirst));
            Sc[18] = total.x;
            Sc[19] = total.y;
            Sc[20] = total.z;
        }
        uint32_t os = W.shOut[i];
        if (os != 0xffffffffu) {
            put3(out, W.n, 13, os, total);
        } else {
            PathState st;
            st.parts = Parts{total, v3(0.0f), v3(0.0f), v3(0.0f), v3(0.0f), v3(0.0f)};
            st.visLit = 0.0f;
            st.visTotal = 0.0f;
            uint32_t pixel = W.shPixel[i];
            st.gi = GI && (pixel & kPendingGI) != 0;
            st.chan = -1; // a dispersive path's GI sample was dropped when it was recorded
            finishPath<false, GI>(c, st, pixel & ~(kPendingGI | kPendingFirst));
        }
    }
}

// ---- neural radiance cache: training paths (plain paths of a pixel subset, every vertex recorded)

__global__ void k_nrc_pick(uint32_t* trainPix, uint32_t nTrain, uint32_t npix, uint32_t frame, uint32_t* counts, uint32_t* sampleCount)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nTrain)
        trainPix[i] = pcg(i ^ pcg(frame * 0x9E3779B9u + 0x2545F491u)) % npix;
    if (i == 0) {
        counts[0] = nTrain;
        for (uint32_t k = 1; k <= kMaxDepthNrc; k++)
            counts[k] = 0;
        for (uint32_t k = 0; k < kMaxDepthNrc; k++)
            counts[kShadowCounts + k] = 0;
        *sampleCount = 0;
    }
}

__global__ __launch_bounds__(kTraceBlock, VH_EXTEND_WAVES) void k_nrc_t_extend(Scene S, Params P, Wavefront W, const uint32_t* trainPix, uint32_t d)
{
    uint32_t count = W.counts[d];
    const uint32_t* st = W.state[d & 1u];
    __shared__ WaveTris waveTris[kTraceBlock / 64];
    WaveTris& sh = waveTris[threadIdx.x >> 6];
    const uint32_t per = raysPerWave(count, W.traceWaves);
    for (uint32_t base = blockIdx.x * per; base < count; base += gridDim.x * per) {
        uint32_t i = base + threadIdx.x;
        bool act = threadIdx.x < per && i < count;
        V3 org = v3(0.0f), dir = v3(0.0f, 0.0f, 1.0f);
        if (act && d == 0) {
            uint32_t p = trainPix[i];
            PathState s;
            generatePath(P, p % P.w[0], p / P.w[0], P.w[6], s);
            org = s.org;
            dir = s.dir;
        } else if (act) {
            org = get3(st, W.n, 0, i);
            dir = get3(st, W.n, 3, i);
        }
        Hit h = traceWave<false>(S.bvh, org, dir, 1.0e30f, act, sh);
        if (act)
            W.hits[i] = make_uint4(h.hit ? 1u : 0u, h.inst, h.prim, asu(h.t));
    }
}

// Shades training paths; a finished path turns each recorded vertex into a training sample
// (inputs, target = (final total - total at the vertex) / throughput at the vertex).
__global__ __launch_bounds__(kWfBlock) void k_nrc_t_shade(Scene S, Params P, Wavefront W, const uint32_t* trainPix, float4* vertices,
                                                          float4* samples, uint32_t* sampleCount, uint32_t sampleCap, uint32_t d)
{
    Frame F{};
    Ctx c{S, F, P};
    uint32_t count = W.counts[d];
    const uint32_t* in = W.state[d & 1u];
    uint32_t* out = W.state[(d + 1) & 1u];
    for (uint32_t base = blockIdx.x * kWfBlock; base < count; base += gridDim.x * kWfBlock) {
        uint32_t i = base + threadIdx.x;
        bool keep = false, finished = false;
        PathState st;
        uint32_t tp = 0;
        if (i < count) {
            if (d == 0) {
                tp = i;
                uint32_t p = trainPix[i];
                generatePath(P, p % P.w[0], p / P.w[0], P.w[6], st);
                for (uint32_t j = 0; j < kMaxDepthNrc; j++)
                    vertices[((size_t)tp * kMaxDepthNrc + j) * kNrcVertexWords + 4].w = 0.0f;
            } else {
                loadPath<false>(in, W.n, i, st, tp);
            }
            uint4 hw = W.hits[i];
            Hit h;
            h.hit = hw.x != 0;
            h.inst = hw.y;
            h.prim = hw.z;
            h.t = asf(hw.w);
            h.b1 = h.b2 = 0.0f;
            h.gid = 0;
            NrcIo io{vertices, nullptr, tp, i, 0u, false};
            keep = shadePath<false, false, 1>(c, st, h, d, tp, &io) && d + 1 < P.w[7];
            finished = !keep;
        }
        uint32_t nvalid = 0;
        const float4* vrec = vertices + (size_t)tp * kMaxDepthNrc * kNrcVertexWords;
        if (finished && !bad(st.parts.total)) {
            for (uint32_t j = 0; j < kMaxDepthNrc; j++)
                nvalid += vrec[j * kNrcVertexWords + 4].w != 0.0f ? 1u : 0u;
        }
        uint32_t first = compactN(nvalid, sampleCount);
        if (nvalid) {
            V3 T = st.parts.total;
            uint32_t o = first;
            for (uint32_t j = 0; j < kMaxDepthNrc && o < sampleCap; j++) {
                const float4* v = vrec + j * kNrcVertexWords;
                if (v[4].w == 0.0f)
                    continue;
                float4 thr = v[3], tae = v[4];
                float4 tgt = make_float4(thr.x > 1.0e-6f ? (T.x - tae.x) / thr.x : 0.0f, thr.y > 1.0e-6f ? (T.y - tae.y) / thr.y : 0.0f,
                                         thr.z > 1.0e-6f ? (T.z - tae.z) / thr.z : 0.0f, 1.0f);
                float4* s = samples + (size_t)o * 4;
                s[0] = v[0];
                s[1] = v[1];
                s[2] = v[2];
                s[3] = tgt;
                o++;
            }
        }
        uint32_t slot = compact(keep, &W.counts[d + 1]);
        if (keep && slot < W.n)
            storePath<false>(out, W.n, slot, st, tp);
    }
}

// The training samples of finished training path tp with final total T: one per recorded
// vertex, target (T - total at the vertex) / throughput at the vertex. Every thread of the block
// calls it (block-wide compaction into the sample buffer).
__device__ __forceinline__ void emitSamples(bool finished, uint32_t tp, V3 T, const float4* vertices, float4* samples,
                                            uint32_t* sampleCount, uint32_t sampleCap)
{
    uint32_t nvalid = 0;
    const float4* vrec = vertices + (size_t)tp * kMaxDepthNrc * kNrcVertexWords;
    if (finished && !bad(T)) {
        for (uint32_t j = 0; j < kMaxDepthNrc; j++)
            nvalid += vrec[j * kNrcVertexWords + 4].w != 0.0f ? 1u : 0u;
    }
    uint32_t first = compactN(nvalid, sampleCount);
    if (nvalid) {
        uint32_t o = first;
        for (uint32_t j = 0; j < kMaxDepthNrc && o < sampleCap; j++) {
            const float4* v = vrec + j * kNrcVertexWords;
            if (v[4].w == 0.0f)
                continue;
            float4 thr = v[3], tae = v[4];
            float4 tgt = make_float4(thr.x > 1.0e-6f ? (T.x - tae.x) / thr.x : 0.0f, thr.y > 1.0e-6f ? (T.y - tae.y) / thr.y : 0.0f,
                                     thr.z > 1.0e-6f ? (T.z - tae.z) / thr.z : 0.0f, 1.0f);
            float4* sp = samples + (size_t)o * 4;
            sp[0] = v[0];
            sp[1] = v[1];
            sp[2] = v[2];
            sp[3] = tgt;
            o++;
        }
    }
}

// Training paths with deferred shadow rays (as k_wf_shade_deferred for the render): shade and
// record the vertices; a path that ends without pending shadow rays gives its samples now, one
// with pending rays gives them in k_nrc_t_resolve once its total is final.
__global__ __launch_bounds__(kWfBlock, VH_SHADE_WAVES) void k_nrc_t_shade_deferred(Scene S, Params P, Wavefront W, const uint32_t* trainPix,
                                                                                 float4* vertices, float4* samples, uint32_t* sampleCount,
                                                                                 uint32_t sampleCap, uint32_t d)
{
    Frame F{};
    Ctx c{S, F, P};
    uint32_t count = W.counts[d];
    const uint32_t* in = W.state[d & 1u];
    uint32_t* out = W.state[(d + 1) & 1u];
    for (uint32_t base = blockIdx.x * kWfBlock; base < count; base += gridDim.x * kWfBlock) {
        uint32_t i = base + threadIdx.x;
        bool keep = false;
        PathState st;
        uint32_t tp = 0, nrec = 0;
        if (i < count) {
            if (d == 0) {
                tp = i;
                uint32_t p = trainPix[i];
                generatePath(P, p % P.w[0], p / P.w[0], P.w[6], st);
                for (uint32_t j = 0; j < kMaxDepthNrc; j++)
                    vertices[((size_t)tp * kMaxDepthNrc + j) * kNrcVertexWords + 4].w = 0.0f;
            } else {
                loadPath<false>(in, W.n, i, st, tp);
            }
            uint4 hw = W.hits[i];
            Hit h;
            h.hit = hw.x != 0;
            h.inst = hw.y;
            h.prim = hw.z;
            h.t = asf(hw.w);
            h.b1 = h.b2 = 0.0f;
            h.gid = 0;
            NrcIo io{vertices, nullptr, tp, i, 0u, false};
            ShadowSink sink{W.shRec, W.n, W.shK, i, 0u};
            keep = shadePath<false, false, 1, true>(c, st, h, d, tp, &io, &sink) && d + 1 < P.w[7];
            nrec = sink.count;
            W.shNRec[i] = nrec;
        }
        emitSamples(i < count && !keep && nrec == 0, tp, i < count ? st.parts.total : v3(0.0f), vertices, samples, sampleCount,
                    sampleCap);
        uint32_t slot = compact(keep, &W.counts[d + 1]);
        if (keep && slot < W.n)
            storePath<false>(out, W.n, slot, st, tp);
        uint32_t q = compactN(nrec, &W.counts[kShadowCounts + d]);
        if (nrec > 0) {
            W.shOut[i] = keep ? slot : 0xffffffffu;
            W.shPixel[i] = tp;
            put3(reinterpret_cast<uint32_t*>(W.shTotal), W.n, 0, i, st.parts.total);
            for (uint32_t k = 0; k < nrec; k++)
                if (q + k < W.shK * W.n)
                    W.shQueue[q + k] = k * W.n + i;
        }
    }
}

// k_wf_resolve for training paths: the visible shadow contributions in record order, then the
// total goes to the path's next slot or, for a finished path, into its training samples.
__global__ __launch_bounds__(kWfBlock) void k_nrc_t_resolve(Wavefront W, float4* vertices, float4* samples, uint32_t* sampleCount,
                                                           uint32_t sampleCap, uint32_t d)
{
    const uint32_t count = W.counts[d];
    uint32_t* out = W.state[(d + 1) & 1u];
    const size_t stride = (size_t)W.shK * W.n;
    for (uint32_t base = blockIdx.x * kWfBlock; base < count; base += gridDim.x * kWfBlock) {
        uint32_t i = base + threadIdx.x;
        uint32_t nrec = i < count ? W.shNRec[i] : 0u;
        V3 total = v3(0.0f);
        bool finished = false;
        uint32_t tp = 0;
        if (nrec > 0) {
            total = get3(reinterpret_cast<const uint32_t*>(W.shTotal), W.n, 0, i);
            for (uint32_t k = 0; k < nrec; k++) {
                uint32_t rid = k * W.n + i;
                if (W.shVis[rid] == 0) {
                    const float* r = W.shRec + rid;
                    total = total + v3(r[7 * stride], r[8 * stride], r[9 * stride]);
                }
            }
            uint32_t os = W.shOut[i];
            tp = W.shPixel[i];
            if (os != 0xffffffffu)
                put3(out, W.n, 13, os, total);
            else
                finished = true;
        }
        emitSamples(finished, tp, total, vertices, samples, sampleCount, sampleCap);
    }
}

// ---- neural radiance cache: render paths, shaded at the terminating depth

__global__ __launch_bounds__(kWfBlock) void k_nrc_q_shade(Scene S, Frame F, Params P, Wavefront W, float4* queries, uint32_t d)
{
    Ctx c{S, F, P};
    uint32_t count = W.counts[d];
    const uint32_t* in = W.state[d & 1u];
    uint32_t* out = W.state[(d + 1) & 1u];
    for (uint32_t base = blockIdx.x * kWfBlock; base < count; base += gridDim.x * kWfBlock) {
        uint32_t i = base + threadIdx.x;
        bool keep = false;
        PathState st;
        uint32_t pixel = 0;
        if (i < count) {
            loadPath<false>(in, W.n, i, st, pixel);
            uint4 hw = W.hits[i];
            Hit h;
            h.hit = hw.x != 0;
            h.inst = hw.y;
            h.prim = hw.z;
            h.t = asf(hw.w);
            h.b1 = h.b2 = 0.0f;
            h.gid = 0;
            NrcIo io{nullptr, queries, 0u, i, d, false};
            bool alive = shadePath<false, false, 2>(c, st, h, d, pixel, &io);
            keep = alive && d + 1 < P.w[7];
            if (!io.terminated) {
                queries[(size_t)i * kNrcQueryWords + 4] = make_float4(0.0f, 0.0f, 0.0f, 0.0f); // no query in this slot
                if (!keep)
                    finishPath<false>(c, st, pixel);
            }
        }
        uint32_t slot = compact(keep, &W.counts[d + 1]);
        if (keep && slot < W.n)
            storePath<false>(out, W.n, slot, st, pixel);
    }
}

__global__ __launch_bounds__(256) void k_primary(Scene S, Params P, uint4* out)
{
    uint32_t n = P.w[0] * P.w[1];
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        PathState s;
        generatePath(P, p % P.w[0], p / P.w[0], P.w[6], s);
        Hit h = trace<false>(S.bvh, s.org, s.dir, 1.0e30f);
        out[p] = h.hit ? make_uint4(1u, h.inst, h.prim, asu(h.t)) : make_uint4(0u, 0u, 0u, 0u);
    }
}

__global__ __launch_bounds__(256) void k_rays(Scene S, const float* rays, uint64_t n, float tmax, uint4* out)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const float* r = rays + 6 * i;
        Hit h = trace<false>(S.bvh, v3(r[0], r[1], r[2]), v3(r[3], r[4], r[5]), tmax);
        out[i] = h.hit ? make_uint4(1u, h.inst, h.prim, asu(h.t)) : make_uint4(0u, 0u, 0u, 0u);
    }
}

} // namespace

int launch_megakernel(const Scene& S, const Frame& F, const Params& P, uint32_t samples, bool restir, bool split, hipStream_t s)
{
    dim3 grid((P.w[0] + 7) / 8, (P.w[1] + 7) / 8);
    if (restir && split)
        k_pathtrace<true, true><<<grid, 64, 0, s>>>(S, F, P, samples);
    else if (restir)
        k_pathtrace<true, false><<<grid, 64, 0, s>>>(S, F, P, samples);
    else if (split)
        k_pathtrace<false, true><<<grid, 64, 0, s>>>(S, F, P, samples);
    else
        k_pathtrace<false, false><<<grid, 64, 0, s>>>(S, F, P, samples);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? VH_OK : fail(VH_ERR_KERNEL, std::string("pathtrace launch: ") + hipGetErrorString(e));
}

int launch_megakernel_samples(const Scene& S, const Frame& F, const Params& P, uint32_t samples, float* staged, uint32_t* next,
                              float* shRec, uint32_t shK, uint32_t cus, hipStream_t s)
{
    // resident waves: kPersistWavesPerCu per CU, the kernel's occupancy
    const uint32_t npix = P.w[0] * P.w[1];
    const uint32_t items = npix * samples;
    const uint32_t waves = std::min<uint32_t>((items + 63) / 64, cus * kPersistWavesPerCu);
    k_zero_u32<<<1, 64, 0, s>>>(next);
    if (shRec && P.w[5] + 1 <= shK)
        k_pathtrace_persist<true><<<waves, 64, 0, s>>>(S, P, staged, next, items, shRec, shK);
    else
        k_pathtrace_persist<false><<<waves, 64, 0, s>>>(S, P, staged, next, items, nullptr, 0u);
    k_film_add_samples<<<blocks_for(npix, 256, 4096), 256, 0, s>>>(F.film, npix, samples, staged);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? VH_OK : fail(VH_ERR_KERNEL, std::string("megakernel samples: ") + hipGetErrorString(e));
}

int launch_wf_sample(const Scene& S, const Frame& F, const Params& P0, const Wavefront& W, bool restir, bool split, const LaunchCfg& cfg,
                     StageMarks* marks)
{
    hipStream_t s = cfg.stream;
    uint32_t grid = cfg.wfGrid ? cfg.wfGrid : 1024;
    const uint32_t tgrid = std::max(grid * kTraceScale, W.traceWaves); // one-wave blocks
    Params P = P0;
    uint32_t maxDepth = P.w[7];
    // ReSTIR DI defers its winner's shadow ray with the others; GI's reuse runs first, in k_gi_prepare
    const bool gi = restir && (P.w[36] & 2u) != 0;
    const bool defer = W.shK > 0 && !split;
    hipError_t me = hipSuccess;
    auto mark = [&](uint8_t stage) {
        if (marks && me == hipSuccess)
            me = marks->mark(s, stage);
    };
    auto shade = [&](uint32_t d) {
        if (gi && d == 0)
            k_gi_prepare<<<grid * kTraceScale, kTraceBlock, 0, s>>>(S, F, P, W);
        if (gi)
            k_wf_shade_deferred<true, true><<<grid, kWfBlock, 0, s>>>(S, F, P, W, d);
        else if (restir && d == 0)
            k_wf_shade_deferred<false, true><<<grid, kWfBlock, 0, s>>>(S, F, P, W, d);
        else
            k_wf_shade_deferred<false, false><<<grid, kWfBlock, 0, s>>>(S, F, P, W, d);
        mark(kStageShade);
    };
    auto resolve = [&](uint32_t d) {
        if (gi)
            k_wf_resolve<true><<<grid, kWfBlock, 0, s>>>(S, F, P, W, d);
        else
            k_wf_resolve<false><<<grid, kWfBlock, 0, s>>>(S, F, P, W, d);
        mark(kStageResolve);
    };
    k_wf_reset<<<1, 64, 0, s>>>(W.counts, W.n, P.w[32], maxDepth);
    if (P.w[32] != 0)
        k_wf_select<<<blocks_for(W.n, kWfBlock, grid), kWfBlock, 0, s>>>(F, P, W);
    mark(kStages);
    if (defer && !S.translucent) {
        // opaque scenes: depth d's closest rays and depth d - 1's shadow rays in one launch, then
        // depth d - 1's resolve, then depth d's shading; a last launch for depth maxDepth - 1's
        // shadow rays (counts[maxDepth] stays 0)
        for (uint32_t d = 0; d <= maxDepth; d++) {
            P.w[30] = d;
            k_wf_trace<<<tgrid, kTraceBlock, 0, s>>>(S, P, W, d);
            mark(kStageTrace);
            if (d > 0)
                resolve(d - 1);
            if (d < maxDepth)
                shade(d);
        }
    } else {
        for (uint32_t d = 0; d < maxDepth; d++) {
            P.w[30] = d;
            k_wf_extend<<<tgrid, kTraceBlock, 0, s>>>(S, P, W, d);
            mark(kStageExtend);
            if (defer) { // translucent scenes: closest-hit shadow rays through the layers, per lane
                shade(d);
                k_wf_shadow<<<tgrid, kTraceBlock, 0, s>>>(S, P, W, d);
                mark(kStageShadow);
                resolve(d);
            } else {
                if (restir && split)
                    k_wf_shade<true, true><<<grid, kWfBlock, 0, s>>>(S, F, P, W, d);
                else if (restir)
                    k_wf_shade<true, false><<<grid, kWfBlock, 0, s>>>(S, F, P, W, d);
                else if (split)
                    k_wf_shade<false, true><<<grid, kWfBlock, 0, s>>>(S, F, P, W, d);
                else
                    k_wf_shade<false, false><<<grid, kWfBlock, 0, s>>>(S, F, P, W, d);
                mark(kStageShade);
            }
        }
    }
    if (me != hipSuccess)
        return fail(VH_ERR_KERNEL, std::string("stage timing: ") + hipGetErrorString(me));
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? VH_OK : fail(VH_ERR_KERNEL, std::string("wavefront launch: ") + hipGetErrorString(e));
}

// The unit direction a directional (or virtual) light's shadow rays take, -normalize(its words
// 5-7), once per light instead of at every shading vertex: the same expression, so the same bits.
__global__ void k_light_dirs(const uint32_t* __restrict__ lights, uint32_t count, float* __restrict__ out)
{
    for (uint32_t li = blockIdx.x * blockDim.x + threadIdx.x; li < count; li += gridDim.x * blockDim.x) {
        const uint32_t* L = lights + li * 16;
        const V3 d = -normalize(v3(asf(L[5]), asf(L[6]), asf(L[7])));
        out[4 * li] = d.x;
        out[4 * li + 1] = d.y;
        out[4 * li + 2] = d.z;
        out[4 * li + 3] = 0.0f;
    }
}

int launch_light_dirs(const uint32_t* lights, uint32_t count, float* out, hipStream_t s)
{
    if (count == 0)
        return VH_OK;
    k_light_dirs<<<(count + 63) / 64, 64, 0, s>>>(lights, count, out);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? VH_OK : fail(VH_ERR_KERNEL, std::string("light directions: ") + hipGetErrorString(e));
}

int launch_film_add_pass(double* film, double* pass, uint64_t n, hipStream_t s)
{
    k_film_add_pass<<<blocks_for(n / 2 + 1, 256, 2048), 256, 0, s>>>(film, pass, n);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? VH_OK : fail(VH_ERR_KERNEL, std::string("film add: ") + hipGetErrorString(e));
}

int launch_nrc_train_paths(const Scene& S, const Params& P, const Wavefront& W, uint32_t* trainPix, uint32_t nTrain, uint32_t npix,
                           uint32_t frame, float4* vertices, float4* samples, uint32_t* sampleCount, uint32_t sampleCap,
                           const LaunchCfg& cfg)
{
    hipStream_t s = cfg.stream;
    uint32_t grid = blocks_for(nTrain, kWfBlock, cfg.wfGrid ? cfg.wfGrid : 1024);
    k_nrc_pick<<<(nTrain + 255) / 256, 256, 0, s>>>(trainPix, nTrain, npix, frame, W.counts, sampleCount);
    const uint32_t tgrid = std::max(grid * kTraceScale, W.traceWaves); // one-wave blocks
    for (uint32_t d = 0; d < P.w[7]; d++) {
        k_nrc_t_extend<<<tgrid, kTraceBlock, 0, s>>>(S, P, W, trainPix, d);
        if (W.shK > 0) { // deferred shadow rays, traced by the wave-cooperative shadow kernel
            k_nrc_t_shade_deferred<<<grid, kWfBlock, 0, s>>>(S, P, W, trainPix, vertices, samples, sampleCount, sampleCap, d);
            k_wf_shadow<<<tgrid, kTraceBlock, 0, s>>>(S, P, W, d);
            k_nrc_t_resolve<<<grid, kWfBlock, 0, s>>>(W, vertices, samples, sampleCount, sampleCap, d);
        } else {
            k_nrc_t_shade<<<grid, kWfBlock, 0, s>>>(S, P, W, trainPix, vertices, samples, sampleCount, sampleCap, d);
        }
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? VH_OK : fail(VH_ERR_KERNEL, std::string("NRC training paths: ") + hipGetErrorString(e));
}

int launch_wf_reset(const Params& P, const Wavefront& W, hipStream_t s)
{
    k_wf_reset<<<1, 64, 0, s>>>(W.counts, W.n, 0u, P.w[7]);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? VH_OK : fail(VH_ERR_KERNEL, std::string("wavefront reset: ") + hipGetErrorString(e));
}

int launch_wf_depth(const Scene& S, const Frame& F, const Params& P0, const Wavefront& W, uint32_t d, bool shade, const LaunchCfg& cfg,
                    bool leaveShadows)
{
    Params P = P0;
    P.w[30] = d;
    uint32_t grid = cfg.wfGrid ? cfg.wfGrid : 1024;
    const uint32_t tgrid = std::max(grid * kTraceScale, W.traceWaves); // one-wave blocks
    k_wf_extend<<<tgrid, kTraceBlock, 0, cfg.stream>>>(S, P, W, d);
    if (shade && W.shK > 0) {
        k_wf_shade_deferred<false, false><<<grid, kWfBlock, 0, cfg.stream>>>(S, F, P, W, d);
        if (!leaveShadows) {
            k_wf_shadow<<<tgrid, kTraceBlock, 0, cfg.stream>>>(S, P, W, d);
            k_wf_resolve<false><<<grid, kWfBlock, 0, cfg.stream>>>(S, F, P, W, d);
        }
    } else if (shade) {
        k_wf_shade<false, false><<<grid, kWfBlock, 0, cfg.stream>>>(S, F, P, W, d);
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? VH_OK : fail(VH_ERR_KERNEL, std::string("wavefront depth: ") + hipGetErrorString(e));
}

int launch_wf_trace_resolve(const Scene& S, const Frame& F, const Params& P0, const Wavefront& W, uint32_t d, const LaunchCfg& cfg)
{
    if (d == 0 || W.shK == 0 || S.translucent)
        return fail(VH_ERR_INVALID_ARGUMENT, "merged trace: needs depth >= 1 of an opaque scene with deferred shadow rays");
    Params P = P0;
    P.w[30] = d;
    uint32_t grid = cfg.wfGrid ? cfg.wfGrid : 1024;
    const uint32_t tgrid = std::max(grid * kTraceScale, W.traceWaves); // one-wave blocks
    k_wf_trace<<<tgrid, kTraceBlock, 0, cfg.stream>>>(S, P, W, d);
    k_wf_resolve<false><<<grid, kWfBlock, 0, cfg.stream>>>(S, F, P, W, d - 1);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? VH_OK : fail(VH_ERR_KERNEL, std::string("wavefront trace: ") + hipGetErrorString(e));
}

int launch_nrc_query_shade(const Scene& S, const Frame& F, const Params& P0, const Wavefront& W, float4* queries, uint32_t d,
                           const LaunchCfg& cfg)
{
    Params P = P0;
    P.w[30] = d;
    k_nrc_q_shade<<<cfg.wfGrid ? cfg.wfGrid : 1024, kWfBlock, 0, cfg.stream>>>(S, F, P, W, queries, d);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? VH_OK : fail(VH_ERR_KERNEL, std::string("NRC query shade: ") + hipGetErrorString(e));
}

int launch_primary_hits(const Scene& S, const Params& P, uint4* out, hipStream_t s)
{
    uint64_t n = (uint64_t)P.w[0] * P.w[1];
    k_primary<<<blocks_for(n, 256, 8192), 256, 0, s>>>(S, P, out);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? VH_OK : fail(VH_ERR_KERNEL, std::string("primary launch: ") + hipGetErrorString(e));
}

int launch_trace_rays(const Scene& S, const float* rays, uint64_t n, float tmax, uint4* out, hipStream_t s)
{
    k_rays<<<blocks_for(n, 256, 8192), 256, 0, s>>>(S, rays, n, tmax, out);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? VH_OK : fail(VH_ERR_KERNEL, std::string("ray launch: ") + hipGetErrorString(e));
}

} // namespace vh
