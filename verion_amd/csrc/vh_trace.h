// vh_trace.h - ray queries over the LBVH: closest hit and shadow (any) hit.
//
// Replaces the driver ray queries of pt_core.slang `traceClosest` (:684-700) and `occluded`
// (:580-613). The BVH is 4-wide: a node is 128 bytes (one L2 line of MI355X), eight 16-byte words:
//   0..5  child boxes as SoA: lo.x[4], hi.x[4], lo.y[4], hi.y[4], lo.z[4], hi.z[4]
//   6     child references: a node index, or for a leaf the first triangle in leaf order
//   7     triangle counts (0 = inner node; empty slots have an inverted box and never hit)
// Triangles are stored in leaf order, 48 bytes each: (p0, global id) (p1, instance) (p2, primitive).
// Boxes are padded at build time, so the slab test is conservative; the triangle test is the
// watertight one of Woop, Benthin and Wald (JCGT 2013) without culling, and the closest hit is the
// smallest t with ties to the smallest global triangle id (= smallest (instance, primitive)),
// which makes the answer independent of the tree.
#pragma once

#include "vh_math.h"

namespace vh {

struct Hit {
    bool hit;
    uint32_t inst, prim;
    float t;
    float b1, b2; // barycentric weights of vertices 1 and 2 (for opacity textures)
    uint32_t gid;
};

struct Bvh {
    const float4* nodes; // 8 float4 per node; node 0 is the root
    const float4* tris;  // 3 float4 per triangle, in leaf order
    const uint2* ids;    // per global triangle id: instance, primitive
    uint32_t ntri;
    // Optional traversal counters (null when off): [0] closest-hit rays, [1] their nodes visited,
    // [2] their triangles tested, [3..5] the same for shadow (any-hit) rays, [6] rays that hit.
    unsigned long long* counters;
};

// Watertight ray/triangle test. Inputs are the world-space vertices; on a hit with 0 < t < tmax
// it returns t and (BARY) the weights of vertices 1 and 2.
struct RayPre {
    int kx, ky, kz;
    float sx, sy, sz;
};

VH_HD RayPre ray_pre(V3 d)
{
    float dv[3] = {d.x, d.y, d.z};
    float ad[3] = {fabsf(d.x), fabsf(d.y), fabsf(d.z)};
    RayPre r;
    r.kz = ad[0] > ad[1] ? (ad[0] > ad[2] ? 0 : 2) : (ad[1] > ad[2] ? 1 : 2);
    r.kx = r.kz + 1 == 3 ? 0 : r.kz + 1;
    r.ky = r.kx + 1 == 3 ? 0 : r.kx + 1;
    if (dv[r.kz] < 0.0f) {
        int t = r.kx;
        r.kx = r.ky;
        r.ky = t;
    }
    r.sx = dv[r.kx] / dv[r.kz];
    r.sy = dv[r.ky] / dv[r.kz];
    r.sz = 1.0f / dv[r.kz];
    return r;
}

VH_HD float pick(V3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// a*b - c*d with a relative error of at most 2 ulp (Cornea, Harrison and Tang 2003; bound by
// Jeannerod, Louvet and Muller 2013): both products are split exactly by FMA, so an exact zero
// stays 0, any other value keeps its sign, and diffProd(c, d, a, b) is exactly -diffProd(a, b, c, d)
// (the same edge seen from the triangle on its other side). The watertight test takes it where an
// edge function rounds to zero in single precision (the case Woop et al. recompute in double).
VH_HD float diffProd(float a, float b, float c, float d)
{
    float p1 = a * b, p2 = c * d;
    float e1 = fmaf(a, b, -p1), e2 = fmaf(c, d, -p2);
    return (p1 - p2) + (e1 - e2);
}

template <bool BARY = true>
VH_HD bool intersect(const RayPre& r, V3 o, V3 p0, V3 p1, V3 p2, float tmax, float& tout, float& b1, float& b2)
{
    V3 A = p0 - o, B = p1 - o, C = p2 - o;
    float Akz = pick(A, r.kz), Bkz = pick(B, r.kz), Ckz = pick(C, r.kz);
    float ax = pick(A, r.kx) - r.sx * Akz, ay = pick(A, r.ky) - r.sy * Akz;
    float bx = pick(B, r.kx) - r.sx * Bkz, by = pick(B, r.ky) - r.sy * Bkz;
    float cx = pick(C, r.kx) - r.sx * Ckz, cy = pick(C, r.ky) - r.sy * Ckz;
    float U = cx * by - cy * bx;
    float V = ax * cy - ay * cx;
    float W = bx * ay - by * ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) { // an edge through the ray within rounding: decide it exactly
        U = diffProd(cx, by, cy, bx);
        V = diffProd(ax, cy, ay, cx);
        W = diffProd(bx, ay, by, ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f))
        return false;
    float det = U + V + W;
    if (det == 0.0f)
        return false;
    float T = U * (r.sz * Akz) + V * (r.sz * Bkz) + W * (r.sz * Ckz);
    // cheap rejections before the division, each implying the exact test below would reject:
    // t = T / det <= 0 when the signs differ or T is 0; t >= tmax when |T| exceeds |det| tmax by
    // more than the rounding of the two products (only while that product is a normal float)
    if (det > 0.0f ? T <= 0.0f : T >= 0.0f)
        return false;
    float lim = fabsf(det) * (tmax * 1.000001f);
    if (lim >= 1.0e-30f && fabsf(T) > lim)
        return false;
    float t = T / det;
    if (!(t > 0.0f && t < tmax))
        return false;
    tout = t;
    if (BARY) {
        b1 = V / det;
        b2 = W / det;
    }
    return true;
}

__device__ __forceinline__ V3 xyz(float4 a) { return v3(a.x, a.y, a.z); }

constexpr int kStack = 96;
// Stack entries a traversal may need in a tree of `depth` levels of 4-wide nodes (the root on level
// 1): a node on level L is reached with at most 3 pending siblings per ancestor on the stack and
// pushes at most 4 children, so at most 3 (depth - 1) + 4; the renderer keeps a margin of 6.
__host__ __device__ constexpr uint32_t stackNeed(uint32_t depth) { return 3 * (depth + 1) + 4; }
// The deepest tree the renderer accepts: stackNeed(kMaxTreeDepth) <= kStack < stackNeed(kMaxTreeDepth + 1).
constexpr uint32_t kMaxTreeDepth = ((uint32_t)kStack - 4) / 3 - 1;
static_assert(stackNeed(kMaxTreeDepth) <= (uint32_t)kStack && stackNeed(kMaxTreeDepth + 1) > (uint32_t)kStack, "depth bound");
// Leaves hold at most kLeafMax triangles (the build makes a leaf of every subtree that small); a
// node's pending leaf triangles are kept as 4 groups of kLeafGroup bits, one group per child slot.
// 1080p procedural scene: leaves of up to 8 visit 8% fewer nodes but test 85% more triangles
// (447 -> 431 Msamples/s, and the larger WaveTris lowers the traversal kernels' occupancy to 4).
#ifndef VH_LEAF_MAX
#define VH_LEAF_MAX 4
#endif
constexpr uint32_t kLeafMax = VH_LEAF_MAX;
static_assert(kLeafMax == 4 || kLeafMax == 8, "leaf size");
constexpr uint32_t kLeafBits = kLeafMax == 8 ? 3u : 2u;
constexpr uint32_t kLeafGroup = 1u << kLeafBits;
// Child reference of an empty slot of a 4-wide node (nodes with fewer than 4 children: the root of
// a scene of at most 4 triangles). Its box is inverted (lo = +inf, hi = -inf), which the FMA slab
// test would take for an infinite box, so the traversal skips it by this reference.
constexpr uint32_t kEmptyChild = 0xFFFFFFFFu;

// One node's eight 16-byte words, the box rows ordered by the ray's entry side: per axis the row
// of the planes the ray enters through (e) and leaves through (f). `near` holds, per axis, which of
// the node's lo / hi rows the ray enters through (bit a: the ray's inverse direction on axis a is
// negative, so it enters through hi); taking each axis's planes by address instead of a min and a
// max per child gives the same distances as min(lo, hi) / max(lo, hi) of the two slab distances for
// every box (lo <= hi), and an empty slot (lo = +inf, hi = -inf) never hits. All eight loads are
// issued before any of them is waited for (the scheduling barrier keeps the compiler from
// interleaving the first uses, which would wait for the child words before the last box rows are
// requested: two memory round trips per node instead of one).
struct NodeRows {
    float4 ex, fx, ey, fy, ez, fz;
    uint4 child, count;
};

__device__ __forceinline__ NodeRows loadNode(const float4* n, uint32_t near)
{
    const uint32_t sx = near & 1u, sy = (near >> 1) & 1u, sz = near >> 2;
    NodeRows r;
    r.child = *reinterpret_cast<const uint4*>(n + 6);
    r.count = *reinterpret_cast<const uint4*>(n + 7);
    r.ex = n[sx];
    r.fx = n[1u - sx];
    r.ey = n[2u + sy];
    r.fy = n[3u - sy];
    r.ez = n[4u + sz];
    r.fz = n[5u - sz];
    __builtin_amdgcn_sched_barrier(0);
    return r;
}

// Box test of the four children: returns a 4-bit mask of hits and their entry distances.
__device__ __forceinline__ uint32_t slab4(const NodeRows& nr, float ix, float iy, float iz, float ox, float oy, float oz, float tmax,
                                          float tn[4])
{
    const float4 ex = nr.ex, fx = nr.fx, ey = nr.ey, fy = nr.fy, ez = nr.ez, fz = nr.fz;
    float a[6][4] = {{ex.x, ex.y, ex.z, ex.w}, {fx.x, fx.y, fx.z, fx.w}, {ey.x, ey.y, ey.z, ey.w},
                     {fy.x, fy.y, fy.z, fy.w}, {ez.x, ez.y, ez.z, ez.w}, {fz.x, fz.y, fz.z, fz.w}};
    uint32_t mask = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        float x0 = fmaf(a[0][c], ix, ox), x1 = fmaf(a[1][c], ix, ox);
        float y0 = fmaf(a[2][c], iy, oy), y1 = fmaf(a[3][c], iy, oy);
        float z0 = fmaf(a[4][c], iz, oz), z1 = fmaf(a[5][c], iz, oz);
        float t0 = fmaxf(fmaxf(x0, y0), fmaxf(z0, 0.0f));
        float t1 = fminf(fminf(x1, y1), fminf(z1, tmax)) * 1.0000004f;
        tn[c] = t0;
        mask |= (t0 <= t1 ? 1u : 0u) << c;
    }
    return mask;
}

__device__ __forceinline__ float safe_inv(float d)
{
    return 1.0f / (fabsf(d) > 1.0e-30f ? d : (d < 0.0f ? -1.0e-30f : 1.0e-30f));
}

template <bool ANY>
__device__ __forceinline__ void count_traversal(const Bvh& bvh, uint32_t rays, uint32_t visited, uint32_t tested, bool hit)
{
    if (bvh.counters) {
        // 64 banks of 8 counters, picked by block, keep the atomics from serialising on one address
        unsigned long long* bank = bvh.counters + 8 * (blockIdx.x & 63u);
        unsigned long long* c = bank + (ANY ? 3 : 0);
        atomicAdd(c, (unsigned long long)rays);
        atomicAdd(c + 1, (unsigned long long)visited);
        atomicAdd(c + 2, (unsigned long long)tested);
        if (hit && !ANY)
            atomicAdd(bank + 6, 1ull);
    }
}

// Traversal of one ray as a resumable state machine: traceBegin sets up the ray, each traceStep
// visits one node (its box tests, the triangles of its leaf children, the push of its inner
// children and the pop of the next node) and returns true when the ray is done. trace() runs the
// steps to the end; the persistent wavefront kernels interleave the steps of many rays per lane.
struct TraceState {
    V3 o;
    float ix, iy, iz, ox, oy, oz;
    uint32_t near; // per axis (bit a): the inverse direction is negative (slab4)
    RayPre pre;
    uint32_t axes;    // kx | ky << 2 | kz << 4 (the cooperative traversal hands the ray to other lanes)
    float tmax, best; // the ray's limit, and the closest accepted t (or tmax)
    Hit h;
    uint32_t node;
    int sp;
    uint32_t visited, tested;
    uint32_t ray; // wave traversal: the lane whose ray this lane's node belongs to
};

__device__ __forceinline__ void traceBegin(TraceState& s, V3 o, V3 d, float tmax)
{
    s.h.hit = false;
    s.h.inst = s.h.prim = s.h.gid = 0;
    s.h.t = 0.0f;
    s.h.b1 = s.h.b2 = 0.0f;
    s.o = o;
    s.pre = ray_pre(d);
    s.axes = (uint32_t)s.pre.kx | ((uint32_t)s.pre.ky << 2) | ((uint32_t)s.pre.kz << 4);
    s.ix = safe_inv(d.x);
    s.iy = safe_inv(d.y);
    s.iz = safe_inv(d.z);
    s.ox = -o.x * s.ix;
    s.oy = -o.y * s.iy;
    s.oz = -o.z * s.iz;
    s.near = (s.ix < 0.0f ? 1u : 0u) | (s.iy < 0.0f ? 2u : 0u) | (s.iz < 0.0f ? 4u : 0u);
    s.tmax = tmax;
    s.best = tmax;
    s.node = 0;
    s.sp = 0;
    s.visited = 0;
    s.tested = 0;
}

__device__ __forceinline__ uint32_t sel4(uint4 v, uint32_t c) { return c == 0 ? v.x : (c == 1 ? v.y : (c == 2 ? v.z : v.w)); }

// Compare-exchange of the sort of the inner children (nearest first).
__device__ __forceinline__ void cmpx(float& ka, uint32_t& va, float& kb, uint32_t& vb)
{
    bool sw = kb < ka;
    float k = sw ? kb : ka;
    uint32_t v = sw ? vb : va;
    kb = sw ? ka : kb;
    vb = sw ? va : vb;
    ka = k;
    va = v;
}

// Closest hit (any = false) or the first hit found (any = true) over (0, tmax). One node per step:
// the box test of its four children; the triangles of the leaf children hit, gathered as bits
// kLeafGroup c + k into one loop (a wave then runs the most any lane has, not one loop per child slot);
// the inner children hit, sorted by a 5-exchange network and pushed farthest first; then the pop
// of the next node that can still hold a closer hit. Every triangle that can hold the closest hit
// (ties up to the next float included) is tested, so the hit does not depend on the visiting order.
template <bool ANY>
__device__ __forceinline__ bool traceStep(const Bvh& bvh, TraceState& s, uint32_t* stackN, float* stackT)
{
    Hit& h = s.h;
    const float4* n = bvh.nodes + 8 * (size_t)s.node;
    s.visited++;
    const NodeRows nr = loadNode(n, s.near);
    const uint4 child = nr.child, count = nr.count;
    float tn[4];
    uint32_t mask = slab4(nr, s.ix, s.iy, s.iz, s.ox, s.oy, s.oz, s.best, tn);
    const float kInf = __builtin_inff();
    uint32_t tris = 0;
    float key[4];
    uint32_t val[4];
#pragma unroll
    for (uint32_t c = 0; c < 4; c++) {
        bool hitc = (mask >> c & 1u) != 0 && sel4(child, c) != kEmptyChild;
        uint32_t cnt = sel4(count, c);
        if (hitc && cnt != 0 && !(h.hit && tn[c] > h.t))
            tris |= ((1u << cnt) - 1u) << (kLeafGroup * c);
        key[c] = hitc && cnt == 0 ? tn[c] : kInf;
        val[c] = sel4(child, c);
    }
    while (tris != 0) {
        uint32_t b = (uint32_t)__builtin_ctz(tris);
        tris &= tris - 1u;
        uint32_t c = b >> kLeafBits;
        float tc = c == 0 ? tn[0] : (c == 1 ? tn[1] : (c == 2 ? tn[2] : tn[3]));
        if (h.hit && tc > h.t) // a closer hit was found in this node meanwhile
            continue;
        s.tested++;
        const float4* tr = bvh.tris + 3 * (size_t)(sel4(child, c) + (b & (kLeafGroup - 1u)));
        float4 a = tr[0], bb = tr[1], cc = tr[2];
        uint32_t gid = asu(a.w);
        float t, b1, b2;
        // ties are kept for the id comparison: test up to the next float above the best t (> 0)
        float lim = h.hit ? asf(asu(h.t) + 1u) : s.tmax;
        if (intersect(s.pre, s.o, xyz(a), xyz(bb), xyz(cc), lim, t, b1, b2) && t < s.tmax &&
            (!h.hit || t < h.t || (t == h.t && gid < h.gid))) {
            h.hit = true;
            h.t = t;
            h.b1 = b1;
            h.b2 = b2;
            h.gid = gid;
            h.inst = asu(bb.w);
            h.prim = asu(cc.w);
            s.best = t;
            if (ANY)
                return true;
        }
    }
    if (!ANY) { // an any-hit ray takes the children in slot order: any occluder ends it
        cmpx(key[0], val[0], key[1], val[1]);
        cmpx(key[2], val[2], key[3], val[3]);
        cmpx(key[0], val[0], key[2], val[2]);
        cmpx(key[1], val[1], key[3], val[3]);
        cmpx(key[1], val[1], key[2], val[2]);
    }
    // the inner children hit, nearest first (closest hit), are pushed farthest first
#pragma unroll
    for (int i = 3; i >= 0; i--) {
        if (key[i] < kInf && s.sp < kStack) { // the build bounds the depth: the stack cannot fill
            stackN[s.sp] = val[i];
            if (!ANY) // an any-hit ray ends at its first hit: entry distances are never compared
                stackT[s.sp] = key[i];
            s.sp++;
        }
    }
    // pop the next node that can still hold a closer hit
    while (s.sp > 0) {
        s.sp--;
        if (!ANY && h.hit && stackT[s.sp] > h.t)
            continue;
        s.node = stackN[s.sp];
        return false;
    }
    return true;
}

// ---------------------------------------------------------------- wave-cooperative traversal
//
// In traceStep a wave runs its triangle loop as many times as its busiest lane has leaf triangles
// in the node (up to 16) while most lanes wait: ~0.45 tests per lane and node step against a
// maximum of 4 to 8. traceWave has the wave's 64 lanes step their rays together and share the
// triangle tests of each step through LDS: the lanes list their pending (ray, triangle) tests, every
// lane takes the next 64 of them (usually all), reads the listing lane's ray from that lane's
// registers (ds_bpermute) and keeps the smallest (t, id) per ray through a 64-bit LDS atomic
// minimum. A ray's tests within one step use the limit of the step's start instead of one tightened
// after each test, which only adds tests whose hits lose: the hit, the smallest t with ties to the
// smallest id, is the same. Lanes whose ray is done keep testing the other lanes' triangles, and
// take over parts of the other rays' traversals (work stealing, below), until the wave's last ray
// ends. Every node a ray reaches is visited once, by one lane, and every test that can hold the
// hit is made whichever lane makes it, so the hit does not depend on who did which part. All 64
// lanes of a wave must call it together (the wave-uniform loops of the wavefront kernels). Its
// callers need the hit's ids and distance only (shading recomputes the barycentrics from the hit
// point), so it returns no barycentric weights (b1 = b2 = 0).
//
// Short stack: the first kShort entries of each lane's traversal stack live in LDS (entry i of
// lane l at [i][l], so a wave's accesses never conflict), the rest in the lane's overflow array.
// On the 1080p procedural scene 0.06% of closest-hit steps hold more than 12 entries (0.5% on the
// 10,000-emitter scene); the deepest stacks seen were 23 and 29 entries. A closest-hit entry keeps
// its box entry distance rounded down to bfloat16 (its upper 16 bits): an entry is skipped when
// that lower bound lies beyond the current hit, so rounding only keeps an entry the exact distance
// would have skipped, never skips one that could hold the hit.

constexpr uint32_t kWaveTests = 64 * 4 * kLeafGroup; // 64 lanes x 4 leaf children x kLeafMax triangles at most
#ifndef VH_SHORT_STACK
#define VH_SHORT_STACK 12
#endif
constexpr int kShort = VH_SHORT_STACK;               // stack entries per lane in LDS
constexpr int kSpill = kStack - kShort;              // the rest, in the lane's overflow array

struct WaveTris {
    unsigned long long key[64];  // per ray (its owner lane): (t bits << 32) | global id of the best hit so far, ~0 for none
    uint32_t stackN[kShort][64]; // the first kShort stack entries of each lane: node
    uint16_t stackT[kShort][64]; // and (closest hit) the upper half of its entry distance's bits
    uint16_t list[kWaveTests];   // pending tests: listing lane << (2 + kLeafBits) | child slot << kLeafBits | triangle;
                                 // while work is handed out, the offered entries: donor lane << 4 | stack index
};
static_assert(sizeof(WaveTris) == 512 + 6 * 64 * kShort + 2048, "LDS of one wave's traversal");

__device__ __forceinline__ void waveSync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// number of set bits of m in the lanes below this one
__device__ __forceinline__ uint32_t lanesBelow(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Inclusive sum of v over the wave's lanes in lane order (every lane executes it): Hillis-Steele
// steps inside each row of 16 lanes by DPP row shifts (lanes before the row's start read 0), then
// row 0's total added to row 1 and row 2's to row 3 (row broadcast of lane 15), then lane 31's
// total to rows 2 and 3 (row broadcast of lane 31). Six DPP adds where a sum of per-bit ballots
// needs one ballot, two lane counts and a shift-add per bit.
__device__ __forceinline__ uint32_t waveInclusiveSum(uint32_t v)
{
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true); // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true); // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true); // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true); // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false); // row_bcast:15 -> rows 1, 3
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false); // row_bcast:31 -> rows 2, 3
    return v;
}

// v of lane `src` (every lane of the wave executes it: an inactive source lane would read as 0)
__device__ __forceinline__ uint32_t fromLane(uint32_t src, uint32_t v) { return (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src << 2), (int)v); }
__device__ __forceinline__ float fromLane(uint32_t src, float v) { return asf(fromLane(src, asu(v))); }

// Work stealing inside a wave: once at least kStealMinIdle lanes have no node, every lane with at
// most kShort stack entries offers its top kOfferMax entries and the idle lanes take them in lane
// order. 1080p procedural scene (wave steps x 64 over node visits): lane use 0.50 -> 0.83 for
// closest-hit and 0.39 -> 0.81 for shadow rays, for 11% and 17% more node visits (a taken subtree
// is searched before the owner's later hits could cull it); per trace launch 0.257 -> 0.175 ms.
// Measured: 1, 4, 8, 16, 24 and 32 idle lanes and 1, 2, 4 and 7 entries per offer; 16 and 2 (or 1
// or 4) within 1% of the best, 1 idle lane 1.5% and 32 2% slower.
constexpr uint32_t kStealMinIdle = 16; // idle lanes a wave needs before it hands out work
constexpr uint32_t kOfferMax = 2;      // stack entries one lane gives away per step
static_assert(kShort <= 16, "offered entries are named by 4-bit stack indices");

// best (t, id) of ray `r` so far (sh.key[r], ~0 for none) as (hit, t)
__device__ __forceinline__ bool bestOf(const WaveTris& sh, uint32_t r, float& t)
{
    const unsigned long long k = sh.key[r];
    t = asf((uint32_t)(k >> 32));
    return k != ~0ull;
}

// One traversal step of the wave (every lane calls it): each live lane visits one node of the ray
// it works for (s.ray: its own, or another lane's it took work from), the wave tests all their
// leaf triangles together into that ray's best (t, id) in LDS, then each lane pushes its inner
// children and pops the next node; `live` turns false when the lane has no node left. Lanes left
// without a node then take the top stack entries of lanes that still have some (work stealing
// inside the wave), with a copy of that ray, so a wave no longer runs as long as its longest ray.
template <bool ANY>
__device__ __forceinline__ void waveStep(const Bvh& bvh, TraceState& s, bool& live, WaveTris& sh, uint32_t* spillN, float* spillT)
{
    const uint32_t lane = threadIdx.x & 63u;
    const float kInf = __builtin_inff();
    uint32_t tris = 0;
    float key[4] = {kInf, kInf, kInf, kInf};
    uint32_t val[4] = {0u, 0u, 0u, 0u};
    uint4 child = make_uint4(0u, 0u, 0u, 0u);
    float bt;
    const bool bh = bestOf(sh, s.ray, bt);
    // the largest t a triangle can have and still matter this step (ties kept for the id comparison)
    const float lim = bh ? asf(asu(bt) + 1u) : s.tmax;
    if (live) {
        const float4* n = bvh.nodes + 8 * (size_t)s.node;
        s.visited++;
        const NodeRows nr = loadNode(n, s.near);
        child = nr.child;
        const uint4 count = nr.count;
        float tn[4];
        uint32_t mask = slab4(nr, s.ix, s.iy, s.iz, s.ox, s.oy, s.oz, bh ? bt : s.tmax, tn);
#pragma unroll
        for (uint32_t c = 0; c < 4; c++) {
            bool hitc = (mask >> c & 1u) != 0 && sel4(child, c) != kEmptyChild;
            uint32_t cnt = sel4(count, c);
            if (hitc && cnt != 0 && !(bh && tn[c] > bt))
                tris |= ((1u << cnt) - 1u) << (kLeafGroup * c);
            key[c] = hitc && cnt == 0 ? tn[c] : kInf;
            val[c] = sel4(child, c);
        }
        if (!ANY) {
            cmpx(key[0], val[0], key[1], val[1]);
            cmpx(key[2], val[2], key[3], val[3]);
            cmpx(key[0], val[0], key[2], val[2]);
            cmpx(key[1], val[1], key[3], val[3]);
            cmpx(key[1], val[1], key[2], val[2]);
        }
        // the inner children hit, nearest first (closest hit), are pushed farthest first; before the
        // triangle tests, which leaves their registers free (the pop after the tests skips what the
        // tests' hits put out of reach, and an any-hit ray that hits ends with its pushes unread).
        // Child i goes to slot sp + (children pushed before it: those hit with a larger index).
        const uint32_t vm = (key[0] < kInf ? 1u : 0u) | (key[1] < kInf ? 2u : 0u) | (key[2] < kInf ? 4u : 0u) | (key[3] < kInf ? 8u : 0u);
        if (__ballot(s.sp > kShort - 4) == 0) {
            // every live lane has room for four entries in LDS: four stores without branches, a child
            // not pushed going to the top slot, which lies above every pushed one (sp + 4 <= kShort)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int slot = ((vm >> i) & 1u) != 0 ? s.sp + __builtin_popcount(vm >> (i + 1)) : kShort - 1;
                sh.stackN[slot][lane] = val[i];
                if (!ANY) // an any-hit ray ends at its first hit: entry distances are never compared
                    sh.stackT[slot][lane] = (uint16_t)(asu(key[i]) >> 16);
            }
            s.sp += __builtin_popcount(vm);
        } else {
#pragma unroll
            for (int i = 3; i >= 0; i--) {
                if (((vm >> i) & 1u) != 0 && s.sp < kStack) { // the build bounds the depth: the stack cannot fill
                    if (s.sp < kShort) {
                        sh.stackN[s.sp][lane] = val[i];
                        if (!ANY)
                            sh.stackT[s.sp][lane] = (uint16_t)(asu(key[i]) >> 16);
                    } else {
                        spillN[s.sp - kShort] = val[i];
                        if (!ANY)
                            spillT[s.sp - kShort] = key[i];
                    }
                    s.sp++;
                }
            }
        }
    }
    // exclusive prefix of the per-lane test counts (at most 4 kLeafMax) over the wave, and the total
    const uint32_t cnt = (uint32_t)__builtin_popcount(tris);
    const uint32_t incl = waveInclusiveSum(cnt);
    const uint32_t pre = incl - cnt;
    const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    if (total != 0) {
        for (uint32_t m = tris, j = pre; m != 0; m &= m - 1u, j++)
            sh.list[j] = (uint16_t)((lane << (2 + kLeafBits)) | (uint32_t)__builtin_ctz(m));
        s.tested += cnt;
        waveSync();
        for (uint32_t r = 0; r < total; r += 64) {
            const uint32_t q = r + lane;
            const uint32_t e = q < total ? (uint32_t)sh.list[q] : 0u;
            const uint32_t owner = e >> (2 + kLeafBits);
            const uint32_t c = (e >> kLeafBits) & 3u, k = e & (kLeafGroup - 1u);
            // from the listing lane's registers (all 64 lanes here): its child references first, so
            // the triangle's fetch waits on them alone, then its ray and limit, which arrive while
            // the fetch is in flight
            const uint32_t c0 = fromLane(owner, child.x), c1 = fromLane(owner, child.y), c2 = fromLane(owner, child.z),
                           c3 = fromLane(owner, child.w);
            const uint32_t first = (c & 2u) != 0 ? ((c & 1u) != 0 ? c3 : c2) : ((c & 1u) != 0 ? c1 : c0);
            // every lane fetches (lanes past the list take triangle 0, which exists whenever a test
            // was listed), so no branch joins the loaded words and the first wait is at their use
            const float4* tr = bvh.tris + 3 * (size_t)(q < total ? first + k : 0u);
            const float4 a = tr[0], bb = tr[1], cc = tr[2];
            __builtin_amdgcn_sched_barrier(0); // the ray's permutes stay behind the fetch
            const V3 o = v3(fromLane(owner, s.o.x), fromLane(owner, s.o.y), fromLane(owner, s.o.z));
            const float olim = fromLane(owner, lim), otmax = fromLane(owner, s.tmax);
            const uint32_t oray = fromLane(owner, s.ray);
            RayPre rp;
            rp.sx = fromLane(owner, s.pre.sx);
            rp.sy = fromLane(owner, s.pre.sy);
            rp.sz = fromLane(owner, s.pre.sz);
            const uint32_t ax = fromLane(owner, s.axes);
            if (q < total) {
                rp.kx = (int)(ax & 3u);
                rp.ky = (int)((ax >> 2) & 3u);
                rp.kz = (int)((ax >> 4) & 3u);
                float t, b1, b2;
                if (intersect<false>(rp, o, xyz(a), xyz(bb), xyz(cc), olim, t, b1, b2) && t < otmax)
                    atomicMin(&sh.key[oray], ((unsigned long long)asu(t) << 32) | asu(a.w)); // t > 0: its bits order like the value
            }
            waveSync();
        }
    }
    if (live) {
        const bool nh = bestOf(sh, s.ray, bt);
        if (ANY && nh) {
            live = false;
        } else {
            // pop the next node that can still hold a closer hit; the LDS entry is read whatever
            // the depth (at a clamped index) and the overflow only past kShort, so the common read
            // stays an LDS read rather than a generic-address load of either array
            // (an empty asm that takes the values keeps both reads where they are, issued together:
            // sunk past the loop the node read became one load through a pointer to either array, a
            // generic-address load)
            auto top = [&](uint32_t& node, float& t) {
                s.sp--;
                const int li = s.sp < kShort ? s.sp : kShort - 1;
                node = sh.stackN[li][lane];
                uint32_t tb = 0;
                if (!ANY)
                    tb = (uint32_t)sh.stackT[li][lane];
                asm volatile("" : "+v"(node), "+v"(tb));
                t = asf(tb << 16);
                if (s.sp >= kShort) {
                    node = spillN[s.sp - kShort];
                    if (!ANY)
                        t = spillT[s.sp - kShort];
                }
            };
            live = false;
            if (s.sp > 0) {
                // the top entry, which most pops take; the loop only when it lies beyond the hit
                uint32_t node;
                float t;
                top(node, t);
                if (ANY || !(nh && t > bt)) {
                    s.node = node;
                    live = true;
                } else {
                    while (s.sp > 0) {
                        top(node, t);
                        if (nh && t > bt)
                            continue;
                        s.node = node;
                        live = true;
                        break;
                    }
                }
            }
        }
    }
    // work stealing: lanes without a node take the top entries (all in LDS) of lanes that have some
    const uint64_t idleM = __ballot(!live);
    const uint32_t offer = live && s.sp <= kShort ? ((uint32_t)s.sp < kOfferMax ? (uint32_t)s.sp : kOfferMax) : 0u;
    const uint64_t offM = __ballot(offer != 0u);
    const uint32_t nIdle = (uint32_t)__builtin_popcountll(idleM);
    if (offM != 0 && nIdle >= kStealMinIdle) {
        const uint32_t oincl = waveInclusiveSum(offer); // the branch is wave-uniform: every lane is here
        const uint32_t opre = oincl - offer;
        const uint32_t ototal = (uint32_t)__builtin_amdgcn_readlane((int)oincl, 63);
        // the offers in lane order go to the idle lanes in lane order
        const uint32_t take = opre >= nIdle ? 0u : (offer < nIdle - opre ? offer : nIdle - opre);
        for (uint32_t i = 0; i < take; i++)
            sh.list[opre + i] = (uint16_t)((lane << 4) | (uint32_t)(s.sp - 1 - (int)i));
        s.sp -= (int)take;
        waveSync();
        const uint32_t rank = lanesBelow(idleM);
        const bool gets = !live && rank < ototal;
        const uint32_t e = gets ? (uint32_t)sh.list[rank] : 0u;
        const uint32_t v = e >> 4, idx = e & 15u;
        // the donor's ray (every lane executes the permutes)
        const float vox = fromLane(v, s.o.x), voy = fromLane(v, s.o.y), voz = fromLane(v, s.o.z);
        const float vix = fromLane(v, s.ix), viy = fromLane(v, s.iy), viz = fromLane(v, s.iz);
        const float vsx = fromLane(v, s.pre.sx), vsy = fromLane(v, s.pre.sy), vsz = fromLane(v, s.pre.sz);
        const uint32_t vax = fromLane(v, s.axes), vray = fromLane(v, s.ray);
        const float vtmax = fromLane(v, s.tmax);
        if (gets) {
            const uint32_t node = sh.stackN[idx][v];
            float et = 0.0f;
            if (!ANY)
                et = asf((uint32_t)sh.stackT[idx][v] << 16);
            float vb;
            const bool vh = bestOf(sh, vray, vb);
            if (ANY ? !vh : (!vh || !(et > vb))) {
                s.o = v3(vox, voy, voz);
                s.ix = vix;
                s.iy = viy;
                s.iz = viz;
                s.ox = -vox * vix;
                s.oy = -voy * viy;
                s.oz = -voz * viz;
                s.near = (vix < 0.0f ? 1u : 0u) | (viy < 0.0f ? 2u : 0u) | (viz < 0.0f ? 4u : 0u);
                s.pre.sx = vsx;
                s.pre.sy = vsy;
                s.pre.sz = vsz;
                s.axes = vax;
                s.ray = vray;
                s.tmax = vtmax;
                s.node = node;
                s.sp = 0;
                live = true;
            }
        }
        waveSync();
    }
}

// One ray per lane, traced by the whole wave together; a ray's best (t, id) lives in sh.key while
// the wave's lanes work on it, and the owner looks up the winner's instance and primitive by its id
// at the end (one load per ray instead of a second barrier and an LDS write per triangle round).
template <bool ANY>
__device__ Hit traceWave(const Bvh& bvh, V3 o, V3 d, float tmax, bool active, WaveTris& sh)
{
    const uint32_t lane = threadIdx.x & 63u;
    TraceState s;
    traceBegin(s, o, d, tmax);
    s.ray = lane;
    bool live = active && bvh.ntri != 0;
    uint32_t spillN[kSpill];
    float spillT[kSpill];
    sh.key[lane] = ~0ull;
    waveSync();
    while (__ballot(live) != 0)
        waveStep<ANY>(bvh, s, live, sh, spillN, spillT);
    waveSync();
    Hit& h = s.h;
    const unsigned long long kk = sh.key[lane];
    if (active && kk != ~0ull) {
        h.hit = true;
        h.t = asf((uint32_t)(kk >> 32));
        h.gid = (uint32_t)kk;
        const uint2 w = bvh.ids[h.gid];
        h.inst = w.x;
        h.prim = w.y;
    }
    waveSync();
    if (active || s.visited != 0)
        count_traversal<ANY>(bvh, active ? 1u : 0u, s.visited, s.tested, active && h.hit);
    return s.h;
}

template <bool ANY>
__device__ Hit trace(const Bvh& bvh, V3 o, V3 d, float tmax)
{
    TraceState s;
    traceBegin(s, o, d, tmax);
    if (bvh.ntri == 0)
        return s.h;
    uint32_t stackN[kStack];
    float stackT[kStack];
    while (!traceStep<ANY>(bvh, s, stackN, stackT)) {
    }
    count_traversal<ANY>(bvh, 1u, s.visited, s.tested, s.h.hit);
    return s.h;
}

} // namespace vh
