/*
 * pt_oracle.c - ORACLE: CPU restatement of Verion's path tracer. TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this (through
 * oracle/__init__.py); the product library never links it and has no CPU fallback.
 *
 * It restates, function by function, shaders/include/pt_core.slang (bindings :12-49, pcg/Rng
 * :69-84, BSDF :115-557, occluded :580-613, generatePathFor :645-682, traceClosest :684-700,
 * ReSTIR DI :702-900, ReSTIR GI :901-980 and :1566-1617, shadePath :992-1543, pixelActive
 * :1551-1559, finishPath :1620-1667), the
 * megakernel loop of shaders/pathtrace.slang:9-25 and the primary-hit AOVs of shaders/aov.slang:84-188
 * and shaders/normals.slang:39-110, with these stated deviations (shared with the
 * GPU library):
 *   - film and adaptive statistics accumulate in FP64 (the reference: FP32);
 *   - the driver's ray queries become a watertight ray/triangle test (Woop, Benthin and Wald,
 *     "Watertight Ray/Triangle Intersection", JCGT 2013) over world-space triangles, closest hit
 *     = smallest t, ties to the smallest (instance, primitive); shadow rays in scenes with a
 *     translucent material take the closest hit per layer (the reference's first-found hit is
 *     implementation-defined);
 *   - transcendental functions come from or_math.h (Cephes), division and sqrt are IEEE;
 *   - ReSTIR GI's own-candidate draw hashes in the render's seed mix (zero for seed 0).
 * Every float literal carries an f suffix: float arithmetic as in the shader, compiled with
 * -ffp-contract=off and -Werror=double-promotion.
 *
 * Parity pins: tests/test_oracle_*.py (closed forms of the reference's own GPU tests, the albedo
 * tables, brute force vs BVH, golden PCG vectors, ReSTIR GI's solid-angle Jacobian, first-frame
 * identity and convergence); tests/test_oracle_reference_pins.py (the reference's assertions on
 * environment maps and importance sampling, dispersion, the thin lens, vertex normals and normal
 * maps, area-light form factors, textures and mips, light slots and the virtual main light, and an
 * isolating scene for ReSTIR GI's first-bounce emission split, pt_core.slang:1162-1168, whose
 * removal a planted copy shows the test catches); tests/test_ies.py (IES profiles against
 * per-sample closed forms); tests/test_oracle_units.py (the exact-sign difference of products of
 * the watertight test against rational arithmetic). Parity unpinned: none.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "or_math.h"

/* ---------------------------------------------------------------- tables (vh_scene_tables layout) */

typedef struct {
    uint32_t light_count, emitter_count, env_width, env_height;
    float env_pdf_scale;
    float sky[3];
    uint32_t virtual_light, transparent_background;
} or_info;

typedef struct {
    float camera[16];
    float aperture, focus;
    float shift[2];
} or_view;

typedef struct {
    const float* positions; uint64_t positions_len;
    const uint32_t* indices; uint64_t indices_len;
    const float* uvs; uint64_t uvs_len;
    const float* vnormals; uint64_t vnormals_len;
    const uint32_t* instances; uint64_t instances_len;
    const float* materials; uint64_t materials_len;
    const uint32_t* lights; uint64_t lights_len;
    const float* emitters; uint64_t emitters_len;
    const float* emit_cdf; uint64_t emit_cdf_len;
    const float* env_pixels; uint64_t env_pixels_len;
    const float* env_marginal; uint64_t env_marginal_len;
    const float* env_cond; uint64_t env_cond_len;
    const float* texels; uint64_t texels_len;
    const uint32_t* tex_info; uint64_t tex_info_len;
    const uint32_t* meshes; uint64_t mesh_count;
    const uint32_t* instance_mesh; uint64_t instance_count;
    or_info info;
    or_view view;
} or_tables;

/* ---------------------------------------------------------------- vectors */

typedef struct { float x, y, z; } f3;

static inline f3 v3(float x, float y, float z) { f3 r = {x, y, z}; return r; }
static inline f3 vadd(f3 a, f3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline f3 vsub(f3 a, f3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline f3 vmul(f3 a, f3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline f3 vscale(f3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline f3 sscale(float s, f3 a) { return v3(s * a.x, s * a.y, s * a.z); }
static inline f3 vdivs(f3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
static inline f3 vneg(f3 a) { return v3(-a.x, -a.y, -a.z); }
static inline float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline f3 cross3(f3 a, f3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline float len3(f3 a) { return sqrtf(dot3(a, a)); }
static inline f3 normalize3(f3 a) { float l = sqrtf(dot3(a, a)); return v3(a.x / l, a.y / l, a.z / l); }
static inline float fmax_(float a, float b) { return fmaxf(a, b); }
static inline float fmin_(float a, float b) { return fminf(a, b); }
static inline float clampf_(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
static inline float saturate_(float x) { return clampf_(x, 0.0f, 1.0f); }
static inline float lerpf_(float a, float b, float t) { return a + t * (b - a); }
static inline f3 lerp3(f3 a, f3 b, float t) { return v3(lerpf_(a.x, b.x, t), lerpf_(a.y, b.y, t), lerpf_(a.z, b.z, t)); }
static inline float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
static inline int bad3(f3 a) { return isnan(a.x) || isnan(a.y) || isnan(a.z) || isinf(a.x) || isinf(a.y) || isinf(a.z); }
/* float -> uint as the GPU converts: negative and NaN to 0, saturating at the top. */
static inline uint32_t f2u(float x) { return x > 0.0f ? (x >= 4294967040.0f ? 0xffffffffu : (uint32_t)x) : 0u; }
static inline float asf(uint32_t u) { return or_b2f(u); }
static inline uint32_t asu(float f) { return or_f2b(f); }

/* ---------------------------------------------------------------- scene, world triangles, BVH */

typedef struct {
    float lo[3], hi[3];
    uint32_t left, right; /* internal: child node ids; leaf: right = 0xffffffff, left = first, */
    uint32_t first, count;
} or_node;

typedef struct or_scene {
    or_tables t;
    uint32_t ntri;
    float* wtri;        /* 9 floats per triangle: world-space p0, p1, p2 */
    uint32_t* tri_inst; /* instance of each world triangle */
    uint32_t* tri_prim; /* primitive index within its instance's mesh */
    or_node* nodes;
    uint32_t nnodes;
    uint32_t* order;    /* triangle ids in leaf order */
    int translucent;    /* some instance's material has opacity < 1 or an opacity texture */
} or_scene;

/* transformPoint (pt_core.slang:90-99) */
static f3 transform_point(const uint32_t* ins, uint32_t base, f3 p)
{
    float r[3];
    for (uint32_t row = 0; row < 3; row++) {
        uint32_t o = base + 4 * row;
        r[row] = asf(ins[o]) * p.x + asf(ins[o + 1]) * p.y + asf(ins[o + 2]) * p.z + asf(ins[o + 3]);
    }
    return v3(r[0], r[1], r[2]);
}

/* Woop/Benthin/Wald watertight test, no culling. Returns 1 and t, b1, b2 (weights of vertices 1
 * and 2) for 0 < t < tmax. */
/* a*b - c*d within 2 ulp (Cornea, Harrison and Tang 2003): the products split exactly by fmaf, so
 * zero stays zero, the sign is right and swapping the products negates the result exactly */
static float diff_prod(float a, float b, float c, float d)
{
    float p1 = a * b, p2 = c * d;
    float e1 = fmaf(a, b, -p1), e2 = fmaf(c, d, -p2);
    return (p1 - p2) + (e1 - e2);
}

static int intersect_tri(const float* v, f3 o, f3 d, float tmax, float* tout, float* b1, float* b2)
{
    float dv[3] = {d.x, d.y, d.z};
    float ov[3] = {o.x, o.y, o.z};
    float ad[3] = {fabsf(d.x), fabsf(d.y), fabsf(d.z)};
    int kz = ad[0] > ad[1] ? (ad[0] > ad[2] ? 0 : 2) : (ad[1] > ad[2] ? 1 : 2);
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    if (dv[kz] < 0.0f) {
        int tmp = kx;
        kx = ky;
        ky = tmp;
    }
    float sx = dv[kx] / dv[kz];
    float sy = dv[ky] / dv[kz];
    float sz = 1.0f / dv[kz];
    float A[3], B[3], C[3];
    for (int k = 0; k < 3; k++) {
        A[k] = v[k] - ov[k];
        B[k] = v[3 + k] - ov[k];
        C[k] = v[6 + k] - ov[k];
    }
    float ax = A[kx] - sx * A[kz], ay = A[ky] - sy * A[kz];
    float bx = B[kx] - sx * B[kz], by = B[ky] - sy * B[kz];
    float cx = C[kx] - sx * C[kz], cy = C[ky] - sy * C[kz];
    float U = cx * by - cy * bx;
    float V = ax * cy - ay * cx;
    float W = bx * ay - by * ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) { /* decide an edge through the ray exactly */
        U = diff_prod(cx, by, cy, bx);
        V = diff_prod(ax, cy, ay, cx);
        W = diff_prod(bx, ay, by, ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f))
        return 0;
    float det = U + V + W;
    if (det == 0.0f)
        return 0;
    float az = sz * A[kz], bz = sz * B[kz], cz = sz * C[kz];
    float T = U * az + V * bz + W * cz;
    float t = T / det;
    if (!(t > 0.0f && t < tmax))
        return 0;
    *tout = t;
    *b1 = V / det;
    *b2 = W / det;
    return 1;
}

static void box_of(const float* v, float lo[3], float hi[3])
{
    for (int k = 0; k < 3; k++) {
        lo[k] = fminf(v[k], fminf(v[3 + k], v[6 + k]));
        hi[k] = fmaxf(v[k], fmaxf(v[3 + k], v[6 + k]));
    }
}

typedef struct {
    or_scene* s;
    float* cent;
    uint32_t next;
    float pad;
} bvh_build;

static int cmp_axis;
static const float* cmp_cent;
static int cmp_tri(const void* a, const void* b)
{
    uint32_t ia = *(const uint32_t*)a, ib = *(const uint32_t*)b;
    float ca = cmp_cent[3 * ia + cmp_axis], cb = cmp_cent[3 * ib + cmp_axis];
    if (ca < cb) return -1;
    if (ca > cb) return 1;
    return ia < ib ? -1 : (ia > ib ? 1 : 0);
}

static uint32_t build_node(bvh_build* b, uint32_t first, uint32_t count)
{
    or_scene* s = b->s;
    uint32_t id = b->next++;
    or_node* n = &s->nodes[id];
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = first; i < first + count; i++) {
        float tl[3], th[3];
        uint32_t tri = s->order[i];
        box_of(s->wtri + 9 * (size_t)tri, tl, th);
        for (int k = 0; k < 3; k++) {
            lo[k] = fminf(lo[k], tl[k]);
            hi[k] = fmaxf(hi[k], th[k]);
            clo[k] = fminf(clo[k], b->cent[3 * tri + k]);
            chi[k] = fmaxf(chi[k], b->cent[3 * tri + k]);
        }
    }
    for (int k = 0; k < 3; k++) {
        s->nodes[id].lo[k] = lo[k] - b->pad;
        s->nodes[id].hi[k] = hi[k] + b->pad;
    }
    if (count <= 4) {
        n->right = 0xffffffffu;
        n->first = first;
        n->count = count;
        n->left = 0;
        return id;
    }
    int axis = 0;
    float ext = chi[0] - clo[0];
    for (int k = 1; k < 3; k++)
        if (chi[k] - clo[k] > ext) {
            ext = chi[k] - clo[k];
            axis = k;
        }
    cmp_axis = axis;
    cmp_cent = b->cent;
    qsort(s->order + first, count, sizeof(uint32_t), cmp_tri);
    uint32_t half = count / 2;
    uint32_t l = build_node(b, first, half);
    uint32_t r = build_node(b, first + half, count - half);
    n = &s->nodes[id];
    n->left = l;
    n->right = r;
    n->first = 0;
    n->count = 0;
    return id;
}

or_scene* or_scene_create(const or_tables* t)
{
    or_scene* s = (or_scene*)calloc(1, sizeof(or_scene));
    s->t = *t;
    uint32_t ninst = (uint32_t)t->instance_count;
    uint64_t ntri = 0;
    for (uint32_t i = 0; i < ninst; i++)
        ntri += t->meshes[4 * t->instance_mesh[i] + 3] / 3;
    s->ntri = (uint32_t)ntri;
    s->wtri = (float*)malloc(sizeof(float) * 9 * (ntri ? ntri : 1));
    s->tri_inst = (uint32_t*)malloc(sizeof(uint32_t) * (ntri ? ntri : 1));
    s->tri_prim = (uint32_t*)malloc(sizeof(uint32_t) * (ntri ? ntri : 1));
    uint64_t g = 0;
    float scale = 0.0f;
    for (uint32_t i = 0; i < ninst; i++) {
        uint32_t mesh = t->instance_mesh[i];
        uint32_t ntm = t->meshes[4 * mesh + 3] / 3;
        uint32_t inst = i * 16;
        uint32_t vo = t->instances[inst];
        for (uint32_t p = 0; p < ntm; p++, g++) {
            uint32_t tri = t->instances[inst + 1] + 3 * p;
            for (int k = 0; k < 3; k++) {
                uint32_t vi = t->indices[tri + k] + vo;
                f3 pos = v3(t->positions[3 * vi], t->positions[3 * vi + 1], t->positions[3 * vi + 2]);
                f3 w = transform_point(t->instances, inst + 4, pos);
                s->wtri[9 * g + 3 * k] = w.x;
                s->wtri[9 * g + 3 * k + 1] = w.y;
                s->wtri[9 * g + 3 * k + 2] = w.z;
                scale = fmaxf(scale, fmaxf(fabsf(w.x), fmaxf(fabsf(w.y), fabsf(w.z))));
            }
            s->tri_inst[g] = i;
            s->tri_prim[g] = p;
        }
        uint32_t mi = t->instances[inst + 2] * 24;
        if (t->materials[mi + 22] < 1.0f || t->materials[mi + 23] > 0.0f)
            s->translucent = 1;
    }
    /* Median-split BVH over the world triangles; boxes padded by a relative epsilon so the slab
     * test never culls a triangle the exact test would hit. */
    s->order = (uint32_t*)malloc(sizeof(uint32_t) * (ntri ? ntri : 1));
    float* cent = (float*)malloc(sizeof(float) * 3 * (ntri ? ntri : 1));
    for (uint32_t i = 0; i < ntri; i++) {
        s->order[i] = i;
        for (int k = 0; k < 3; k++)
            cent[3 * i + k] = (s->wtri[9 * i + k] + s->wtri[9 * i + 3 + k] + s->wtri[9 * i + 6 + k]) / 3.0f;
    }
    s->nodes = (or_node*)calloc(2 * (ntri ? ntri : 1), sizeof(or_node));
    bvh_build b = {s, cent, 0, scale * 1.0e-5f + 1.0e-6f};
    if (ntri)
        build_node(&b, 0, (uint32_t)ntri);
    s->nnodes = b.next;
    free(cent);
    return s;
}

void or_scene_destroy(or_scene* s)
{
    if (!s)
        return;
    free(s->wtri);
    free(s->tri_inst);
    free(s->tri_prim);
    free(s->nodes);
    free(s->order);
    free(s);
}

uint32_t or_scene_triangles(const or_scene* s) { return s->ntri; }
const float* or_scene_world(const or_scene* s) { return s->wtri; }

typedef struct {
    int hit;
    uint32_t inst, prim;
    float t;
    float b1, b2;
    uint32_t gid;
} or_hit;

static inline int better(float t, uint32_t g, const or_hit* h)
{
    return !h->hit || t < h->t || (t == h->t && g < h->gid);
}

static void test_tri(const or_scene* s, uint32_t g, f3 o, f3 d, float tmax, or_hit* h)
{
    float t, b1, b2;
    float lim = h->hit ? h->t : tmax;
    /* ties are kept for the id comparison: test against the next float above the current best */
    if (h->hit)
        lim = nextafterf(lim, INFINITY);
    if (intersect_tri(s->wtri + 9 * (size_t)g, o, d, lim, &t, &b1, &b2) && t < tmax && better(t, g, h)) {
        h->hit = 1;
        h->t = t;
        h->b1 = b1;
        h->b2 = b2;
        h->gid = g;
        h->inst = s->tri_inst[g];
        h->prim = s->tri_prim[g];
    }
}

static int slab(const or_node* n, f3 o, f3 inv, float tmax, float* tnear)
{
    float ov[3] = {o.x, o.y, o.z}, iv[3] = {inv.x, inv.y, inv.z};
    float t0 = 0.0f, t1 = tmax;
    for (int k = 0; k < 3; k++) {
        float a = (n->lo[k] - ov[k]) * iv[k];
        float b = (n->hi[k] - ov[k]) * iv[k];
        float lo = fminf(a, b), hi = fmaxf(a, b);
        t0 = fmaxf(t0, lo);
        t1 = fminf(t1, hi * 1.0000004f);
    }
    *tnear = t0;
    return t0 <= t1;
}

static inline float safe_inv(float d) { return 1.0f / (fabsf(d) > 1.0e-30f ? d : (d < 0.0f ? -1.0e-30f : 1.0e-30f)); }

/* Closest hit (any = 0) or first found hit (any = 1) over (0, tmax). */
static or_hit trace_bvh(const or_scene* s, f3 o, f3 d, float tmax, int any)
{
    or_hit h;
    memset(&h, 0, sizeof(h));
    if (!s->ntri)
        return h;
    f3 inv = v3(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const or_node* n = &s->nodes[stack[--sp]];
        float tn;
        if (!slab(n, o, inv, tmax, &tn))
            continue;
        if (h.hit && tn > h.t)
            continue;
        if (n->right == 0xffffffffu) {
            for (uint32_t i = 0; i < n->count; i++) {
                test_tri(s, s->order[n->first + i], o, d, tmax, &h);
                if (any && h.hit)
                    return h;
            }
        } else {
            stack[sp++] = n->right;
            stack[sp++] = n->left;
        }
    }
    return h;
}

static or_hit trace_brute(const or_scene* s, f3 o, f3 d, float tmax)
{
    or_hit h;
    memset(&h, 0, sizeof(h));
    for (uint32_t g = 0; g < s->ntri; g++)
        test_tri(s, g, o, d, tmax, &h);
    return h;
}

/* ---------------------------------------------------------------- render context */

typedef struct {
    const or_scene* s;
    const or_tables* t;
    const uint32_t* P;
    double* film;
    double* stats;
    float* lighting;
    float* restir;
    uint32_t seed_mix;
    int brute;
} ctx_t;

#define PF(i) asf(c->P[i])

/* pcg, Rng (pt_core.slang:69-84) */
uint32_t or_pcg(uint32_t v)
{
    uint32_t s = v * 747796405u + 2891336453u;
    uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}

static inline float rng_next(uint32_t* st)
{
    *st = or_pcg(*st);
    return (float)(*st >> 8) * (1.0f / 16777216.0f);
}

static inline f3 vertex_pos(const ctx_t* c, uint32_t v) { return v3(c->t->positions[3 * v], c->t->positions[3 * v + 1], c->t->positions[3 * v + 2]); }
static inline f3 vertex_normal(const ctx_t* c, uint32_t v) { return v3(c->t->vnormals[3 * v], c->t->vnormals[3 * v + 1], c->t->vnormals[3 * v + 2]); }

/* transformNormal (pt_core.slang:103-111) */
static f3 transform_normal(const ctx_t* c, uint32_t base, f3 n)
{
    const uint32_t* I = c->t->instances;
    f3 r0 = v3(asf(I[base]), asf(I[base + 1]), asf(I[base + 2]));
    f3 r1 = v3(asf(I[base + 4]), asf(I[base + 5]), asf(I[base + 6]));
    f3 r2 = v3(asf(I[base + 8]), asf(I[base + 9]), asf(I[base + 10]));
    f3 c0 = cross3(r1, r2);
    float sg = dot3(r0, c0) < 0.0f ? -1.0f : 1.0f;
    return sscale(sg, v3(dot3(c0, n), dot3(cross3(r2, r0), n), dot3(cross3(r0, r1), n)));
}

/* BSDF pieces (pt_core.slang:115-177) */
static f3 cosine_local(float u1, float u2)
{
    float r = sqrtf(u1);
    float phi = 2.0f * OR_PI * u2;
    return v3(r * or_cosf(phi), r * or_sinf(phi), sqrtf(fmax_(0.0f, 1.0f - u1)));
}

float or_G1(float cosT, float alpha)
{
    float c2 = cosT * cosT;
    float tan2 = (1.0f - c2) / fmax_(c2, 1.0e-8f);
    return 2.0f / (1.0f + sqrtf(1.0f + alpha * alpha * tan2));
}

float or_D(float hz, float alpha)
{
    float a2 = alpha * alpha;
    float d = hz * hz * (a2 - 1.0f) + 1.0f;
    return a2 / (OR_PI * d * d);
}

static f3 schlick3(f3 f0, float cs)
{
    float m = 1.0f - saturate_(cs);
    float m2 = m * m;
    float k = m2 * m2 * m;
    return v3(f0.x + (1.0f - f0.x) * k, f0.y + (1.0f - f0.y) * k, f0.z + (1.0f - f0.z) * k);
}

float or_fresnel(float cosi, float eta)
{
    float s2 = (1.0f - cosi * cosi) / (eta * eta);
    if (s2 >= 1.0f)
        return 1.0f;
    float cost = sqrtf(1.0f - s2);
    float rs = (cosi - eta * cost) / (cosi + eta * cost);
    float rp = (eta * cosi - cost) / (eta * cosi + cost);
    return 0.5f * (rs * rs + rp * rp);
}

static void basis(f3 n, f3* t, f3* b)
{
    f3 a = fabsf(n.x) > 0.9f ? v3(0.0f, 1.0f, 0.0f) : v3(1.0f, 0.0f, 0.0f);
    *b = normalize3(cross3(n, a));
    *t = cross3(*b, n);
}

static f3 sample_vndf(f3 wo, float alpha, float u1, float u2)
{
    f3 vh = normalize3(v3(alpha * wo.x, alpha * wo.y, wo.z));
    float lensq = vh.x * vh.x + vh.y * vh.y;
    f3 t1 = lensq > 0.0f ? vscale(v3(-vh.y, vh.x, 0.0f), 1.0f / sqrtf(lensq)) : v3(1.0f, 0.0f, 0.0f);
    f3 t2 = cross3(vh, t1);
    float r = sqrtf(u1);
    float phi = 2.0f * OR_PI * u2;
    float a = r * or_cosf(phi);
    float b = r * or_sinf(phi);
    float s = 0.5f * (1.0f + vh.z);
    b = (1.0f - s) * sqrtf(1.0f - a * a) + s * b;
    f3 nh = vadd(vadd(sscale(a, t1), sscale(b, t2)), sscale(sqrtf(fmax_(0.0f, 1.0f - a * a - b * b)), vh));
    return normalize3(v3(alpha * nh.x, alpha * nh.y, fmax_(0.0f, nh.z)));
}

void or_sample_vndf(const float* wo, float alpha, float u1, float u2, float* out)
{
    f3 h = sample_vndf(v3(wo[0], wo[1], wo[2]), alpha, u1, u2);
    out[0] = h.x;
    out[1] = h.y;
    out[2] = h.z;
}

static inline float lum(f3 c) { return dot3(c, v3(0.2126f, 0.7152f, 0.0722f)); }

#define GGX_N 32u
#define SHEEN_N 16u
#define DIEL_N 16u
#define DIEL_NIOR 9u
#define ALBEDO_GGX 0u
#define ALBEDO_SHEEN (GGX_N * GGX_N)
#define ALBEDO_DIEL (ALBEDO_SHEEN + SHEEN_N * SHEEN_N)
#define ALBEDO_WORDS (ALBEDO_DIEL + DIEL_N * DIEL_N * DIEL_NIOR * 2u)

static inline float albedo_at(const ctx_t* c, uint32_t i) { return asf(c->t->lights[c->P[5] * 16 + i]); }
static inline uint32_t recv_mask(const ctx_t* c, uint32_t inst) { return c->t->lights[c->P[5] * 16 + ALBEDO_WORDS + 2 * inst]; }
static inline uint32_t emit_mask(const ctx_t* c, uint32_t inst) { return c->t->lights[c->P[5] * 16 + ALBEDO_WORDS + 2 * inst + 1]; }

/* iesFactor (pt_core.slang:199-216) */
static float ies_factor(const ctx_t* c, uint32_t po, f3 d)
{
    const uint32_t* L = c->t->lights;
    f3 l = v3(dot3(d, v3(asf(L[po]), asf(L[po + 1]), asf(L[po + 2]))),
              dot3(d, v3(asf(L[po + 3]), asf(L[po + 4]), asf(L[po + 5]))),
              dot3(d, v3(asf(L[po + 6]), asf(L[po + 7]), asf(L[po + 8]))));
    float th = or_acosf(clampf_(-l.z, -1.0f, 1.0f)) * (180.0f / OR_PI);
    float ph = or_atan2f(l.y, l.x) * (180.0f / OR_PI);
    if (ph < 0.0f)
        ph += 360.0f;
    uint32_t i = f2u(th) < 179u ? f2u(th) : 179u;
    uint32_t j = f2u(ph) < 359u ? f2u(ph) : 359u;
    float ft = th - (float)i;
    float fp = ph - (float)j;
    uint32_t t = po + 9;
    float a0 = lerpf_(asf(L[t + j * 181 + i]), asf(L[t + j * 181 + i + 1]), ft);
    float a1 = lerpf_(asf(L[t + (j + 1) * 181 + i]), asf(L[t + (j + 1) * 181 + i + 1]), ft);
    return lerpf_(a0, a1, fp);
}

/* addSlot (pt_core.slang:220-239) */
static void add_slot(const ctx_t* c, uint32_t pixel, uint32_t slot, f3 u, f3 s, f3 wi, float lit, float r)
{
    if (bad3(u) || bad3(s))
        return;
    float* L = c->lighting + (size_t)65 * pixel + 17 + 12 * slot;
    float w = lum(u);
    float wd = w + lum(s);
    L[0] += u.x;
    L[1] += u.y;
    L[2] += u.z;
    L[3] += w;
    L[4] += w * r;
    L[5] += s.x;
    L[6] += s.y;
    L[7] += s.z;
    L[8] += wd * wi.x;
    L[9] += wd * wi.y;
    L[10] += wd * wi.z;
    L[11] += wd * lit;
}

/* environment map (pt_core.slang:241-306) */
static f3 env_texel(const ctx_t* c, uint32_t i, uint32_t j)
{
    uint32_t o = 3 * (j * c->P[24] + i);
    return v3(c->t->env_pixels[o], c->t->env_pixels[o + 1], c->t->env_pixels[o + 2]);
}

static void env_coords(const ctx_t* c, f3 d, uint32_t* i, uint32_t* j)
{
    uint32_t w = c->P[24], h = c->P[25];
    float u = or_atan2f(d.x, -d.z) * (0.5f / OR_PI) + 0.5f;
    float v = or_acosf(clampf_(d.y, -1.0f, 1.0f)) / OR_PI;
    uint32_t iu = f2u(u * (float)w), jv = f2u(v * (float)h);
    *i = iu < w - 1 ? iu : w - 1;
    *j = jv < h - 1 ? jv : h - 1;
}

static f3 env_radiance(const ctx_t* c, f3 d)
{
    uint32_t i, j;
    env_coords(c, d, &i, &j);
    return env_texel(c, i, j);
}

static float env_pdf(const ctx_t* c, f3 d)
{
    uint32_t i, j;
    env_coords(c, d, &i, &j);
    return lum(env_texel(c, i, j)) * PF(26);
}

static f3 env_sample(const ctx_t* c, float u1, float u2, float u3, float u4, float* pdf, f3* le)
{
    uint32_t w = c->P[24], h = c->P[25];
    uint32_t lo = 0, hi = h - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2;
        if (c->t->env_marginal[mid] >= u1)
            hi = mid;
        else
            lo = mid + 1;
    }
    uint32_t j = lo;
    lo = 0;
    hi = w - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2;
        if (c->t->env_cond[j * w + mid] >= u2)
            hi = mid;
        else
            lo = mid + 1;
    }
    uint32_t i = lo;
    float phi = (((float)i + u3) / (float)w - 0.5f) * 2.0f * OR_PI;
    float theta = (((float)j + u4) / (float)h) * OR_PI;
    float st = or_sinf(theta);
    *le = env_texel(c, i, j);
    *pdf = lum(*le) * PF(26);
    return v3(st * or_sinf(phi), or_cosf(theta), -st * or_cosf(phi));
}

/* textures (pt_core.slang:309-356) */
static f3 texel_at(const ctx_t* c, uint32_t base, uint32_t w, uint32_t h, int x, int y)
{
    uint32_t xi = (uint32_t)(((x % (int)w) + (int)w) % (int)w);
    uint32_t yi = (uint32_t)(((y % (int)h) + (int)h) % (int)h);
    uint32_t o = 3 * (base + yi * w + xi);
    return v3(c->t->texels[o], c->t->texels[o + 1], c->t->texels[o + 2]);
}

static f3 bilinear(const ctx_t* c, uint32_t base, uint32_t w, uint32_t h, float ux, float uy)
{
    float fx = ux * (float)w - 0.5f;
    float fy = uy * (float)h - 0.5f;
    int x0 = (int)floorf(fx);
    int y0 = (int)floorf(fy);
    float tx = fx - (float)x0;
    float ty = fy - (float)y0;
    f3 a = lerp3(texel_at(c, base, w, h, x0, y0), texel_at(c, base, w, h, x0 + 1, y0), tx);
    f3 b = lerp3(texel_at(c, base, w, h, x0, y0 + 1), texel_at(c, base, w, h, x0 + 1, y0 + 1), tx);
    return lerp3(a, b, ty);
}

static f3 sample_texture(const ctx_t* c, uint32_t tex, float ux, float uy, float footprint)
{
    uint32_t ti = tex * 4;
    uint32_t base = c->t->tex_info[ti], w = c->t->tex_info[ti + 1], h = c->t->tex_info[ti + 2], levels = c->t->tex_info[ti + 3];
    float lod = clampf_(or_log2f(fmax_(footprint * (float)(w > h ? w : h), 1.0e-8f)), 0.0f, (float)(levels - 1));
    uint32_t l0 = f2u(floorf(lod));
    float t = lod - (float)l0;
    uint32_t off = base, lw = w, lh = h;
    for (uint32_t k = 0; k < l0; k++) {
        off += lw * lh;
        lw = lw / 2 > 1u ? lw / 2 : 1u;
        lh = lh / 2 > 1u ? lh / 2 : 1u;
    }
    f3 a = bilinear(c, off, lw, lh, ux, uy);
    if (l0 + 1 >= levels || t <= 0.0f)
        return a;
    off += lw * lh;
    f3 b = bilinear(c, off, lw / 2 > 1u ? lw / 2 : 1u, lh / 2 > 1u ? lh / 2 : 1u, ux, uy);
    return lerp3(a, b, t);
}

typedef struct {
    f3 albedo;
    float metallic, alpha, transmission, ior, eta, fSel, tComp, coat, coatAlpha;
    f3 sheenColor;
    float sheenAlpha, sheenP, coatP;
} surface_t;

typedef struct {
    f3 diffuse, specular;
} lobes_t;

/* albedo-table lookups (pt_core.slang:384-445) */
static float ggx_albedo(const ctx_t* c, float cosO, float alpha)
{
    float x = saturate_(cosO) * (float)(GGX_N - 1);
    float y = saturate_(sqrtf(alpha)) * (float)(GGX_N - 1);
    uint32_t x0 = f2u(x) < GGX_N - 2 ? f2u(x) : GGX_N - 2;
    uint32_t y0 = f2u(y) < GGX_N - 2 ? f2u(y) : GGX_N - 2;
    float fx = x - (float)x0, fy = y - (float)y0;
    float a = lerpf_(albedo_at(c, ALBEDO_GGX + y0 * GGX_N + x0), albedo_at(c, ALBEDO_GGX + y0 * GGX_N + x0 + 1), fx);
    float b = lerpf_(albedo_at(c, ALBEDO_GGX + (y0 + 1) * GGX_N + x0), albedo_at(c, ALBEDO_GGX + (y0 + 1) * GGX_N + x0 + 1), fx);
    return lerpf_(a, b, fy);
}

static f3 multiscatter_factor(const ctx_t* c, const surface_t* s, float cosO)
{
    if (c->P[29] == 0)
        return v3(1.0f, 1.0f, 1.0f);
    float k = 1.0f / fmax_(ggx_albedo(c, cosO, s->alpha), 0.05f) - 1.0f;
    return v3(1.0f + s->albedo.x * k, 1.0f + s->albedo.y * k, 1.0f + s->albedo.z * k);
}

static float dielectric_albedo(const ctx_t* c, float cosO, float rough, float ior, int outside)
{
    float x = saturate_(cosO) * (float)(DIEL_N - 1);
    float y = saturate_(rough) * (float)(DIEL_N - 1);
    float z = clampf_((ior - 1.0f) * 4.0f, 0.0f, (float)(DIEL_NIOR - 1) - 1.0e-3f);
    uint32_t x0 = f2u(x) < DIEL_N - 2 ? f2u(x) : DIEL_N - 2;
    uint32_t y0 = f2u(y) < DIEL_N - 2 ? f2u(y) : DIEL_N - 2;
    uint32_t z0 = f2u(z);
    float fx = x - (float)x0, fy = y - (float)y0, fz = z - (float)z0;
    uint32_t base = (outside ? 0u : DIEL_NIOR) * DIEL_N * DIEL_N;
    float cc[2];
    for (uint32_t k = 0; k < 2; k++) {
        uint32_t o = base + ((z0 + k) * DIEL_N + y0) * DIEL_N + x0;
        float a = lerpf_(albedo_at(c, ALBEDO_DIEL + o), albedo_at(c, ALBEDO_DIEL + o + 1), fx);
        float b = lerpf_(albedo_at(c, ALBEDO_DIEL + o + DIEL_N), albedo_at(c, ALBEDO_DIEL + o + DIEL_N + 1), fx);
        cc[k] = lerpf_(a, b, fy);
    }
    return lerpf_(cc[0], cc[1], fz);
}

static float sheen_albedo(const ctx_t* c, float cosO, float r)
{
    float x = saturate_(cosO) * (float)(SHEEN_N - 1);
    float y = saturate_(r) * (float)(SHEEN_N - 1);
    uint32_t x0 = f2u(x) < SHEEN_N - 2 ? f2u(x) : SHEEN_N - 2;
    uint32_t y0 = f2u(y) < SHEEN_N - 2 ? f2u(y) : SHEEN_N - 2;
    float fx = x - (float)x0, fy = y - (float)y0;
    float a = lerpf_(albedo_at(c, ALBEDO_SHEEN + y0 * SHEEN_N + x0), albedo_at(c, ALBEDO_SHEEN + y0 * SHEEN_N + x0 + 1), fx);
    float b = lerpf_(albedo_at(c, ALBEDO_SHEEN + (y0 + 1) * SHEEN_N + x0), albedo_at(c, ALBEDO_SHEEN + (y0 + 1) * SHEEN_N + x0 + 1), fx);
    return lerpf_(a, b, fy);
}

/* sheenCore (pt_core.slang:449-455) */
float or_sheen_core(float woz, float wiz, float hz, float alpha)
{
    float sinH = sqrtf(fmax_(1.0f - hz * hz, 0.0f));
    float d = (2.0f + 1.0f / alpha) * or_powf(sinH, 1.0f / alpha) / (2.0f * OR_PI);
    float v = 1.0f / (4.0f * (wiz + woz - wiz * woz));
    return d * v * wiz;
}

/* evalLobes (pt_core.slang:459-486) */
static lobes_t eval_lobes(const ctx_t* c, const surface_t* s, f3 woL, f3 wiL)
{
    f3 hh = normalize3(vadd(woL, wiL));
    float woh = fmax_(dot3(woL, hh), 0.0f);
    float core = or_D(hh.z, s->alpha) * or_G1(woL.z, s->alpha) * or_G1(wiL.z, s->alpha) / (4.0f * woL.z);
    float t = s->transmission;
    lobes_t l;
    f3 sch = schlick3(s->albedo, woh);
    f3 ms = multiscatter_factor(c, s, woL.z);
    float diel = (1.0f - s->metallic) * (t * or_fresnel(woh, s->eta) * s->tComp + (1.0f - t) * or_fresnel(woh, s->ior)) * core;
    f3 m = vscale(vmul(sscale(s->metallic, sch), ms), core);
    l.specular = v3(m.x + diel, m.y + diel, m.z + diel);
    float kd = (1.0f - s->metallic) * (1.0f - t) * (1.0f - s->fSel);
    f3 dd = sscale(kd, s->albedo);
    l.diffuse = vscale(dd, wiL.z / OR_PI);
    if (s->sheenP > 0.0f) {
        float sc = or_sheen_core(woL.z, wiL.z, hh.z, s->sheenAlpha);
        l.specular = vadd(sscale(1.0f - s->sheenP, l.specular), vscale(s->sheenColor, sc));
        l.diffuse = vscale(l.diffuse, 1.0f - s->sheenP);
    }
    if (s->coat > 0.0f) {
        float coreC = or_D(hh.z, s->coatAlpha) * or_G1(woL.z, s->coatAlpha) * or_G1(wiL.z, s->coatAlpha) / (4.0f * woL.z);
        float cs = s->coat * or_fresnel(woh, 1.5f) * coreC;
        f3 k = sscale(1.0f - s->coatP, l.specular);
        l.specular = v3(k.x + cs, k.y + cs, k.z + cs);
        l.diffuse = vscale(l.diffuse, 1.0f - s->coatP);
    }
    return l;
}

static f3 eval_reflective(const ctx_t* c, const surface_t* s, f3 woL, f3 wiL)
{
    lobes_t l = eval_lobes(c, s, woL, wiL);
    return vadd(l.specular, l.diffuse);
}

/* Parts (pt_core.slang:496-536) */
typedef struct {
    f3 total, directDiffuse, indirectDiffuse, specular, transmission, emission;
} parts_t;

static void parts_file(parts_t* p, uint32_t depth, uint32_t cls, f3 c)
{
    if (depth == 0)
        p->emission = vadd(p->emission, c);
    else if (cls == 0)
        p->indirectDiffuse = vadd(p->indirectDiffuse, c);
    else if (cls == 1)
        p->specular = vadd(p->specular, c);
    else
        p->transmission = vadd(p->transmission, c);
}

static void parts_add(parts_t* p, uint32_t depth, uint32_t cls, f3 c)
{
    p->total = vadd(p->total, c);
    parts_file(p, depth, cls, c);
}

static void parts_add_light(parts_t* p, uint32_t depth, uint32_t cls, f3 c, f3 cd, f3 cs)
{
    p->total = vadd(p->total, c);
    if (depth == 0) {
        p->directDiffuse = vadd(p->directDiffuse, cd);
        p->specular = vadd(p->specular, cs);
    } else {
        parts_file(p, depth, cls, vadd(cd, cs));
    }
}

/* reflectivePdf (pt_core.slang:540-557) */
static float reflective_pdf(const surface_t* s, f3 woL, f3 wiL)
{
    f3 hh = normalize3(vadd(woL, wiL));
    float woh = fmax_(dot3(woL, hh), 0.0f);
    float spec = or_G1(woL.z, s->alpha) * or_D(hh.z, s->alpha) / (4.0f * woL.z);
    float t = s->transmission;
    float pSpec = s->metallic + (1.0f - s->metallic) * (t * or_fresnel(woh, s->eta) + (1.0f - t) * s->fSel);
    float pDiff = (1.0f - s->metallic) * (1.0f - t) * (1.0f - s->fSel) * wiL.z / OR_PI;
    float q = pSpec * spec + pDiff;
    if (s->sheenP > 0.0f)
        q = (1.0f - s->sheenP) * q + s->sheenP * wiL.z / OR_PI;
    if (s->coat > 0.0f) {
        float specC = or_G1(woL.z, s->coatAlpha) * or_D(hh.z, s->coatAlpha) / (4.0f * woL.z);
        q = (1.0f - s->coatP) * q + s->coatP * specC;
    }
    return q;
}

/* opacityAt (pt_core.slang:561-576) */
static float opacity_at(const ctx_t* c, uint32_t instIdx, uint32_t prim, float b1, float b2)
{
    const or_tables* t = c->t;
    uint32_t mi = t->instances[instIdx * 16 + 2] * 24;
    float op = t->materials[mi + 22];
    uint32_t tex = f2u(t->materials[mi + 23]);
    if (tex > 0 && op > 0.0f) {
        uint32_t tri = t->instances[instIdx * 16 + 1] + 3 * prim;
        uint32_t vo = t->instances[instIdx * 16];
        uint32_t i0 = t->indices[tri] + vo, i1 = t->indices[tri + 1] + vo, i2 = t->indices[tri + 2] + vo;
        float u0x = t->uvs[2 * i0], u0y = t->uvs[2 * i0 + 1];
        float u1x = t->uvs[2 * i1], u1y = t->uvs[2 * i1 + 1];
        float u2x = t->uvs[2 * i2], u2y = t->uvs[2 * i2 + 1];
        float ux = u0x + b1 * (u1x - u0x) + b2 * (u2x - u0x);
        float uy = u0y + b1 * (u1y - u0y) + b2 * (u2y - u0y);
        op *= sample_texture(c, tex - 1, ux, uy, 0.0f).x;
    }
    return op;
}

static or_hit trace(const ctx_t* c, f3 o, f3 d, float tmax, int any)
{
    if (c->brute)
        return trace_brute(c->s, o, d, tmax);
    return trace_bvh(c->s, o, d, tmax, any);
}

/* occluded (pt_core.slang:580-613) */
static int occluded(const ctx_t* c, f3 o, f3 d, float tmax)
{
    int any = !c->s->translucent;
    for (uint32_t layer = 0; layer < 8; layer++) {
        or_hit q = trace(c, o, d, tmax, any);
        if (!q.hit)
            return 0;
        float opacity = opacity_at(c, q.inst, q.prim, q.b1, q.b2);
        if (opacity >= 1.0f)
            return 1;
        uint32_t h = asu(o.x) * 73856093u ^ asu(o.y) * 19349663u ^ asu(o.z) * 83492791u ^ asu(d.x) * 2654435761u
            ^ asu(d.y) * 40503u ^ asu(d.z) * 2246822519u ^ layer * 3266489917u;
        h ^= h >> 16;
        h *= 0x7feb352du;
        h ^= h >> 15;
        h *= 0x846ca68bu;
        h ^= h >> 16;
        if ((float)(h >> 8) * (1.0f / 16777216.0f) < opacity)
            return 1;
        float adv = q.t + 1.0e-4f;
        o = vadd(o, vscale(d, adv));
        tmax -= adv;
        if (tmax <= 0.0f)
            return 0;
    }
    return 1;
}

/* PathState (pt_core.slang:616-633) */
typedef struct {
    f3 org, dir, thr;
    uint32_t rng, cls;
    int chan;
    float pathLen;
    int misPrev, gi, restir;
    float qPrev;
    uint32_t prevInst;
    float visLit, visTotal;
    parts_t parts;
} path_t;

/* generatePathFor (pt_core.slang:645-682), with the seed mix XORed into the pixel hash */
static path_t generate_path(const ctx_t* c, uint32_t px, uint32_t py, uint32_t sample)
{
    uint32_t w = c->P[0], h = c->P[1];
    uint32_t pixel = py * w + px;
    uint32_t rng = or_pcg(pixel ^ or_pcg(sample * 0x9E3779B9u + 0x7F4A7C15u) ^ c->seed_mix);
    float t = PF(2);
    float aspect = (float)w / (float)h;
    float u = (2.0f * ((float)px + rng_next(&rng)) / (float)w - 1.0f + PF(39)) * aspect * t;
    float v = (1.0f - 2.0f * ((float)py + rng_next(&rng)) / (float)h + PF(40)) * t;
    f3 focal = vscale(v3(u, v, -1.0f), PF(4));
    float lensR = PF(3) * sqrtf(rng_next(&rng));
    float lensA = 2.0f * OR_PI * rng_next(&rng);
    f3 lens = v3(lensR * or_cosf(lensA), lensR * or_sinf(lensA), 0.0f);
    f3 dc = normalize3(vsub(focal, lens));
    path_t s;
    memset(&s, 0, sizeof(s));
    float org[3], dir[3];
    for (uint32_t r = 0; r < 3; r++) {
        org[r] = PF(8 + 4 * r) * lens.x + PF(9 + 4 * r) * lens.y + PF(10 + 4 * r) * lens.z + PF(11 + 4 * r);
        dir[r] = PF(8 + 4 * r) * dc.x + PF(9 + 4 * r) * dc.y + PF(10 + 4 * r) * dc.z;
    }
    s.org = v3(org[0], org[1], org[2]);
    s.dir = normalize3(v3(dir[0], dir[1], dir[2]));
    s.thr = v3(1.0f, 1.0f, 1.0f);
    s.rng = rng;
    s.chan = -1;
    return s;
}

/* ---------------------------------------------------------------- ReSTIR DI (pt_core.slang:702-995) */

#define RESTIR_MIN_ALPHA 0.15f
#define RESTIR_RADIUS 6
#define RESTIR_NEIGHBOURS 2u
#define RESTIR_HALF 28u
#define RESTIR_STRIDE 80u

/* Random neighbours per reuse: bits 8-15 of the ReSTIR mode word params[36] (an extension of this
   build; 0 = the reference's fixed RESTIR_NEIGHBOURS = 2, pt_core.slang:706). */
static inline uint32_t restir_neighbours(const ctx_t* c)
{
    uint32_t k = (c->P[36] >> 8) & 0xffu;
    return k ? k : RESTIR_NEIGHBOURS;
}

static inline uint32_t restir_base(uint32_t pixel, uint32_t which) { return pixel * RESTIR_STRIDE + which * RESTIR_HALF; }

static int restir_light(const ctx_t* c, uint32_t e, uint32_t recv, f3 x, f3 p, f3 n, f3* wi, float* dist, float* cosi, f3* g, float* plsa)
{
    const float* E = c->t->emitters + (size_t)e * 16;
    f3 a = v3(E[0], E[1], E[2]), b = v3(E[3], E[4], E[5]), cc = v3(E[6], E[7], E[8]);
    f3 toL = vsub(x, p);
    float d2 = dot3(toL, toL);
    *dist = sqrtf(d2);
    *wi = vdivs(toL, fmax_(*dist, 1.0e-6f));
    float cosl = fabsf(dot3(normalize3(cross3(vsub(b, a), vsub(cc, a))), *wi));
    *cosi = dot3(n, *wi);
    *g = vscale(v3(E[9], E[10], E[11]), cosl / fmax_(d2, 1.0e-12f));
    *plsa = E[13] * d2 / (fmax_(E[12], 1.0e-30f) * fmax_(cosl, 1.0e-5f));
    return *cosi > 0.0f && cosl > 1.0e-5f && d2 > 1.0e-12f && (asu(E[14]) & recv) != 0;
}

static float restir_mis(const surface_t* surf, f3 woL, f3 wiL, float plsa)
{
    float qb = reflective_pdf(surf, woL, wiL);
    return plsa * plsa / fmax_(plsa * plsa + qb * qb, 1.0e-30f);
}

typedef struct {
    int valid;
    uint32_t e;
    f3 x;
    float W;
    int occ;
} restir_pick;

static restir_pick restir_di(const ctx_t* c, uint32_t pixel, uint32_t recv, f3 p, f3 n, float pHitT, const surface_t* surf, f3 T, f3 B, f3 woL, uint32_t* rng)
{
    uint32_t w = c->P[0];
    uint32_t cur = c->P[6] & 1u;
    uint32_t emitterCount = c->P[23];
    float cap = (float)(c->P[38] > 1u ? c->P[38] : 1u);
    float wsum = 0.0f, mTotal = 0.0f;
    uint32_t selE = 0;
    f3 selX = v3(0.0f, 0.0f, 0.0f);
    float selPhat = 0.0f;
    uint32_t ncand = c->P[37] > 1u ? c->P[37] : 1u;
    for (uint32_t k = 0; k < ncand; k++) {
        float ue = rng_next(rng), ua = rng_next(rng), ub = rng_next(rng);
        uint32_t lo = 0, hi = emitterCount - 1;
        while (lo < hi) {
            uint32_t mid = (lo + hi) / 2;
            if (c->t->emit_cdf[mid] >= ue)
                hi = mid;
            else
                lo = mid + 1;
        }
        const float* E = c->t->emitters + (size_t)lo * 16;
        f3 a = v3(E[0], E[1], E[2]), b = v3(E[3], E[4], E[5]), cc = v3(E[6], E[7], E[8]);
        float su = sqrtf(ua);
        f3 x = vadd(vadd(vscale(a, 1.0f - su), vscale(b, su * (1.0f - ub))), vscale(cc, su * ub));
        float area = E[12];
        float q = area > 0.0f ? E[13] / area : 0.0f;
        f3 wi, g;
        float dist, cosi, plsa;
        float phat = 0.0f;
        if (q > 0.0f && restir_light(c, lo, recv, x, p, n, &wi, &dist, &cosi, &g, &plsa)) {
            f3 wiL = v3(dot3(wi, T), dot3(wi, B), cosi);
            phat = lum(vmul(eval_reflective(c, surf, woL, wiL), g)) * restir_mis(surf, woL, wiL, plsa);
        }
        float wgt = phat / fmax_(q, 1.0e-30f);
        mTotal += 1.0f;
        if (phat > 0.0f) {
            wsum += wgt;
            if (rng_next(rng) * wsum < wgt) {
                selE = lo + 1;
                selX = x;
                selPhat = phat;
            }
        }
    }
    if (c->P[6] > 0) {
        int px = (int)(pixel % w), py = (int)(pixel / w);
        for (uint32_t k = 0; k < 1 + restir_neighbours(c); k++) {
            int nx = px, ny = py;
            if (k > 0) {
                nx += (int)(rng_next(rng) * (float)(2 * RESTIR_RADIUS + 1)) - RESTIR_RADIUS;
                ny += (int)(rng_next(rng) * (float)(2 * RESTIR_RADIUS + 1)) - RESTIR_RADIUS;
                if (nx < 0 || ny < 0 || nx >= (int)w || ny >= (int)c->P[1])
                    continue;
            }
            const float* R = c->restir + restir_base((uint32_t)ny * w + (uint32_t)nx, cur ^ 1u);
            uint32_t ei = asu(R[0]);
            if (asu(R[14]) == 0 || ei == 0)
                continue;
            f3 pn = v3(R[11], R[12], R[13]);
            f3 pp = v3(R[8], R[9], R[10]);
            if (dot3(pn, n) < 0.9f || len3(vsub(pp, p)) > 0.05f * fmax_(pHitT, 0.01f))
                continue;
            f3 xc = v3(R[1], R[2], R[3]);
            float mc = fmin_(R[5], cap);
            f3 wi, g;
            float dist, cosi, plsa;
            float phat = 0.0f;
            if (restir_light(c, ei - 1, recv, xc, p, n, &wi, &dist, &cosi, &g, &plsa)) {
                f3 wiL = v3(dot3(wi, T), dot3(wi, B), cosi);
                phat = lum(vmul(eval_reflective(c, surf, woL, wiL), g)) * restir_mis(surf, woL, wiL, plsa);
            }
            float wgt = phat * R[4] * mc;
            mTotal += mc;
            if (wgt > 0.0f) {
                wsum += wgt;
                if (rng_next(rng) * wsum < wgt) {
                    selE = ei;
                    selX = xc;
                    selPhat = phat;
                }
            }
        }
    }
    restir_pick r;
    r.valid = selE != 0 && selPhat > 0.0f;
    r.e = selE;
    r.x = selX;
    r.W = r.valid ? wsum / (mTotal * selPhat) : 0.0f;
    r.occ = 0;
    if (r.valid) {
        f3 wi, g;
        float dist, cosi, plsa;
        restir_light(c, selE - 1, recv, selX, p, n, &wi, &dist, &cosi, &g, &plsa);
        r.occ = occluded(c, p, wi, dist * (1.0f - 1.0e-3f));
    }
    float* C = c->restir + restir_base(pixel, cur);
    C[0] = asf(selE);
    C[1] = selX.x;
    C[2] = selX.y;
    C[3] = selX.z;
    C[4] = r.occ ? 0.0f : r.W;
    C[5] = mTotal;
    return r;
}

static void restir_gbuffer(const ctx_t* c, uint32_t pixel, f3 p, f3 n)
{
    float* C = c->restir + restir_base(pixel, c->P[6] & 1u);
    C[8] = p.x;
    C[9] = p.y;
    C[10] = p.z;
    C[11] = n.x;
    C[12] = n.y;
    C[13] = n.z;
    C[14] = asf(1u);
}

static void restir_clear(const ctx_t* c, uint32_t pixel)
{
    float* C = c->restir + restir_base(pixel, c->P[6] & 1u);
    C[14] = asf(0u);
    C[27] = asf(0u);
}

/* ---------------------------------------------------------------- ReSTIR GI (pt_core.slang:901-980, 1566-1617) */

#define RESTIR_MAX_JACOBIAN 10.0f

/* restirScratch (pt_core.slang:713): the per-pixel words after the two frame halves */
static inline uint32_t restir_scratch(uint32_t pixel) { return pixel * RESTIR_STRIDE + 2u * RESTIR_HALF; }

/* The reconnection test of restirGIPrepare (pt_core.slang:938-951): the direction wi and distance
   dR from the shading point p (normal n) to a neighbour's second hit x2 (normal n2), found from the
   neighbour's shading point pq, and the Jacobian of the shift, the ratio of the solid angles that
   an area element at x2 subtends at p and at pq: (cosR / cosQ) (dQ / dR)^2. Returns whether the
   sample may be moved: it faces p, is not too close, and the Jacobian is within [1/10, 10]. */
static int gi_reconnect(f3 p, f3 n, f3 pq, f3 x2, f3 n2, f3* wi, float* cosT, float* dR, float* jac)
{
    f3 toX = vsub(x2, p);
    *dR = len3(toX);
    *wi = vdivs(toX, fmax_(*dR, 1.0e-6f));
    float dQ = len3(vsub(x2, pq));
    *cosT = dot3(n, *wi);
    float cosR = fabsf(dot3(n2, *wi));
    float cosQ = fabsf(dot3(n2, normalize3(vsub(pq, x2))));
    *jac = 0.0f;
    if (*cosT <= 0.0f || *dR < 1.0e-3f || cosR < 1.0e-3f || cosQ < 1.0e-3f)
        return 0;
    *jac = (cosR / cosQ) * (dQ * dQ) / (*dR * *dR);
    return !(*jac > RESTIR_MAX_JACOBIAN || *jac < 1.0f / RESTIR_MAX_JACOBIAN);
}

/* gi_reconnect for tests: p, n, pq, x2, n2 are 3 floats each; writes the Jacobian (0 when a cosine
   or distance test fails) and returns the accept flag */
int or_gi_reconnect(const float* p, const float* n, const float* pq, const float* x2, const float* n2, float* jac)
{
    f3 wi;
    float cosT, dR;
    return gi_reconnect(v3(p[0], p[1], p[2]), v3(n[0], n[1], n[2]), v3(pq[0], pq[1], pq[2]), v3(x2[0], x2[1], x2[2]),
                        v3(n2[0], n2[1], n2[2]), &wi, &cosT, &dR, jac);
}

/* restirGIPrepare (pt_core.slang:901-972): the previous frame's second-hit reservoirs of this pixel
   and two random neighbours, reconnected to this shading point with the solid-angle Jacobian
   (cosR / cosQ) (dQ / dR)^2, kept within [1/10, 10]; the resampled pick goes to the scratch words */
static void restir_gi_prepare(const ctx_t* c, uint32_t pixel, f3 p, f3 n, float pHitT, const surface_t* surf, f3 T, f3 B, f3 woL, uint32_t* rng)
{
    uint32_t w = c->P[0];
    uint32_t cur = c->P[6] & 1u;
    float cap = (float)(c->P[38] > 1u ? c->P[38] : 1u);
    float wsum = 0.0f, mTotal = 0.0f, pSel = 0.0f;
    f3 cSel = v3(0.0f, 0.0f, 0.0f), xSel = cSel, nSel = cSel, lSel = cSel;
    if (c->P[6] > 0) {
        int px = (int)(pixel % w), py = (int)(pixel / w);
        for (uint32_t k = 0; k < 1 + restir_neighbours(c); k++) {
            int nx = px, ny = py;
            if (k > 0) {
                nx += (int)(rng_next(rng) * (float)(2 * RESTIR_RADIUS + 1)) - RESTIR_RADIUS;
                ny += (int)(rng_next(rng) * (float)(2 * RESTIR_RADIUS + 1)) - RESTIR_RADIUS;
                if (nx < 0 || ny < 0 || nx >= (int)w || ny >= (int)c->P[1])
                    continue;
            }
            const float* R = c->restir + restir_base((uint32_t)ny * w + (uint32_t)nx, cur ^ 1u);
            if (asu(R[14]) == 0 || asu(R[27]) == 0)
                continue;
            f3 pq = v3(R[8], R[9], R[10]);
            f3 nq = v3(R[11], R[12], R[13]);
            if (dot3(nq, n) < 0.9f || len3(vsub(pq, p)) > 0.05f * fmax_(pHitT, 0.01f))
                continue;
            f3 x2 = v3(R[16], R[17], R[18]);
            f3 n2 = v3(R[19], R[20], R[21]);
            f3 l2 = v3(R[22], R[23], R[24]);
            float mq = fmin_(R[26], cap);
            mTotal += mq;
            f3 wi;
            float cosT, dR, jac;
            if (!gi_reconnect(p, n, pq, x2, n2, &wi, &cosT, &dR, &jac))
                continue;
            f3 cc = vmul(eval_reflective(c, surf, woL, v3(dot3(wi, T), dot3(wi, B), cosT)), l2);
            float phat = lum(cc);
            if (phat <= 0.0f || occluded(c, p, wi, dR * (1.0f - 1.0e-3f)))
                continue;
            float wgt = phat * R[25] * mq * jac;
            wsum += wgt;
            if (rng_next(rng) * wsum < wgt) {
                pSel = phat;
                cSel = cc;
                xSel = x2;
                nSel = n2;
                lSel = l2;
            }
        }
    }
    float* S = c->restir + restir_scratch(pixel);
    S[0] = wsum;
    S[1] = mTotal;
    S[2] = pSel;
    float cv[3] = {cSel.x, cSel.y, cSel.z}, xv[3] = {xSel.x, xSel.y, xSel.z};
    float nv[3] = {nSel.x, nSel.y, nSel.z}, lv[3] = {lSel.x, lSel.y, lSel.z};
    for (uint32_t k = 0; k < 3; k++) {
        S[3 + k] = cv[k];
        S[6 + k] = xv[k];
        S[9 + k] = nv[k];
        S[12 + k] = lv[k];
    }
    S[22] = asf(0u);
}

/* giFinish (pt_core.slang:1566-1617): the path's own candidate (its first-bounce surface, carrying
   what the rest of the path gathered, divided by the first bounce's throughput) joins the
   resampled neighbours; the pixel gets what the first hit gathered itself plus the pick's indirect
   light times its weight W = wsum / (M pSel); the new reservoir goes to this frame's half. The
   candidate stream hashes the pixel with the sample index (and the render's seed mix, zero for
   the reference's own seed) */
static f3 gi_finish(const ctx_t* c, f3 pathTotal, uint32_t pixel)
{
    float* C = c->restir + restir_base(pixel, c->P[6] & 1u);
    float* S = c->restir + restir_scratch(pixel);
    float wsum = S[0];
    float m = S[1] + 1.0f;
    float pSel = S[2];
    f3 cSel = v3(S[3], S[4], S[5]);
    f3 xSel = v3(S[6], S[7], S[8]);
    f3 nSel = v3(S[9], S[10], S[11]);
    f3 lSel = v3(S[12], S[13], S[14]);
    f3 thr1 = v3(S[15], S[16], S[17]);
    f3 d0 = v3(S[18], S[19], S[20]);
    float qPrev = S[21];
    int own = asu(S[22]) != 0 && qPrev > 0.0f;
    f3 giTotal = vsub(pathTotal, d0);
    f3 result = d0;
    if (own) {
        f3 cOwn = vscale(giTotal, qPrev);
        float pOwn = lum(cOwn);
        if (pOwn > 0.0f) {
            float wOwn = pOwn / qPrev;
            wsum += wOwn;
            uint32_t r = or_pcg(pixel ^ or_pcg(c->P[6] * 0x85EBCA6Bu + 0x1234u) ^ c->seed_mix);
            if (rng_next(&r) * wsum < wOwn) {
                pSel = pOwn;
                cSel = cOwn;
                xSel = v3(C[16], C[17], C[18]);
                nSel = v3(C[19], C[20], C[21]);
                lSel = v3(thr1.x > 1.0e-6f ? giTotal.x / thr1.x : 0.0f, thr1.y > 1.0e-6f ? giTotal.y / thr1.y : 0.0f,
                          thr1.z > 1.0e-6f ? giTotal.z / thr1.z : 0.0f);
            }
        }
    } else
        result = vadd(result, giTotal);
    float W = pSel > 0.0f ? wsum / (m * pSel) : 0.0f;
    result = vadd(result, vscale(cSel, W));
    C[16] = xSel.x;
    C[17] = xSel.y;
    C[18] = xSel.z;
    C[19] = nSel.x;
    C[20] = nSel.y;
    C[21] = nSel.z;
    C[22] = lSel.x;
    C[23] = lSel.y;
    C[24] = lSel.z;
    C[25] = W;
    C[26] = m;
    C[27] = asf(pSel > 0.0f ? 1u : 0u);
    return result;
}

/* ---------------------------------------------------------------- shadePath (pt_core.slang:992-1543) */

static int shade_path(const ctx_t* c, path_t* st, or_hit hq, uint32_t depth, uint32_t pixel)
{
    const or_tables* t = c->t;
    f3 sky = v3(PF(20), PF(21), PF(22));
    int lightingOn = c->P[28] != 0;
    uint32_t lightCount = c->P[5];
    uint32_t emitterCount = c->P[23];
    int hasEnv = c->P[24] > 0;
    float pEnv = hasEnv ? (emitterCount > 0 ? 0.5f : 1.0f) : 0.0f;
    int restirOn = c->P[36] != 0;
    uint32_t rng = st->rng;
    f3 org = st->org, dir = st->dir, thr = st->thr;
    uint32_t cls = st->cls;
    int chan = st->chan;
    int alive = 1;
    do {
        if (!hq.hit) {
            if (depth == 0 && restirOn)
                restir_clear(c, pixel);
            if (depth == 0 && c->P[42] != 0) {
                alive = 0;
                break;
            }
            parts_add(&st->parts, depth, cls, vmul(thr, sky));
            if (hasEnv) {
                float wB = 1.0f;
                if (st->misPrev) {
                    float pl = (st->restir && depth == 1 ? 1.0f : pEnv) * env_pdf(c, dir);
                    float den = st->qPrev * st->qPrev + pl * pl;
                    wB = den > 0.0f ? st->qPrev * st->qPrev / den : 1.0f;
                }
                parts_add(&st->parts, depth, cls, vscale(vmul(thr, env_radiance(c, dir)), wB));
            }
            alive = 0;
            break;
        }
        uint32_t inst = hq.inst * 16;
        uint32_t tri = t->instances[inst + 1] + 3 * hq.prim;
        uint32_t vo = t->instances[inst];
        uint32_t i0 = t->indices[tri] + vo, i1 = t->indices[tri + 1] + vo, i2 = t->indices[tri + 2] + vo;
        f3 p0 = transform_point(t->instances, inst + 4, vertex_pos(c, i0));
        f3 p1 = transform_point(t->instances, inst + 4, vertex_pos(c, i1));
        f3 p2 = transform_point(t->instances, inst + 4, vertex_pos(c, i2));
        f3 nOut = normalize3(cross3(vsub(p1, p0), vsub(p2, p0)));
        f3 wo = vneg(dir);
        int outside = dot3(nOut, wo) > 0.0f;
        f3 ng = outside ? nOut : vneg(nOut);
        float pHitT = hq.t;
        st->pathLen += pHitT;
        f3 hit = vadd(org, vscale(dir, pHitT));
        f3 e1 = vsub(p1, p0), e2 = vsub(p2, p0), e3 = vsub(hit, p0);
        float d11 = dot3(e1, e1), d12 = dot3(e1, e2), d22 = dot3(e2, e2);
        float den = fmax_(d11 * d22 - d12 * d12, 1.0e-20f);
        float b1 = (d22 * dot3(e3, e1) - d12 * dot3(e3, e2)) / den;
        float b2 = (d11 * dot3(e3, e2) - d12 * dot3(e3, e1)) / den;

        float opacity = opacity_at(c, hq.inst, hq.prim, b1, b2);
        if (opacity < 1.0f && rng_next(&rng) >= opacity) {
            org = vadd(hit, vscale(dir, 1.0e-4f));
            break;
        }
        f3 p = vadd(hit, vscale(ng, 1.0e-4f));
        if (depth == 1 && st->gi) {
            /* the first bounce reached a surface: ReSTIR GI's candidate point and facing normal (pt_core.slang:1072-1082) */
            float* C = c->restir + restir_base(pixel, c->P[6] & 1u);
            C[16] = hit.x;
            C[17] = hit.y;
            C[18] = hit.z;
            C[19] = ng.x;
            C[20] = ng.y;
            C[21] = ng.z;
            c->restir[restir_scratch(pixel) + 22] = asf(1u);
        }

        f3 nsObj = vadd(vadd(sscale(1.0f - b1 - b2, vertex_normal(c, i0)), sscale(b1, vertex_normal(c, i1))), sscale(b2, vertex_normal(c, i2)));
        f3 ns = nOut;
        if (dot3(nsObj, nsObj) > 1.0e-12f) {
            ns = normalize3(transform_normal(c, inst + 4, nsObj));
            if (dot3(ns, nOut) < 0.0f)
                ns = vneg(ns);
        }
        f3 n = outside ? ns : vneg(ns);

        uint32_t mi = t->instances[inst + 2] * 24;
        const float* M = t->materials + mi;
        f3 albedo = v3(M[0], M[1], M[2]);
        uint32_t texBase = f2u(M[10]), texMR = f2u(M[11]), texEm = f2u(M[12]), texNm = f2u(M[13]);
        float texRMx = 1.0f, texRMy = 1.0f;
        f3 texEmission = v3(1.0f, 1.0f, 1.0f);
        if (texBase + texMR + texEm + texNm > 0) {
            float u0x = t->uvs[2 * i0], u0y = t->uvs[2 * i0 + 1];
            float u1x = t->uvs[2 * i1], u1y = t->uvs[2 * i1 + 1];
            float u2x = t->uvs[2 * i2], u2y = t->uvs[2 * i2 + 1];
            float ux = u0x + b1 * (u1x - u0x) + b2 * (u2x - u0x);
            float uy = u0y + b1 * (u1y - u0y) + b2 * (u2y - u0y);
            float d1x = u1x - u0x, d1y = u1y - u0y, d2x = u2x - u0x, d2y = u2y - u0y;
            float uvPerWorld = sqrtf(fabsf(d1x * d2y - d1y * d2x) / fmax_(len3(cross3(e1, e2)), 1.0e-20f));
            float footprint = st->pathLen * PF(27) / fmax_(fabsf(dot3(nOut, dir)), 0.1f) * uvPerWorld;
            if (texBase > 0)
                albedo = vmul(albedo, sample_texture(c, texBase - 1, ux, uy, footprint));
            if (texMR > 0) {
                f3 mr = sample_texture(c, texMR - 1, ux, uy, footprint);
                texRMx = mr.y;
                texRMy = mr.z;
            }
            if (texEm > 0)
                texEmission = sample_texture(c, texEm - 1, ux, uy, footprint);
            if (texNm > 0) {
                float det = d1x * d2y - d2x * d1y;
                f3 dPdu = vdivs(vsub(vscale(e1, d2y), vscale(e2, d1y)), det);
                f3 dPdv = vdivs(vsub(vscale(e2, d1x), vscale(e1, d2x)), det);
                f3 tg = vsub(dPdu, vscale(ns, dot3(ns, dPdu)));
                if (fabsf(det) > 1.0e-20f && dot3(tg, tg) > 1.0e-20f) {
                    tg = normalize3(tg);
                    f3 bt = cross3(ns, tg);
                    if (dot3(bt, dPdv) > 0.0f)
                        bt = vneg(bt);
                    f3 ntx = sample_texture(c, texNm - 1, ux, uy, footprint);
                    f3 nt = v3(ntx.x * 2.0f - 1.0f, ntx.y * 2.0f - 1.0f, ntx.z * 2.0f - 1.0f);
                    ns = normalize3(vadd(vadd(vscale(tg, nt.x * M[14]), vscale(bt, nt.y * M[14])), vscale(ns, nt.z)));
                    n = outside ? ns : vneg(ns);
                }
            }
        }
        float woN = dot3(n, wo);
        if (woN < 0.05f)
            n = normalize3(vadd(n, vscale(wo, 0.05f - woN)));

        f3 le = vmul(v3(M[3], M[4], M[5]), texEmission);
        if (st->prevInst != 0 && (emit_mask(c, hq.inst) & recv_mask(c, st->prevInst - 1)) == 0)
            le = v3(0.0f, 0.0f, 0.0f);
        if (le.x > 0.0f || le.y > 0.0f || le.z > 0.0f) {
            float wB = 1.0f;
            uint32_t eb = t->instances[inst + 3];
            if (st->misPrev && eb != 0) {
                const float* E = t->emitters + (size_t)(eb - 1 + hq.prim) * 16;
                float cosl = fmax_(fabsf(dot3(nOut, dir)), 1.0e-4f);
                float pl = (st->restir && depth == 1 ? 1.0f : 1.0f - pEnv) * E[13] * pHitT * pHitT / (E[12] * cosl);
                wB = st->qPrev * st->qPrev / (st->qPrev * st->qPrev + pl * pl);
            }
            parts_add(&st->parts, depth, cls, vscale(vmul(thr, le), wB));
            if (depth == 1 && st->gi) {
                /* emission seen by the first bounce stays with the first hit's own light (pt_core.slang:1162-1168) */
                float* S = c->restir + restir_scratch(pixel) + 18;
                f3 a = vscale(vmul(thr, le), wB);
                S[0] += a.x;
                S[1] += a.y;
                S[2] += a.z;
            }
        }
        float metallic = M[6] * texRMy;
        float rough = M[7] * texRMx;
        float transmission = M[8];
        float ior = M[9];
        float dispersion = M[21];
        if (dispersion > 0.0f && transmission > 0.0f) {
            if (chan < 0) {
                int ch = (int)(rng_next(&rng) * 3.0f);
                chan = ch < 2 ? ch : 2;
                thr = vmul(thr, sscale(3.0f, v3(chan == 0 ? 1.0f : 0.0f, chan == 1 ? 1.0f : 0.0f, chan == 2 ? 1.0f : 0.0f)));
            }
            ior = fmax_(ior + (float)(chan - 1) * (ior - 1.0f) * 0.025f * dispersion, 1.0f);
        }
        /* a glass second hit is no reconnection candidate (pt_core.slang:1187-1190) */
        if (depth == 1 && st->gi && transmission > 0.0f)
            c->restir[restir_scratch(pixel) + 22] = asf(0u);
        float alpha = fmax_(rough * rough, 1.0e-3f);
        float eta = outside ? ior : 1.0f / ior;

        f3 T, B;
        basis(n, &T, &B);
        f3 woL = v3(dot3(wo, T), dot3(wo, B), fmax_(dot3(wo, n), 1.0e-4f));
        float fSel = or_fresnel(woL.z, ior);
        surface_t surf;
        surf.albedo = albedo;
        surf.metallic = metallic;
        surf.alpha = alpha;
        surf.transmission = transmission;
        surf.ior = ior;
        surf.eta = eta;
        surf.fSel = fSel;
        surf.tComp = 1.0f;
        if (c->P[29] != 0 && transmission > 0.0f && rough > 0.0f)
            surf.tComp = 1.0f / fmax_(dielectric_albedo(c, woL.z, rough, ior, outside), 0.1f);
        surf.coat = outside ? M[15] : 0.0f;
        surf.coatAlpha = fmax_(M[16] * M[16], 1.0e-3f);
        float coatF = or_fresnel(woL.z, 1.5f);
        surf.coatP = surf.coat * coatF;
        surf.sheenColor = outside ? v3(M[18], M[19], M[20]) : v3(0.0f, 0.0f, 0.0f);
        float sheenMax = fmax_(surf.sheenColor.x, fmax_(surf.sheenColor.y, surf.sheenColor.z));
        float sheenR = fmax_(M[17], 0.07f);
        surf.sheenAlpha = sheenR * sheenR;
        surf.sheenP = sheenMax > 0.0f ? sheenMax * sheen_albedo(c, woL.z, sheenR) : 0.0f;

        /* delta lights (pt_core.slang:1224-1282) */
        for (uint32_t li = 0; li < lightCount; li++) {
            const uint32_t* L = t->lights + li * 16;
            f3 col = vscale(v3(asf(L[1]), asf(L[2]), asf(L[3])), asf(L[4]));
            f3 vec = v3(asf(L[5]), asf(L[6]), asf(L[7]));
            int virt = L[0] == 3;
            if (virt && !(lightingOn && depth == 0))
                continue;
            if ((L[14] & recv_mask(c, hq.inst)) == 0)
                continue;
            f3 wi, li_;
            float dist;
            if (L[0] == 1 || virt) {
                wi = vneg(normalize3(vec));
                dist = 1.0e30f;
                li_ = col;
            } else {
                f3 toL = vsub(vec, p);
                float d2 = dot3(toL, toL);
                dist = sqrtf(d2);
                wi = vdivs(toL, dist);
                li_ = vdivs(col, d2);
                if (L[15] != 0)
                    li_ = vscale(li_, ies_factor(c, L[15] - 1, vneg(wi)));
                if (L[0] == 2) {
                    f3 axis = v3(asf(L[8]), asf(L[9]), asf(L[10]));
                    float att = saturate_(dot3(axis, vneg(wi)) * asf(L[11]) + asf(L[12]));
                    li_ = vscale(li_, att * att);
                }
            }
            float cc = dot3(n, wi);
            if (cc > 0.0f) {
                int occ = occluded(c, p, wi, dist - 1.0e-3f);
                int track = lightingOn && depth == 0;
                if (!occ || track) {
                    lobes_t l = eval_lobes(c, &surf, woL, v3(dot3(wi, T), dot3(wi, B), cc));
                    f3 cd = vmul(vmul(thr, l.diffuse), li_);
                    f3 cs = vmul(vmul(thr, l.specular), li_);
                    if (!occ && !virt)
                        parts_add_light(&st->parts, depth, cls, vmul(vmul(thr, vadd(l.diffuse, l.specular)), li_), cd, cs);
                    if (track) {
                        float vv = virt ? 0.0f : lum(vadd(cd, cs));
                        st->visTotal += vv;
                        st->visLit += occ ? 0.0f : vv;
                        add_slot(c, pixel, L[13], vdivs(cd, cc), vdivs(cs, cc), wi, occ ? 0.0f : 1.0f, occ ? 0.0f : cc);
                    }
                }
            }
        }

        /* ReSTIR DI at the first hit (pt_core.slang:1286-1332) */
        int reDI = 0;
        if (depth == 0 && restirOn) {
            int suits = PF(3) == 0.0f && transmission == 0.0f && surf.alpha >= RESTIR_MIN_ALPHA
                && (surf.coat == 0.0f || surf.coatAlpha >= RESTIR_MIN_ALPHA);
            reDI = suits && (c->P[36] & 1u) != 0 && emitterCount > 0;
            int reGI = suits && (c->P[36] & 2u) != 0 && !lightingOn;
            st->restir = reDI;
            st->gi = reGI;
            if (suits) {
                restir_gbuffer(c, pixel, p, n);
                if (reGI)
                    restir_gi_prepare(c, pixel, p, n, pHitT, &surf, T, B, woL, &rng);
            } else
                restir_clear(c, pixel);
        }
        if (reDI) {
            restir_pick pick = restir_di(c, pixel, recv_mask(c, hq.inst), p, n, pHitT, &surf, T, B, woL, &rng);
            if (pick.valid) {
                f3 wi, g;
                float dist, cosi, plsa;
                restir_light(c, pick.e - 1, recv_mask(c, hq.inst), pick.x, p, n, &wi, &dist, &cosi, &g, &plsa);
                int track = lightingOn;
                if (!pick.occ || track) {
                    f3 wiL = v3(dot3(wi, T), dot3(wi, B), cosi);
                    float wl = restir_mis(&surf, woL, wiL, plsa);
                    lobes_t l = eval_lobes(c, &surf, woL, wiL);
                    f3 cd = vscale(vmul(vmul(thr, l.diffuse), g), wl * pick.W);
                    f3 cs = vscale(vmul(vmul(thr, l.specular), g), wl * pick.W);
                    if (!pick.occ)
                        parts_add_light(&st->parts, depth, cls, vadd(cd, cs), cd, cs);
                    if (track) {
                        float vv = lum(vadd(cd, cs));
                        st->visTotal += vv;
                        st->visLit += pick.occ ? 0.0f : vv;
                        add_slot(c, pixel, 3u, vdivs(cd, cosi), vdivs(cs, cosi), wi, pick.occ ? 0.0f : 1.0f, pick.occ ? 0.0f : cosi);
                    }
                }
            }
        }

        /* emitter NEE (pt_core.slang:1335-1390) */
        float chooseLight = rng_next(&rng);
        float pEnvHere = reDI ? (hasEnv ? 1.0f : 0.0f) : pEnv;
        int chooseEnv = hasEnv && chooseLight < pEnvHere;
        if (emitterCount > 0 && !chooseEnv && !reDI) {
            float ue = rng_next(&rng), ua = rng_next(&rng), ub = rng_next(&rng);
            uint32_t lo = 0, hi = emitterCount - 1;
            while (lo < hi) {
                uint32_t mid = (lo + hi) / 2;
                if (t->emit_cdf[mid] >= ue)
                    hi = mid;
                else
                    lo = mid + 1;
            }
            const float* E = t->emitters + (size_t)lo * 16;
            f3 a = v3(E[0], E[1], E[2]), b = v3(E[3], E[4], E[5]), cc = v3(E[6], E[7], E[8]);
            float su = sqrtf(ua);
            f3 xl = vadd(vadd(vscale(a, 1.0f - su), vscale(b, su * (1.0f - ub))), vscale(cc, su * ub));
            f3 toL = vsub(xl, p);
            float d2 = dot3(toL, toL);
            float dist = sqrtf(d2);
            f3 wi = vdivs(toL, dist);
            float cosl = fabsf(dot3(normalize3(cross3(vsub(b, a), vsub(cc, a))), wi));
            float cosi = dot3(n, wi);
            if (cosi > 0.0f && cosl > 1.0e-5f && (asu(E[14]) & recv_mask(c, hq.inst)) != 0) {
                int occ = occluded(c, p, wi, dist * (1.0f - 1.0e-3f));
                int track = lightingOn && depth == 0;
                if (!occ || track) {
                    float pl = (1.0f - pEnv) * E[13] * d2 / (E[12] * cosl);
                    f3 wiL = v3(dot3(wi, T), dot3(wi, B), cosi);
                    float qb = reflective_pdf(&surf, woL, wiL);
                    float wl = pl * pl / (pl * pl + qb * qb);
                    f3 lev = v3(E[9], E[10], E[11]);
                    lobes_t l = eval_lobes(c, &surf, woL, wiL);
                    f3 cd = vscale(vdivs(vmul(vmul(thr, l.diffuse), lev), pl), wl);
                    f3 cs = vscale(vdivs(vmul(vmul(thr, l.specular), lev), pl), wl);
                    if (!occ)
                        parts_add_light(&st->parts, depth, cls, vscale(vdivs(vmul(vmul(thr, eval_reflective(c, &surf, woL, wiL)), lev), pl), wl), cd, cs);
                    if (track) {
                        float vv = lum(vadd(cd, cs));
                        st->visTotal += vv;
                        st->visLit += occ ? 0.0f : vv;
                        add_slot(c, pixel, 3u, vdivs(cd, cosi), vdivs(cs, cosi), wi, occ ? 0.0f : 1.0f, occ ? 0.0f : cosi);
                    }
                }
            }
        }

        /* environment NEE (pt_core.slang:1392-1427) */
        if (chooseEnv) {
            float u1e = rng_next(&rng), u2e = rng_next(&rng), u3e = rng_next(&rng), u4e = rng_next(&rng);
            float pdfE;
            f3 leE;
            f3 wi = env_sample(c, u1e, u2e, u3e, u4e, &pdfE, &leE);
            float cosi = dot3(n, wi);
            if (pdfE > 0.0f && cosi > 0.0f) {
                int occ = occluded(c, p, wi, 1.0e30f);
                int track = lightingOn && depth == 0;
                if (!occ || track) {
                    float pl = pEnvHere * pdfE;
                    f3 wiL = v3(dot3(wi, T), dot3(wi, B), cosi);
                    float qb = reflective_pdf(&surf, woL, wiL);
                    float wl = pl * pl / (pl * pl + qb * qb);
                    lobes_t l = eval_lobes(c, &surf, woL, wiL);
                    f3 cd = vscale(vdivs(vmul(vmul(thr, l.diffuse), leE), pl), wl);
                    f3 cs = vscale(vdivs(vmul(vmul(thr, l.specular), leE), pl), wl);
                    if (!occ)
                        parts_add_light(&st->parts, depth, cls, vscale(vdivs(vmul(vmul(thr, eval_reflective(c, &surf, woL, wiL)), leE), pl), wl), cd, cs);
                    if (track) {
                        float vv = lum(vadd(cd, cs));
                        st->visTotal += vv;
                        st->visLit += occ ? 0.0f : vv;
                        if (c->P[41] == 0)
                            add_slot(c, pixel, 3u, vdivs(cd, cosi), vdivs(cs, cosi), wi, occ ? 0.0f : 1.0f, occ ? 0.0f : cosi);
                    }
                }
            }
        }

        /* one-lobe BSDF sampling (pt_core.slang:1431-1523) */
        float lobe = rng_next(&rng);
        float u1 = rng_next(&rng), u2 = rng_next(&rng), u3 = rng_next(&rng);
        f3 wiL = v3(0.0f, 0.0f, 1.0f);
        f3 wgt = v3(1.0f, 1.0f, 1.0f);
        int ok = 1, refracted = 0;
        uint32_t lobeCls = 1;
        float uc = surf.coat > 0.0f ? rng_next(&rng) : 1.0f;
        float us = surf.sheenP > 0.0f ? rng_next(&rng) : 1.0f;
        if (uc < surf.coatP) {
            f3 h = sample_vndf(woL, surf.coatAlpha, u1, u2);
            float woh = dot3(woL, h);
            wiL = vsub(vscale(h, 2.0f * woh), woL);
            ok = wiL.z > 0.0f;
            float k = or_fresnel(woh, 1.5f) / coatF * or_G1(fmax_(wiL.z, 0.0f), surf.coatAlpha);
            wgt = v3(k, k, k);
        } else if (us < surf.sheenP) {
            wiL = cosine_local(u1, u2);
            f3 hh = normalize3(vadd(woL, wiL));
            float sc = or_sheen_core(woL.z, wiL.z, hh.z, surf.sheenAlpha);
            float den = wiL.z * surf.sheenP;
            wgt = vdivs(vscale(vscale(surf.sheenColor, sc), OR_PI), den);
        } else if (lobe < metallic) {
            f3 h = sample_vndf(woL, alpha, u1, u2);
            float woh = dot3(woL, h);
            wiL = vsub(vscale(h, 2.0f * woh), woL);
            ok = wiL.z > 0.0f;
            wgt = vscale(vmul(schlick3(albedo, woh), multiscatter_factor(c, &surf, woL.z)), or_G1(fmax_(wiL.z, 0.0f), alpha));
        } else if (lobe < metallic + (1.0f - metallic) * transmission) {
            f3 h = sample_vndf(woL, alpha, u1, u2);
            float woh = dot3(woL, h);
            if (u3 < or_fresnel(woh, eta)) {
                wiL = vsub(vscale(h, 2.0f * woh), woL);
                ok = wiL.z > 0.0f;
                float k = or_G1(fmax_(wiL.z, 0.0f), alpha) * surf.tComp;
                wgt = v3(k, k, k);
            } else {
                float r = 1.0f / eta;
                float k = 1.0f - r * r * (1.0f - woh * woh);
                wiL = vadd(vscale(woL, -r), vscale(h, r * woh - sqrtf(fmax_(k, 0.0f))));
                ok = wiL.z < 0.0f && k > 0.0f;
                refracted = 1;
                lobeCls = 2;
                wgt = vscale(vscale(vscale(vscale(albedo, or_G1(fmax_(-wiL.z, 0.0f), alpha)), r), r), surf.tComp);
            }
        } else if (u3 < fSel) {
            f3 h = sample_vndf(woL, alpha, u1, u2);
            float woh = dot3(woL, h);
            wiL = vsub(vscale(h, 2.0f * woh), woL);
            ok = wiL.z > 0.0f;
            float k = or_fresnel(woh, ior) / fSel * or_G1(fmax_(wiL.z, 0.0f), alpha);
            wgt = v3(k, k, k);
        } else {
            wiL = cosine_local(u1, u2);
            wgt = albedo;
            lobeCls = 0;
        }
        if (!ok) {
            alive = 0;
            break;
        }
        if (depth == 0)
            cls = lobeCls;
        st->misPrev = !refracted;
        st->qPrev = refracted ? 0.0f : reflective_pdf(&surf, woL, wiL);
        st->prevInst = hq.inst + 1;

        dir = normalize3(vadd(vadd(vscale(T, wiL.x), vscale(B, wiL.y)), vscale(n, wiL.z)));
        float side = dot3(dir, ng);
        if (refracted ? side >= 0.0f : side <= 0.0f) {
            alive = 0;
            break;
        }
        org = vadd(hit, vscale(ng, side > 0.0f ? 1.0e-4f : -1.0e-4f));
        thr = vmul(thr, wgt);

        if (depth >= 2) {
            float pr = clampf_(fmax_(thr.x, fmax_(thr.y, thr.z)), 0.05f, 0.95f);
            if (rng_next(&rng) > pr) {
                alive = 0;
                break;
            }
            thr = vdivs(thr, pr);
        }
    } while (0);
    st->rng = rng;
    st->org = org;
    st->dir = dir;
    st->thr = thr;
    st->cls = cls;
    st->chan = chan;
    if (depth == 0 && st->gi) {
        /* the first hit's own gathered light and the first bounce's throughput and pdf (pt_core.slang:1531-1540) */
        float* S = c->restir + restir_scratch(pixel);
        S[15] = thr.x;
        S[16] = thr.y;
        S[17] = thr.z;
        S[18] = st->parts.total.x;
        S[19] = st->parts.total.y;
        S[20] = st->parts.total.z;
        S[21] = st->qPrev;
    }
    return alive;
}

/* pixelActive (pt_core.slang:1551-1559), in FP64 */
static int pixel_active(const ctx_t* c, uint32_t p)
{
    double n = c->stats[3 * (size_t)p];
    if (n < (double)c->P[34])
        return 1;
    double mean = c->stats[3 * (size_t)p + 1] / n;
    double var = fmax(c->stats[3 * (size_t)p + 2] / n - mean * mean, 0.0) * n / (n - 1.0);
    return sqrt(var / n) > (double)PF(33) * fmax(mean, 0.05);
}

/* finishPath (pt_core.slang:1620-1667), FP64 film */
static void finish_path(const ctx_t* c, const path_t* st, uint32_t pixel)
{
    parts_t parts = st->parts;
    float visLit = st->visLit, visTotal = st->visTotal;
    f3 radiance = parts.total;
    if (st->gi) {
        /* a path that went through dispersive glass carries one tripled channel: not shared (pt_core.slang:1626-1632) */
        if (st->chan >= 0)
            c->restir[restir_scratch(pixel) + 22] = asf(0u);
        radiance = gi_finish(c, radiance, pixel);
    }
    if (bad3(radiance)) {
        radiance = v3(0.0f, 0.0f, 0.0f);
        memset(&parts, 0, sizeof(parts));
        visLit = 0.0f;
        visTotal = 0.0f;
    }
    double* F = c->film + 3 * (size_t)pixel;
    F[0] += (double)radiance.x;
    F[1] += (double)radiance.y;
    F[2] += (double)radiance.z;
    if (c->P[32] != 0) {
        double l = (double)lum(radiance);
        double* S = c->stats + 3 * (size_t)pixel;
        S[0] += 1.0;
        S[1] += l;
        S[2] += l * l;
    }
    if (c->P[28] != 0) {
        float* L = c->lighting + (size_t)65 * pixel;
        f3 split[5] = {parts.directDiffuse, parts.indirectDiffuse, parts.specular, parts.transmission, parts.emission};
        for (uint32_t k = 0; k < 5; k++) {
            L[3 * k] += split[k].x;
            L[3 * k + 1] += split[k].y;
            L[3 * k + 2] += split[k].z;
        }
        L[15] += visLit;
        L[16] += visTotal;
    }
}

static uint32_t seed_mix(uint32_t seed) { return seed == 0 ? 0u : or_pcg(seed ^ 0x5851F42Du); }

/* One sample of every pixel (one megakernel dispatch with params[35] = 1, pathtrace.slang:9-25).
 * `params` is the 43-word block of pack::base_params (pack.rs:828-874) with word 6 = the sample. */
static void render_sample(ctx_t* c, int threads)
{
    uint32_t w = c->P[0], h = c->P[1];
    int64_t npx = (int64_t)w * h;
    (void)threads;
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
    for (int64_t pi = 0; pi < npx; pi++) {
        uint32_t pixel = (uint32_t)pi;
        if (c->P[32] != 0 && !pixel_active(c, pixel))
            continue;
        path_t st = generate_path(c, pixel % w, pixel / w, c->P[6]);
        for (uint32_t depth = 0; depth < c->P[7]; depth++) {
            or_hit hq = c->brute ? trace_brute(c->s, st.org, st.dir, 1.0e30f) : trace_bvh(c->s, st.org, st.dir, 1.0e30f, 0);
            if (!shade_path(c, &st, hq, depth, pixel))
                break;
        }
        finish_path(c, &st, pixel);
    }
}

/* Renders `samples` samples from `first_sample` into the running sums (any of which may be NULL
 * when its feature is off). Samples run one after another, pixels in parallel. */
int or_render(const or_scene* s, const uint32_t* params, uint32_t seed, uint32_t first_sample, uint32_t samples,
              double* film, double* stats, float* lighting, float* restir, int threads, int brute)
{
    uint32_t P[43];
    memcpy(P, params, sizeof(P));
    if ((P[32] && !stats) || (P[28] && !lighting) || (P[36] && !restir))
        return 1;
    for (uint32_t k = 0; k < samples; k++) {
        P[6] = first_sample + k;
        ctx_t c = {s, &s->t, P, film, stats, lighting, restir, seed_mix(seed), brute};
        render_sample(&c, threads > 0 ? threads : 1);
    }
    return 0;
}

/* Closest hits for `n` rays (org xyz, dir xyz each): 4 words out per ray (hit, inst, prim, t bits). */
void or_trace_rays(const or_scene* s, uint64_t n, const float* rays, float tmax, int brute, uint32_t* out)
{
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        const float* r = rays + 6 * i;
        f3 o = v3(r[0], r[1], r[2]), d = v3(r[3], r[4], r[5]);
        or_hit h = brute ? trace_brute(s, o, d, tmax) : trace_bvh(s, o, d, tmax, 0);
        out[4 * i] = h.hit ? 1u : 0u;
        out[4 * i + 1] = h.hit ? h.inst : 0u;
        out[4 * i + 2] = h.hit ? h.prim : 0u;
        out[4 * i + 3] = h.hit ? asu(h.t) : 0u;
    }
}

/* Camera rays of sample `sample` for every pixel, then their closest hits (4 words per pixel). */
void or_primary_hits(const or_scene* s, const uint32_t* params, uint32_t seed, uint32_t sample, int brute, uint32_t* out, float* rays)
{
    uint32_t P[43];
    memcpy(P, params, sizeof(P));
    P[6] = sample;
    ctx_t c = {s, &s->t, P, NULL, NULL, NULL, NULL, seed_mix(seed), brute};
    uint32_t w = P[0], h = P[1];
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t pi = 0; pi < (int64_t)w * h; pi++) {
        path_t st = generate_path(&c, (uint32_t)(pi % w), (uint32_t)(pi / w), sample);
        or_hit hq = brute ? trace_brute(s, st.org, st.dir, 1.0e30f) : trace_bvh(s, st.org, st.dir, 1.0e30f, 0);
        out[4 * pi] = hq.hit ? 1u : 0u;
        out[4 * pi + 1] = hq.hit ? hq.inst : 0u;
        out[4 * pi + 2] = hq.hit ? hq.prim : 0u;
        out[4 * pi + 3] = hq.hit ? asu(hq.t) : 0u;
        if (rays) {
            float* r = rays + 6 * pi;
            r[0] = st.org.x; r[1] = st.org.y; r[2] = st.org.z;
            r[3] = st.dir.x; r[4] = st.dir.y; r[5] = st.dir.z;
        }
    }
}

/* ---------------------------------------------------------------- primary-hit AOVs */

/* pixelDir (aov.slang:73-81; the same expressions inline in normals.slang:47-57): the camera-space
 * unit direction through pixel position (pxx, pxy) and, in `dir`, its world-space direction. */
static f3 pixel_dir(const ctx_t* c, float pxx, float pxy, f3* dir)
{
    uint32_t w = c->P[0], h = c->P[1];
    float t = PF(2);
    float aspect = (float)w / (float)h;
    float u = (2.0f * pxx / (float)w - 1.0f + PF(39)) * aspect * t;
    float v = (1.0f - 2.0f * pxy / (float)h + PF(40)) * t;
    f3 d = normalize3(v3(u, v, -1.0f));
    float r3[3];
    for (uint32_t r = 0; r < 3; r++)
        r3[r] = PF(8 + 4 * r) * d.x + PF(9 + 4 * r) * d.y + PF(10 + 4 * r) * d.z;
    *dir = v3(r3[0], r3[1], r3[2]);
    return d;
}

static f3 camera_origin(const ctx_t* c) { return v3(PF(11), PF(15), PF(19)); }

/* World-space triangle of a hit and the barycentric weights of its vertices 1 and 2 at `hit`
 * (normals.slang:77-95, aov.slang:117-132). */
static void hit_frame(const ctx_t* c, uint32_t instance, uint32_t prim, f3 hit, f3 p[3], uint32_t vi[3], float* b1, float* b2)
{
    const or_tables* T = c->t;
    uint32_t inst = instance * 16;
    uint32_t tri = T->instances[inst + 1] + 3 * prim;
    uint32_t vo = T->instances[inst];
    for (int k = 0; k < 3; k++) {
        vi[k] = T->indices[tri + k] + vo;
        p[k] = transform_point(T->instances, inst + 4, vertex_pos(c, vi[k]));
    }
    f3 e1 = vsub(p[1], p[0]), e2 = vsub(p[2], p[0]), e3 = vsub(hit, p[0]);
    float d11 = dot3(e1, e1), d12 = dot3(e1, e2), d22 = dot3(e2, e2);
    float den = fmax_(d11 * d22 - d12 * d12, 1.0e-20f);
    *b1 = (d22 * dot3(e3, e1) - d12 * dot3(e3, e2)) / den;
    *b2 = (d11 * dot3(e3, e2) - d12 * dot3(e3, e1)) / den;
}

/* normals (normals.slang:39-110): the pixel-centre ray's interpolated vertex normal (geometric
 * normal without vertex normals), on the geometric normal's side, facing the viewer; 0 on a miss.
 * 3 floats per pixel. `params` is the 43-word block (words 0-1 the image size, the camera). */
void or_render_normals(const or_scene* s, const uint32_t* params, int brute, float* out)
{
    ctx_t cc = {s, &s->t, params, NULL, NULL, NULL, NULL, 0u, brute};
    const ctx_t* c = &cc;
    uint32_t w = params[0], h = params[1];
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t pi = 0; pi < (int64_t)w * h; pi++) {
        uint32_t x = (uint32_t)(pi % w), y = (uint32_t)(pi / w);
        f3 dir;
        pixel_dir(c, (float)x + 0.5f, (float)y + 0.5f, &dir);
        f3 org = camera_origin(c);
        dir = normalize3(dir);
        or_hit q = trace(c, org, dir, 1.0e30f, 0);
        f3 n = v3(0.0f, 0.0f, 0.0f);
        if (q.hit) {
            f3 p[3];
            uint32_t vi[3];
            float b1, b2;
            hit_frame(c, q.inst, q.prim, vadd(org, vscale(dir, q.t)), p, vi, &b1, &b2);
            f3 ng = normalize3(cross3(vsub(p[1], p[0]), vsub(p[2], p[0])));
            n = ng;
            f3 nsObj = vadd(vadd(sscale(1.0f - b1 - b2, vertex_normal(c, vi[0])), sscale(b1, vertex_normal(c, vi[1]))),
                            sscale(b2, vertex_normal(c, vi[2])));
            if (dot3(nsObj, nsObj) > 1.0e-12f) {
                n = normalize3(transform_normal(c, q.inst * 16 + 4, nsObj));
                if (dot3(n, ng) < 0.0f)
                    n = vneg(n);
            }
            if (dot3(ng, dir) > 0.0f)
                n = vneg(n);
        }
        out[3 * pi] = n.x;
        out[3 * pi + 1] = n.y;
        out[3 * pi + 2] = n.z;
    }
}

/* aov (aov.slang:84-188): per pixel, from the pixel-centre ray, albedo (base colour times the
 * level-0 base colour texture) and depth along the camera's forward axis (4 floats), instance + 1
 * and material + 1 (0 on a miss); and the coverage of 16 stratified rays through up to 8
 * translucent layers, summed in ray order and divided by 16. */
void or_render_aovs(const or_scene* s, const uint32_t* params, int brute, float* albedo_depth, uint32_t* object_id,
                    uint32_t* material_id, float* coverage)
{
    ctx_t cc = {s, &s->t, params, NULL, NULL, NULL, NULL, 0u, brute};
    const ctx_t* c = &cc;
    const or_tables* T = &s->t;
    uint32_t w = params[0], h = params[1];
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t pi = 0; pi < (int64_t)w * h; pi++) {
        uint32_t x = (uint32_t)(pi % w), y = (uint32_t)(pi / w);
        f3 dir;
        f3 d = pixel_dir(c, (float)x + 0.5f, (float)y + 0.5f, &dir);
        f3 org = camera_origin(c);
        dir = normalize3(dir);
        or_hit q = trace(c, org, dir, 1.0e30f, 0);
        f3 albedo = v3(0.0f, 0.0f, 0.0f);
        float depth = 0.0f;
        uint32_t id = 0, matId = 0;
        if (q.hit) {
            uint32_t inst = q.inst * 16;
            id = q.inst + 1;
            matId = T->instances[inst + 2] + 1;
            depth = q.t * -d.z;
            const float* M = T->materials + (size_t)T->instances[inst + 2] * 24;
            albedo = v3(M[0], M[1], M[2]);
            uint32_t texBase = f2u(M[10]);
            if (texBase > 0) {
                f3 p[3];
                uint32_t vi[3];
                float b1, b2;
                hit_frame(c, q.inst, q.prim, vadd(org, vscale(dir, q.t)), p, vi, &b1, &b2);
                float u0x = T->uvs[2 * vi[0]], u0y = T->uvs[2 * vi[0] + 1];
                float u1x = T->uvs[2 * vi[1]], u1y = T->uvs[2 * vi[1] + 1];
                float u2x = T->uvs[2 * vi[2]], u2y = T->uvs[2 * vi[2] + 1];
                float ux = u0x + b1 * (u1x - u0x) + b2 * (u2x - u0x);
                float uy = u0y + b1 * (u1y - u0y) + b2 * (u2y - u0y);
                albedo = vmul(albedo, sample_texture(c, texBase - 1, ux, uy, 0.0f));
            }
        }
        float cover = 0.0f;
        for (uint32_t k = 0; k < 16; k++) {
            f3 sdir;
            pixel_dir(c, (float)x + ((float)(k % 4) + 0.5f) / 4.0f, (float)y + ((float)(k / 4) + 0.5f) / 4.0f, &sdir);
            f3 so = org, sd = normalize3(sdir);
            float seen = 0.0f, through = 1.0f;
            for (uint32_t layer = 0; layer < 8 && through > 0.0f; layer++) {
                or_hit sq = trace(c, so, sd, 1.0e30f, 0);
                if (!sq.hit)
                    break;
                float op = opacity_at(c, sq.inst, sq.prim, sq.b1, sq.b2);
                seen += through * op;
                through *= 1.0f - op;
                so = vadd(so, vscale(sd, sq.t + 1.0e-4f));
            }
            cover += seen;
        }
        coverage[pi] = cover / 16.0f;
        albedo_depth[4 * pi] = albedo.x;
        albedo_depth[4 * pi + 1] = albedo.y;
        albedo_depth[4 * pi + 2] = albedo.z;
        albedo_depth[4 * pi + 3] = depth;
        object_id[pi] = id;
        material_id[pi] = matId;
    }
}

/* ---------------------------------------------------------------- unit-test exports */

float or_sin(float x) { return or_sinf(x); }
float or_cos(float x) { return or_cosf(x); }
float or_acos(float x) { return or_acosf(x); }
float or_atan2(float y, float x) { return or_atan2f(y, x); }
float or_log2(float x) { return or_log2f(x); }
float or_exp2(float x) { return or_exp2f(x); }
float or_pow(float x, float y) { return or_powf(x, y); }
float or_rng_next(uint32_t* state) { return rng_next(state); }
float or_diff_prod(float a, float b, float c, float d) { return diff_prod(a, b, c, d); }

int or_intersect(const float* tri9, const float* org, const float* dir, float tmax, float* out3)
{
    float t, b1, b2;
    int hit = intersect_tri(tri9, v3(org[0], org[1], org[2]), v3(dir[0], dir[1], dir[2]), tmax, &t, &b1, &b2);
    if (hit) {
        out3[0] = t;
        out3[1] = b1;
        out3[2] = b2;
    }
    return hit;
}
int or_pixel_active(const double* stats3, uint32_t min_spp, float threshold)
{
    uint32_t P[43] = {0};
    P[34] = min_spp;
    P[33] = asu(threshold);
    ctx_t c = {NULL, NULL, P, NULL, (double*)stats3, NULL, NULL, 0, 0};
    return pixel_active(&c, 0);
}
