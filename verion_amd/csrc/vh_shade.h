// vh_shade.h - one path bounce on the GPU: camera rays, BSDF, light sampling, ReSTIR DI and GI, film.
//
// The device form of shaders/include/pt_core.slang: generatePathFor (:645-682), the BSDF
// (:115-557), occluded (:580-613), ReSTIR DI (:702-900), ReSTIR GI (:901-980, 1566-1617),
// shadePath (:992-1543), pixelActive (:1551-1559), finishPath (:1620-1667). Expressions keep the shader's evaluation order (the
// library is built with contraction off), so the megakernel and the wavefront kernels, which both
// call these functions, give identical bits, and the CPU oracle reproduces them.
// Deviations shared with the oracle: FP64 film and adaptive statistics; closest hits with ties to
// the smallest (instance, primitive); closest-hit shadow rays when a material is translucent;
// ReSTIR GI's own-candidate draw hashes in the render's seed mix (zero for the reference's stream).
#pragma once

#include "vh_trace.h"

namespace vh {

constexpr uint32_t kParamWords = 43;
constexpr uint32_t kLightingStride = 65;
constexpr uint32_t kSlotWords = 12;
constexpr uint32_t kGgxN = 32, kSheenN = 16, kDielN = 16, kDielNIor = 9;
constexpr uint32_t kAlbedoGgx = 0;
constexpr uint32_t kAlbedoSheen = kGgxN * kGgxN;
constexpr uint32_t kAlbedoDiel = kAlbedoSheen + kSheenN * kSheenN;
constexpr uint32_t kAlbedoWords = kAlbedoDiel + kDielN * kDielN * kDielNIor * 2;
constexpr float kRestirMinAlpha = 0.15f;
constexpr int kRestirRadius = 6;
constexpr uint32_t kRestirNeighbours = 2;
constexpr uint32_t kRestirMaxNeighbours = 16;
constexpr uint32_t kRestirHalf = 28;
constexpr uint32_t kRestirStride = 80;
constexpr float kRestirMaxJacobian = 10.0f;

// Scene tables in device memory (the bindings of pt_core.slang:12-49, minus the RT structure).
struct Scene {
    const float* positions;
    const uint32_t* indices;
    const float* uvs;
    const float* vnormals;
    const uint32_t* instances;
    const float* materials;
    const uint32_t* lights;
    const float* lightDirs; // per delta light, -normalize(its direction words) (4 floats; directional lights read it)
    const float* emitters;
    const float* emitCdf;
    const uint32_t* emitGuide; // 2^emitGuideBits + 1 cutpoints into emitCdf (emitterIndex)
    const float* envPixels;
    const float* envMarginal;
    const float* envCond;
    const uint32_t* envGuide; // cutpoints of envMarginal, then of each row of envCond (envSample)
    const float* texels;
    const uint32_t* texInfo;
    Bvh bvh;
    uint32_t translucent; // some material has opacity < 1 or an opacity texture
    uint32_t emitGuideBits;
    uint32_t envGuideBits; // bits of the marginal's guide | bits of a row's guide << 8
};

// Running sums of one accumulation.
struct Frame {
    double* film;     // 3 per pixel
    double* stats;    // 3 per pixel (adaptive only)
    float* lighting;  // 65 per pixel (lighting AOVs only)
    float* restir;    // 80 per pixel (ReSTIR only)
};

// The 43-word parameter block of pack::base_params (pack.rs:828-874) plus the seed mix.
struct Params {
    uint32_t w[kParamWords];
    uint32_t seedMix;
    __device__ __forceinline__ float f(int i) const { return asf(w[i]); }
};

struct Ctx {
    const Scene& S;
    const Frame& F;
    const Params& P;
};

struct Surface {
    V3 albedo;
    float metallic, alpha, transmission, ior, eta, fSel, tComp, coat, coatAlpha;
    V3 sheenColor;
    float sheenAlpha, sheenP, coatP;
    V3 ms;         // multiscatterFactor at the vertex's outgoing direction (the same for every lobe evaluation there)
    float g1o, g1oC; // G1 of the outgoing direction for the base and the coat roughness (likewise)
};

struct Lobes {
    V3 diffuse, specular;
};

struct Parts {
    V3 total, directDiffuse, indirectDiffuse, specular, transmission, emission;
};

// PathState (pt_core.slang:616-633)
struct PathState {
    V3 org, dir, thr;
    uint32_t rng, cls;
    int chan;
    float pathLen;
    bool misPrev, gi, restir;
    float qPrev;
    uint32_t prevInst;
    float visLit, visTotal;
    Parts parts;
};

__device__ __forceinline__ V3 vtx(const float* a, uint32_t v) { return v3(a[3 * v], a[3 * v + 1], a[3 * v + 2]); }

__device__ __forceinline__ V3 transformPoint(const uint32_t* I, uint32_t base, V3 p)
{
    float r[3];
#pragma unroll
    for (uint32_t row = 0; row < 3; row++) {
        uint32_t o = base + 4 * row;
        r[row] = asf(I[o]) * p.x + asf(I[o + 1]) * p.y + asf(I[o + 2]) * p.z + asf(I[o + 3]);
    }
    return v3(r[0], r[1], r[2]);
}

// The same two transforms on a row-major 3 x 4 matrix already in registers (same operations, same
// order, same bits).
__device__ __forceinline__ V3 transformPoint(const float (&X)[12], V3 p)
{
    float r[3];
#pragma unroll
    for (uint32_t row = 0; row < 3; row++)
        r[row] = X[4 * row] * p.x + X[4 * row + 1] * p.y + X[4 * row + 2] * p.z + X[4 * row + 3];
    return v3(r[0], r[1], r[2]);
}

__device__ __forceinline__ V3 transformNormal(const float (&X)[12], V3 n)
{
    V3 r0 = v3(X[0], X[1], X[2]);
    V3 r1 = v3(X[4], X[5], X[6]);
    V3 r2 = v3(X[8], X[9], X[10]);
    V3 c0 = cross(r1, r2);
    float s = dot(r0, c0) < 0.0f ? -1.0f : 1.0f;
    return s * v3(dot(c0, n), dot(cross(r2, r0), n), dot(cross(r0, r1), n));
}

__device__ __forceinline__ V3 transformNormal(const uint32_t* I, uint32_t base, V3 n)
{
    V3 r0 = v3(asf(I[base]), asf(I[base + 1]), asf(I[base + 2]));
    V3 r1 = v3(asf(I[base + 4]), asf(I[base + 5]), asf(I[base + 6]));
    V3 r2 = v3(asf(I[base + 8]), asf(I[base + 9]), asf(I[base + 10]));
    V3 c0 = cross(r1, r2);
    float s = dot(r0, c0) < 0.0f ? -1.0f : 1.0f;
    return s * v3(dot(c0, n), dot(cross(r2, r0), n), dot(cross(r0, r1), n));
}

__device__ __forceinline__ V3 cosineLocal(float u1, float u2)
{
    float r = sqrtf(u1);
    float phi = 2.0f * kPi * u2;
    return v3(r * vcos(phi), r * vsin(phi), sqrtf(fmaxf(0.0f, 1.0f - u1)));
}

__device__ __forceinline__ float G1(float cosT, float alpha)
{
    float c2 = cosT * cosT;
    float tan2 = (1.0f - c2) / fmaxf(c2, 1.0e-8f);
    return 2.0f / (1.0f + sqrtf(1.0f + alpha * alpha * tan2));
}

__device__ __forceinline__ float Dggx(float hz, float alpha)
{
    float a2 = alpha * alpha;
    float d = hz * hz * (a2 - 1.0f) + 1.0f;
    return a2 / (kPi * d * d);
}

__device__ __forceinline__ V3 schlick(V3 f0, float c)
{
    float m = 1.0f - saturate(c);
    float m2 = m * m;
    return f0 + (1.0f - f0) * (m2 * m2 * m);
}

__device__ __forceinline__ float fresnelDielectric(float cosi, float eta)
{
    float s2 = (1.0f - cosi * cosi) / (eta * eta);
    if (s2 >= 1.0f)
        return 1.0f;
    float cost = sqrtf(1.0f - s2);
    float rs = (cosi - eta * cost) / (cosi + eta * cost);
    float rp = (eta * cosi - cost) / (eta * cosi + cost);
    return 0.5f * (rs * rs + rp * rp);
}

__device__ __forceinline__ void basis(V3 n, V3& t, V3& b)
{
    V3 a = fabsf(n.x) > 0.9f ? v3(0.0f, 1.0f, 0.0f) : v3(1.0f, 0.0f, 0.0f);
    b = normalize(cross(n, a));
    t = cross(b, n);
}

__device__ __forceinline__ V3 sampleVndf(V3 wo, float alpha, float u1, float u2)
{
    V3 vh = normalize(v3(alpha * wo.x, alpha * wo.y, wo.z));
    float lensq = vh.x * vh.x + vh.y * vh.y;
    V3 t1 = lensq > 0.0f ? v3(-vh.y, vh.x, 0.0f) * (1.0f / sqrtf(lensq)) : v3(1.0f, 0.0f, 0.0f);
    V3 t2 = cross(vh, t1);
    float r = sqrtf(u1);
    float phi = 2.0f * kPi * u2;
    float a = r * vcos(phi);
    float b = r * vsin(phi);
    float s = 0.5f * (1.0f + vh.z);
    b = (1.0f - s) * sqrtf(1.0f - a * a) + s * b;
    V3 nh = a * t1 + b * t2 + sqrtf(fmaxf(0.0f, 1.0f - a * a - b * b)) * vh;
    return normalize(v3(alpha * nh.x, alpha * nh.y, fmaxf(0.0f, nh.z)));
}

__device__ __forceinline__ float albedoAt(const Ctx& c, uint32_t i) { return asf(c.S.lights[c.P.w[5] * 16 + i]); }
__device__ __forceinline__ uint32_t recvMask(const Ctx& c, uint32_t inst) { return c.S.lights[c.P.w[5] * 16 + kAlbedoWords + 2 * inst]; }
__device__ __forceinline__ uint32_t emitMask(const Ctx& c, uint32_t inst) { return c.S.lights[c.P.w[5] * 16 + kAlbedoWords + 2 * inst + 1]; }

__device__ float iesFactor(const Ctx& c, uint32_t po, V3 d)
{
    const uint32_t* L = c.S.lights;
    V3 l = v3(dot(d, v3(asf(L[po]), asf(L[po + 1]), asf(L[po + 2]))), dot(d, v3(asf(L[po + 3]), asf(L[po + 4]), asf(L[po + 5]))),
              dot(d, v3(asf(L[po + 6]), asf(L[po + 7]), asf(L[po + 8]))));
    float th = vacos(clampf(-l.z, -1.0f, 1.0f)) * (180.0f / kPi);
    float ph = vatan2(l.y, l.x) * (180.0f / kPi);
    if (ph < 0.0f)
        ph += 360.0f;
    uint32_t i = umin(f2u(th), 179u);
    uint32_t j = umin(f2u(ph), 359u);
    float ft = th - (float)i;
    float fp = ph - (float)j;
    uint32_t t = po + 9;
    float a0 = lerp(asf(L[t + j * 181 + i]), asf(L[t + j * 181 + i + 1]), ft);
    float a1 = lerp(asf(L[t + (j + 1) * 181 + i]), asf(L[t + (j + 1) * 181 + i + 1]), ft);
    return lerp(a0, a1, fp);
}

__device__ void addSlot(const Ctx& c, uint32_t pixel, uint32_t slot, V3 u, V3 s, V3 wi, float lit, float r)
{
    if (bad(u) || bad(s))
        return;
    float* L = c.F.lighting + (size_t)kLightingStride * pixel + 17 + kSlotWords * slot;
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

__device__ __forceinline__ V3 envTexel(const Ctx& c, uint32_t i, uint32_t j)
{
    uint32_t o = 3 * (j * c.P.w[24] + i);
    return v3(c.S.envPixels[o], c.S.envPixels[o + 1], c.S.envPixels[o + 2]);
}

__device__ __forceinline__ void envCoords(const Ctx& c, V3 d, uint32_t& i, uint32_t& j)
{
    uint32_t w = c.P.w[24], h = c.P.w[25];
    float u = vatan2(d.x, -d.z) * (0.5f / kPi) + 0.5f;
    float v = vacos(clampf(d.y, -1.0f, 1.0f)) / kPi;
    i = umin(f2u(u * (float)w), w - 1);
    j = umin(f2u(v * (float)h), h - 1);
}

__device__ __forceinline__ V3 envRadiance(const Ctx& c, V3 d)
{
    uint32_t i, j;
    envCoords(c, d, i, j);
    return envTexel(c, i, j);
}

__device__ __forceinline__ float envPdf(const Ctx& c, V3 d)
{
    uint32_t i, j;
    envCoords(c, d, i, j);
    return lum(envTexel(c, i, j)) * c.P.f(26);
}

// The environment's texel for draws (u1, u2): the row whose marginal CDF value first reaches u1,
// then the column whose CDF in that row first reaches u2, or the last one (the binary searches of
// pt_core.slang:274-306). Each search starts from the cutpoints of u's interval in
// the CDF's guide table (Chen and Asau's method, as emitterIndex), which hold the answer between
// them: the same texel in ~1-2 dependent loads per search instead of log2(h) and log2(w).
__device__ V3 envSample(const Ctx& c, float u1, float u2, float u3, float u4, float& pdf, V3& le)
{
    uint32_t w = c.P.w[24], h = c.P.w[25];
    const uint32_t bm = c.S.envGuideBits & 0xffu, bc = c.S.envGuideBits >> 8;
    const uint32_t km = umin((uint32_t)(u1 * (float)(1u << bm)), (1u << bm) - 1u);
    uint32_t lo = c.S.envGuide[km], hi = c.S.envGuide[km + 1];
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2;
        if (c.S.envMarginal[mid] >= u1)
            hi = mid;
        else
            lo = mid + 1;
    }
    uint32_t j = lo;
    const uint32_t* g = c.S.envGuide + (1u << bm) + 1u + j * ((1u << bc) + 1u);
    const uint32_t kc = umin((uint32_t)(u2 * (float)(1u << bc)), (1u << bc) - 1u);
    lo = g[kc];
    hi = g[kc + 1];
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2;
        if (c.S.envCond[j * w + mid] >= u2)
            hi = mid;
        else
            lo = mid + 1;
    }
    uint32_t i = lo;
    float phi = (((float)i + u3) / (float)w - 0.5f) * 2.0f * kPi;
    float theta = (((float)j + u4) / (float)h) * kPi;
    float st = vsin(theta);
    le = envTexel(c, i, j);
    pdf = lum(le) * c.P.f(26);
    return v3(st * vsin(phi), vcos(theta), -st * vcos(phi));
}

__device__ __forceinline__ V3 texelAt(const Ctx& c, uint32_t base, uint32_t w, uint32_t h, int x, int y)
{
    uint32_t xi = (uint32_t)(((x % (int)w) + (int)w) % (int)w);
    uint32_t yi = (uint32_t)(((y % (int)h) + (int)h) % (int)h);
    uint32_t o = 3 * (base + yi * w + xi);
    return v3(c.S.texels[o], c.S.texels[o + 1], c.S.texels[o + 2]);
}

__device__ V3 bilinear(const Ctx& c, uint32_t base, uint32_t w, uint32_t h, float ux, float uy)
{
    float fx = ux * (float)w - 0.5f;
    float fy = uy * (float)h - 0.5f;
    int x0 = (int)floorf(fx);
    int y0 = (int)floorf(fy);
    float tx = fx - (float)x0;
    float ty = fy - (float)y0;
    V3 a = lerp(texelAt(c, base, w, h, x0, y0), texelAt(c, base, w, h, x0 + 1, y0), tx);
    V3 b = lerp(texelAt(c, base, w, h, x0, y0 + 1), texelAt(c, base, w, h, x0 + 1, y0 + 1), tx);
    return lerp(a, b, ty);
}

__device__ V3 sampleTexture(const Ctx& c, uint32_t tex, float ux, float uy, float footprint)
{
    uint32_t ti = tex * 4;
    uint32_t base = c.S.texInfo[ti], w = c.S.texInfo[ti + 1], h = c.S.texInfo[ti + 2], levels = c.S.texInfo[ti + 3];
    float lod = clampf(vlog2(fmaxf(footprint * (float)umax(w, h), 1.0e-8f)), 0.0f, (float)(levels - 1));
    uint32_t l0 = f2u(floorf(lod));
    float t = lod - (float)l0;
    uint32_t off = base, lw = w, lh = h;
    for (uint32_t k = 0; k < l0; k++) {
        off += lw * lh;
        lw = umax(lw / 2, 1u);
        lh = umax(lh / 2, 1u);
    }
    V3 a = bilinear(c, off, lw, lh, ux, uy);
    if (l0 + 1 >= levels || t <= 0.0f)
        return a;
    off += lw * lh;
    V3 b = bilinear(c, off, umax(lw / 2, 1u), umax(lh / 2, 1u), ux, uy);
    return lerp(a, b, t);
}

__device__ float ggxAlbedo(const Ctx& c, float cosO, float alpha)
{
    float x = saturate(cosO) * (float)(kGgxN - 1);
    float y = saturate(sqrtf(alpha)) * (float)(kGgxN - 1);
    uint32_t x0 = umin(f2u(x), kGgxN - 2);
    uint32_t y0 = umin(f2u(y), kGgxN - 2);
    float fx = x - (float)x0, fy = y - (float)y0;
    float a = lerp(albedoAt(c, kAlbedoGgx + y0 * kGgxN + x0), albedoAt(c, kAlbedoGgx + y0 * kGgxN + x0 + 1), fx);
    float b = lerp(albedoAt(c, kAlbedoGgx + (y0 + 1) * kGgxN + x0), albedoAt(c, kAlbedoGgx + (y0 + 1) * kGgxN + x0 + 1), fx);
    return lerp(a, b, fy);
}

__device__ __forceinline__ V3 multiscatterFactor(const Ctx& c, const Surface& s, float cosO)
{
    if (c.P.w[29] == 0)
        return v3(1.0f);
    return 1.0f + s.albedo * (1.0f / fmaxf(ggxAlbedo(c, cosO, s.alpha), 0.05f) - 1.0f);
}

__device__ float dielectricAlbedo(const Ctx& c, float cosO, float rough, float ior, bool outside)
{
    float x = saturate(cosO) * (float)(kDielN - 1);
    float y = saturate(rough) * (float)(kDielN - 1);
    float z = clampf((ior - 1.0f) * 4.0f, 0.0f, (float)(kDielNIor - 1) - 1.0e-3f);
    uint32_t x0 = umin(f2u(x), kDielN - 2);
    uint32_t y0 = umin(f2u(y), kDielN - 2);
    uint32_t z0 = f2u(z);
    float fx = x - (float)x0, fy = y - (float)y0, fz = z - (float)z0;
    uint32_t base = (outside ? 0u : kDielNIor) * kDielN * kDielN;
    float cc[2];
#pragma unroll
    for (uint32_t k = 0; k < 2; k++) {
        uint32_t o = base + ((z0 + k) * kDielN + y0) * kDielN + x0;
        float a = lerp(albedoAt(c, kAlbedoDiel + o), albedoAt(c, kAlbedoDiel + o + 1), fx);
        float b = lerp(albedoAt(c, kAlbedoDiel + o + kDielN), albedoAt(c, kAlbedoDiel + o + kDielN + 1), fx);
        cc[k] = lerp(a, b, fy);
    }
    return lerp(cc[0], cc[1], fz);
}

__device__ float sheenAlbedo(const Ctx& c, float cosO, float r)
{
    float x = saturate(cosO) * (float)(kSheenN - 1);
    float y = saturate(r) * (float)(kSheenN - 1);
    uint32_t x0 = umin(f2u(x), kSheenN - 2);
    uint32_t y0 = umin(f2u(y), kSheenN - 2);
    float fx = x - (float)x0, fy = y - (float)y0;
    float a = lerp(albedoAt(c, kAlbedoSheen + y0 * kSheenN + x0), albedoAt(c, kAlbedoSheen + y0 * kSheenN + x0 + 1), fx);
    float b = lerp(albedoAt(c, kAlbedoSheen + (y0 + 1) * kSheenN + x0), albedoAt(c, kAlbedoSheen + (y0 + 1) * kSheenN + x0 + 1), fx);
    return lerp(a, b, fy);
}

__device__ __forceinline__ float sheenCore(float woz, float wiz, float hz, float alpha)
{
    float sinH = sqrtf(fmaxf(1.0f - hz * hz, 0.0f));
    float d = (2.0f + 1.0f / alpha) * vpow(sinH, 1.0f / alpha) / (2.0f * kPi);
    float v = 1.0f / (4.0f * (wiz + woz - wiz * woz));
    return d * v * wiz;
}

// t fresnelDielectric(woh, eta) (times tComp in the lobe values): skipped for an opaque surface
// (t = 0) when eta is positive and finite, where the product is exactly 0 (the Fresnel term is then
// finite and not negative); computed otherwise, so a degenerate eta gives the same bits as before
__device__ __forceinline__ bool transmissionTerm(const Surface& s)
{
    return s.transmission != 0.0f || !(s.eta > 0.0f && s.eta <= 3.40282347e38f);
}

__device__ Lobes evalLobes(const Ctx& c, const Surface& s, V3 woL, V3 wiL)
{
    V3 hh = normalize(woL + wiL);
    float woh = fmaxf(dot(woL, hh), 0.0f);
    float core = Dggx(hh.z, s.alpha) * s.g1o * G1(wiL.z, s.alpha) / (4.0f * woL.z);
    float t = s.transmission;
    Lobes l;
    const float ft = transmissionTerm(s) ? t * fresnelDielectric(woh, s.eta) * s.tComp : 0.0f;
    l.specular = s.metallic * schlick(s.albedo, woh) * s.ms * core + (1.0f - s.metallic) * (ft + (1.0f - t) * fresnelDielectric(woh, s.ior)) * core;
    l.diffuse = (1.0f - s.metallic) * (1.0f - t) * (1.0f - s.fSel) * s.albedo * (wiL.z / kPi);
    if (s.sheenP > 0.0f) {
        l.specular = (1.0f - s.sheenP) * l.specular + s.sheenColor * sheenCore(woL.z, wiL.z, hh.z, s.sheenAlpha);
        l.diffuse = l.diffuse * (1.0f - s.sheenP);
    }
    if (s.coat > 0.0f) {
        float coreC = Dggx(hh.z, s.coatAlpha) * s.g1oC * G1(wiL.z, s.coatAlpha) / (4.0f * woL.z);
        l.specular = (1.0f - s.coatP) * l.specular + s.coat * fresnelDielectric(woh, 1.5f) * coreC;
        l.diffuse = l.diffuse * (1.0f - s.coatP);
    }
    return l;
}

__device__ __forceinline__ V3 evalReflective(const Ctx& c, const Surface& s, V3 woL, V3 wiL)
{
    Lobes l = evalLobes(c, s, woL, wiL);
    return l.specular + l.diffuse;
}

__device__ __forceinline__ void partsFile(Parts& p, uint32_t depth, uint32_t cls, V3 c)
{
    if (depth == 0)
        p.emission = p.emission + c;
    else if (cls == 0)
        p.indirectDiffuse = p.indirectDiffuse + c;
    else if (cls == 1)
        p.specular = p.specular + c;
    else
        p.transmission = p.transmission + c;
}

__device__ __forceinline__ void partsAdd(Parts& p, uint32_t depth, uint32_t cls, V3 c, bool split)
{
    p.total = p.total + c;
    if (split)
        partsFile(p, depth, cls, c);
}

__device__ __forceinline__ void partsAddLight(Parts& p, uint32_t depth, uint32_t cls, V3 c, V3 cd, V3 cs, bool split)
{
    p.total = p.total + c;
    if (!split)
        return;
    if (depth == 0) {
        p.directDiffuse = p.directDiffuse + cd;
        p.specular = p.specular + cs;
    } else {
        partsFile(p, depth, cls, cd + cs);
    }
}

__device__ float reflectivePdf(const Surface& s, V3 woL, V3 wiL)
{
    V3 hh = normalize(woL + wiL);
    float woh = fmaxf(dot(woL, hh), 0.0f);
    float spec = s.g1o * Dggx(hh.z, s.alpha) / (4.0f * woL.z);
    float t = s.transmission;
    const float ft = transmissionTerm(s) ? t * fresnelDielectric(woh, s.eta) : 0.0f;
    float pSpec = s.metallic + (1.0f - s.metallic) * (ft + (1.0f - t) * s.fSel);
    float pDiff = (1.0f - s.metallic) * (1.0f - t) * (1.0f - s.fSel) * wiL.z / kPi;
    float q = pSpec * spec + pDiff;
    if (s.sheenP > 0.0f)
        q = (1.0f - s.sheenP) * q + s.sheenP * wiL.z / kPi;
    if (s.coat > 0.0f) {
        float specC = s.g1oC * Dggx(hh.z, s.coatAlpha) / (4.0f * woL.z);
        q = (1.0f - s.coatP) * q + s.coatP * specC;
    }
    return q;
}

__device__ float opacityAt(const Ctx& c, uint32_t instIdx, uint32_t prim, float b1, float b2)
{
    const Scene& S = c.S;
    uint32_t mi = S.instances[instIdx * 16 + 2] * 24;
    float op = S.materials[mi + 22];
    uint32_t tex = f2u(S.materials[mi + 23]);
    if (tex > 0 && op > 0.0f) {
        uint32_t tri = S.instances[instIdx * 16 + 1] + 3 * prim;
        uint32_t vo = S.instances[instIdx * 16];
        uint32_t i0 = S.indices[tri] + vo, i1 = S.indices[tri + 1] + vo, i2 = S.indices[tri + 2] + vo;
        float u0x = S.uvs[2 * i0], u0y = S.uvs[2 * i0 + 1];
        float u1x = S.uvs[2 * i1], u1y = S.uvs[2 * i1 + 1];
        float u2x = S.uvs[2 * i2], u2y = S.uvs[2 * i2 + 1];
        float ux = u0x + b1 * (u1x - u0x) + b2 * (u2x - u0x);
        float uy = u0y + b1 * (u1y - u0y) + b2 * (u2y - u0y);
        op *= sampleTexture(c, tex - 1, ux, uy, 0.0f).x;
    }
    return op;
}

__device__ bool occluded(const Ctx& c, V3 o, V3 d, float tmax)
{
    for (uint32_t layer = 0; layer < 8; layer++) {
        Hit q = c.S.translucent ? trace<false>(c.S.bvh, o, d, tmax) : trace<true>(c.S.bvh, o, d, tmax);
        if (!q.hit)
            return false;
        if (!c.S.translucent)
            return true;
        float opacity = opacityAt(c, q.inst, q.prim, q.b1, q.b2);
        if (opacity >= 1.0f)
            return true;
        uint32_t h = asu(o.x) * 73856093u ^ asu(o.y) * 19349663u ^ asu(o.z) * 83492791u ^ asu(d.x) * 2654435761u ^
                     asu(d.y) * 40503u ^ asu(d.z) * 2246822519u ^ layer * 3266489917u;
        h ^= h >> 16;
        h *= 0x7feb352du;
        h ^= h >> 15;
        h *= 0x846ca68bu;
        h ^= h >> 16;
        if ((float)(h >> 8) * (1.0f / 16777216.0f) < opacity)
            return true;
        float adv = q.t + 1.0e-4f;
        o = o + d * adv;
        tmax -= adv;
        if (tmax <= 0.0f)
            return false;
    }
    return true;
}

// generatePathFor (pt_core.slang:645-682); the seed mix is XORed into the pixel hash (0 = the reference stream)
__device__ void generatePath(const Params& P, uint32_t px, uint32_t py, uint32_t sample, PathState& s)
{
    uint32_t w = P.w[0], h = P.w[1];
    uint32_t pixel = py * w + px;
    uint32_t rng = pcg(pixel ^ pcg(sample * 0x9E3779B9u + 0x7F4A7C15u) ^ P.seedMix);
    float t = P.f(2);
    float aspect = (float)w / (float)h;
    float u = (2.0f * ((float)px + rnext(rng)) / (float)w - 1.0f + P.f(39)) * aspect * t;
    float v = (1.0f - 2.0f * ((float)py + rnext(rng)) / (float)h + P.f(40)) * t;
    V3 focal = v3(u, v, -1.0f) * P.f(4);
    float lensR = P.f(3) * sqrtf(rnext(rng));
    float lensA = 2.0f * kPi * rnext(rng);
    V3 lens = v3(lensR * vcos(lensA), lensR * vsin(lensA), 0.0f);
    V3 dc = normalize(focal - lens);
    float org[3], dir[3];
#pragma unroll
    for (uint32_t r = 0; r < 3; r++) {
        org[r] = P.f(8 + 4 * r) * lens.x + P.f(9 + 4 * r) * lens.y + P.f(10 + 4 * r) * lens.z + P.f(11 + 4 * r);
        dir[r] = P.f(8 + 4 * r) * dc.x + P.f(9 + 4 * r) * dc.y + P.f(10 + 4 * r) * dc.z;
    }
    s.org = v3(org[0], org[1], org[2]);
    s.dir = normalize(v3(dir[0], dir[1], dir[2]));
    s.thr = v3(1.0f);
    s.rng = rng;
    s.cls = 0;
    s.chan = -1;
    s.pathLen = 0.0f;
    s.misPrev = false;
    s.restir = false;
    s.gi = false;
    s.qPrev = 0.0f;
    s.prevInst = 0;
    s.visLit = 0.0f;
    s.visTotal = 0.0f;
    s.parts = Parts{v3(0.0f), v3(0.0f), v3(0.0f), v3(0.0f), v3(0.0f), v3(0.0f)};
}

// ---------------------------------------------------------------- ReSTIR DI (pt_core.slang:702-995)

// The emitter a power-CDF draw u in [0, 1) selects: the first index whose CDF value reaches u, or
// the last emitter (the binary search of pt_core.slang:1343-1351 and :744-753). The guide table
// (Chen and Asau's cutpoint method) holds, for each of the 2^bits intervals [k, k + 1) / 2^bits,
// the first index whose CDF reaches k / 2^bits, so the answer for u lies between the cutpoints of
// u's interval and the same binary search runs over that range only: the same index in ~2 steps
// instead of log2(n). u * 2^bits is exact (u is a multiple of 2^-24), so the interval is exact.
__device__ __forceinline__ uint32_t emitterIndex(const Scene& S, float u)
{
    uint32_t k = (uint32_t)(u * (float)(1u << S.emitGuideBits));
    k = umin(k, (1u << S.emitGuideBits) - 1u);
    uint32_t lo = S.emitGuide[k], hi = S.emitGuide[k + 1];
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2;
        if (S.emitCdf[mid] >= u)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

// Random neighbours per reuse: bits 8-15 of word 36 (the ReSTIR mode word); 0 = the reference's 2.
__device__ __forceinline__ uint32_t restirNeighbours(const Params& P)
{
    uint32_t k = (P.w[36] >> 8) & 0xffu;
    return k ? k : kRestirNeighbours;
}

__device__ __forceinline__ uint32_t restirBase(uint32_t pixel, uint32_t which) { return pixel * kRestirStride + which * kRestirHalf; }

__device__ bool restirLight(const Ctx& c, uint32_t e, uint32_t recv, V3 x, V3 p, V3 n, V3& wi, float& dist, float& cosi, V3& g, float& plsa)
{
    const float* E = c.S.emitters + (size_t)e * 16;
    V3 a = v3(E[0], E[1], E[2]), b = v3(E[3], E[4], E[5]), cc = v3(E[6], E[7], E[8]);
    V3 toL = x - p;
    float d2 = dot(toL, toL);
    dist = sqrtf(d2);
    wi = toL / fmaxf(dist, 1.0e-6f);
    float cosl = fabsf(dot(normalize(cross(b - a, cc - a)), wi));
    cosi = dot(n, wi);
    g = v3(E[9], E[10], E[11]) * (cosl / fmaxf(d2, 1.0e-12f));
    plsa = E[13] * d2 / (fmaxf(E[12], 1.0e-30f) * fmaxf(cosl, 1.0e-5f));
    return cosi > 0.0f && cosl > 1.0e-5f && d2 > 1.0e-12f && (asu(E[14]) & recv) != 0;
}

__device__ __forceinline__ float restirMis(const Surface& surf, V3 woL, V3 wiL, float plsa)
{
    float qb = reflectivePdf(surf, woL, wiL);
    return plsa * plsa / fmaxf(plsa * plsa + qb * qb, 1.0e-30f);
}

struct RestirPick {
    bool valid;
    uint32_t e;
    V3 x;
    float W;
    bool occ;
};

// DEFER: the winner's shadow ray is left to the caller's shadow record (stored W as if visible;
// k_wf_resolve zeroes it when the ray is blocked).
template <bool DEFER = false>
__device__ RestirPick restirDI(const Ctx& c, uint32_t pixel, uint32_t recv, V3 p, V3 n, float pHitT, const Surface& surf, V3 T, V3 B,
                               V3 woL, uint32_t& rng)
{
    const Params& P = c.P;
    uint32_t w = P.w[0];
    uint32_t cur = P.w[6] & 1u;
    float cap = (float)umax(P.w[38], 1u);
    float wsum = 0.0f, mTotal = 0.0f;
    uint32_t selE = 0;
    V3 selX = v3(0.0f);
    float selPhat = 0.0f;
    uint32_t ncand = umax(P.w[37], 1u);
    for (uint32_t k = 0; k < ncand; k++) {
        float ue = rnext(rng), ua = rnext(rng), ub = rnext(rng);
        uint32_t lo = emitterIndex(c.S, ue);
        const float* E = c.S.emitters + (size_t)lo * 16;
        V3 a = v3(E[0], E[1], E[2]), b = v3(E[3], E[4], E[5]), cc = v3(E[6], E[7], E[8]);
        float su = sqrtf(ua);
        V3 x = a * (1.0f - su) + b * (su * (1.0f - ub)) + cc * (su * ub);
        float area = E[12];
        float q = area > 0.0f ? E[13] / area : 0.0f;
        V3 wi, g;
        float dist, cosi, plsa;
        float phat = 0.0f;
        if (q > 0.0f && restirLight(c, lo, recv, x, p, n, wi, dist, cosi, g, plsa)) {
            V3 wiL = v3(dot(wi, T), dot(wi, B), cosi);
            phat = lum(evalReflective(c, surf, woL, wiL) * g) * restirMis(surf, woL, wiL, plsa);
        }
        float wgt = phat / fmaxf(q, 1.0e-30f);
        mTotal += 1.0f;
        if (phat > 0.0f) {
            wsum += wgt;
            if (rnext(rng) * wsum < wgt) {
                selE = lo + 1;
                selX = x;
                selPhat = phat;
            }
        }
    }
    if (P.w[6] > 0) {
        int px = (int)(pixel % w), py = (int)(pixel / w);
        for (uint32_t k = 0; k < 1 + restirNeighbours(P); k++) {
            int nx = px, ny = py;
            if (k > 0) {
                nx += (int)(rnext(rng) * (float)(2 * kRestirRadius + 1)) - kRestirRadius;
                ny += (int)(rnext(rng) * (float)(2 * kRestirRadius + 1)) - kRestirRadius;
                if (nx < 0 || ny < 0 || nx >= (int)w || ny >= (int)P.w[1])
                    continue;
            }
            const float* R = c.F.restir + restirBase((uint32_t)ny * w + (uint32_t)nx, cur ^ 1u);
            uint32_t ei = asu(R[0]);
            if (asu(R[14]) == 0 || ei == 0)
                continue;
            V3 pn = v3(R[11], R[12], R[13]);
            V3 pp = v3(R[8], R[9], R[10]);
            if (dot(pn, n) < 0.9f || length(pp - p) > 0.05f * fmaxf(pHitT, 0.01f))
                continue;
            V3 xc = v3(R[1], R[2], R[3]);
            float mc = fminf(R[5], cap);
            V3 wi, g;
            float dist, cosi, plsa;
            float phat = 0.0f;
            if (restirLight(c, ei - 1, recv, xc, p, n, wi, dist, cosi, g, plsa)) {
                V3 wiL = v3(dot(wi, T), dot(wi, B), cosi);
                phat = lum(evalReflective(c, surf, woL, wiL) * g) * restirMis(surf, woL, wiL, plsa);
            }
            float wgt = phat * R[4] * mc;
            mTotal += mc;
            if (wgt > 0.0f) {
                wsum += wgt;
                if (rnext(rng) * wsum < wgt) {
                    selE = ei;
                    selX = xc;
                    selPhat = phat;
                }
            }
        }
    }
    RestirPick r;
    r.valid = selE != 0 && selPhat > 0.0f;
    r.e = selE;
    r.x = selX;
    r.W = r.valid ? wsum / (mTotal * selPhat) : 0.0f;
    r.occ = false;
    if (r.valid && !DEFER) {
        V3 wi, g;
        float dist, cosi, plsa;
        restirLight(c, selE - 1, recv, selX, p, n, wi, dist, cosi, g, plsa);
        r.occ = occluded(c, p, wi, dist * (1.0f - 1.0e-3f));
    }
    float* C = c.F.restir + restirBase(pixel, cur);
    C[0] = asf(selE);
    C[1] = selX.x;
    C[2] = selX.y;
    C[3] = selX.z;
    C[4] = r.occ ? 0.0f : r.W;
    C[5] = mTotal;
    return r;
}

__device__ __forceinline__ void restirGbuffer(const Ctx& c, uint32_t pixel, V3 p, V3 n)
{
    float* C = c.F.restir + restirBase(pixel, c.P.w[6] & 1u);
    C[8] = p.x;
    C[9] = p.y;
    C[10] = p.z;
    C[11] = n.x;
    C[12] = n.y;
    C[13] = n.z;
    C[14] = asf(1u);
}

__device__ __forceinline__ void restirClear(const Ctx& c, uint32_t pixel)
{
    float* C = c.F.restir + restirBase(pixel, c.P.w[6] & 1u);
    C[14] = asf(0u);
    C[27] = asf(0u);
}

// ---------------------------------------------------------------- ReSTIR GI (pt_core.slang:901-980, 1566-1617)
//
// Per pixel, each frame half holds after the DI words a second-hit reservoir: 16..18 the point x2,
// 19..21 its facing normal, 22..24 the radiance it sends back (per unit throughput), 25 W, 26 M,
// 27 valid. The 24 scratch words after the two halves carry the frame's resampled pick from the
// first hit to the end of the path: 0 wsum, 1 M, 2 pSel, 3..5 its contribution, 6..8 x2,
// 9..11 n2, 12..14 radiance, 15..17 the first bounce's throughput, 18..20 what the first hit
// gathered itself, 21 the first bounce's pdf, 22 "the first bounce reached a reusable surface".

__device__ __forceinline__ uint32_t restirScratch(uint32_t pixel) { return pixel * kRestirStride + 2 * kRestirHalf; }

// restirGIPrepare (pt_core.slang:901-972): the previous frame's reservoirs of this pixel and two
// random neighbours, reconnected to this shading point; a neighbour's sample is weighed by the
// change of solid angle (cosR / cosQ) (dQ / dR)^2, which must stay within [1/10, 10], and must be
// visible from the point.
__device__ void restirGIPrepare(const Ctx& c, uint32_t pixel, V3 p, V3 n, float pHitT, const Surface& surf, V3 T, V3 B, V3 woL, uint32_t& rng)
{
    const Params& P = c.P;
    uint32_t w = P.w[0];
    uint32_t cur = P.w[6] & 1u;
    float cap = (float)max(P.w[38], 1u);
    float wsum = 0.0f, mTotal = 0.0f, pSel = 0.0f;
    V3 cSel = v3(0.0f), xSel = v3(0.0f), nSel = v3(0.0f), lSel = v3(0.0f);
    if (P.w[6] > 0) {
        int px = (int)(pixel % w), py = (int)(pixel / w);
        for (uint32_t k = 0; k < 1 + restirNeighbours(P); k++) {
            int nx = px, ny = py;
            if (k > 0) {
                nx += (int)(rnext(rng) * (float)(2 * kRestirRadius + 1)) - kRestirRadius;
                ny += (int)(rnext(rng) * (float)(2 * kRestirRadius + 1)) - kRestirRadius;
                if (nx < 0 || ny < 0 || nx >= (int)w || ny >= (int)P.w[1])
                    continue;
            }
            const float* R = c.F.restir + restirBase((uint32_t)ny * w + (uint32_t)nx, cur ^ 1u);
            if (asu(R[14]) == 0 || asu(R[27]) == 0)
                continue;
            V3 pq = v3(R[8], R[9], R[10]);
            V3 nq = v3(R[11], R[12], R[13]);
            if (dot(nq, n) < 0.9f || length(pq - p) > 0.05f * fmaxf(pHitT, 0.01f))
                continue;
            V3 x2 = v3(R[16], R[17], R[18]);
            V3 n2 = v3(R[19], R[20], R[21]);
            V3 l2 = v3(R[22], R[23], R[24]);
            float mq = fminf(R[26], cap);
            mTotal += mq;
            V3 toX = x2 - p;
            float dR = length(toX);
            V3 wi = toX / fmaxf(dR, 1.0e-6f);
            float dQ = length(x2 - pq);
            float cosT = dot(n, wi);
            float cosR = fabsf(dot(n2, wi));
            float cosQ = fabsf(dot(n2, normalize(pq - x2)));
            if (cosT <= 0.0f || dR < 1.0e-3f || cosR < 1.0e-3f || cosQ < 1.0e-3f)
                continue;
            float jac = (cosR / cosQ) * (dQ * dQ) / (dR * dR);
            if (jac > kRestirMaxJacobian || jac < 1.0f / kRestirMaxJacobian)
                continue;
            V3 cc = evalReflective(c, surf, woL, v3(dot(wi, T), dot(wi, B), cosT)) * l2;
            float phat = lum(cc);
            if (phat <= 0.0f || occluded(c, p, wi, dR * (1.0f - 1.0e-3f)))
                continue;
            float wgt = phat * R[25] * mq * jac;
            wsum += wgt;
            if (rnext(rng) * wsum < wgt) {
                pSel = phat;
                cSel = cc;
                xSel = x2;
                nSel = n2;
                lSel = l2;
            }
        }
    }
    float* Sc = c.F.restir + restirScratch(pixel);
    Sc[0] = wsum;
    Sc[1] = mTotal;
    Sc[2] = pSel;
    Sc[3] = cSel.x;
    Sc[4] = cSel.y;
    Sc[5] = cSel.z;
    Sc[6] = xSel.x;
    Sc[7] = xSel.y;
    Sc[8] = xSel.z;
    Sc[9] = nSel.x;
    Sc[10] = nSel.y;
    Sc[11] = nSel.z;
    Sc[12] = lSel.x;
    Sc[13] = lSel.y;
    Sc[14] = lSel.z;
    Sc[22] = asf(0u);
}

// What restirGIPrepare reads at the first hit, captured by shadePath's probe mode (GIP = 1).
struct GiSite {
    bool active; // the first hit takes ReSTIR GI
    V3 p, n, woL, T, B;
    float pHitT;
    Surface surf;
    uint32_t rng;
};

// restirGIPrepare for the wave's 64 first hits together: the same candidates, weights and random
// draws per lane, with each candidate's visibility ray traced by the wave-cooperative traversal
// (every lane of the wave calls this, `site.active` or not). The lane's random state after the
// reuse goes to scratch word 23, where shadePath's prepared mode (GIP = 2) picks it up.
__device__ void restirGIPrepareWave(const Ctx& c, const GiSite& site, uint32_t pixel, WaveTris& sh)
{
    const Params& P = c.P;
    uint32_t w = P.w[0];
    uint32_t cur = P.w[6] & 1u;
    float cap = (float)max(P.w[38], 1u);
    float wsum = 0.0f, mTotal = 0.0f, pSel = 0.0f;
    V3 cSel = v3(0.0f), xSel = v3(0.0f), nSel = v3(0.0f), lSel = v3(0.0f);
    uint32_t rng = site.rng;
    const bool act = site.active;
    const V3 p = site.p, n = site.n;
    if (P.w[6] > 0) {
        int px = (int)(pixel % w), py = (int)(pixel / w);
        for (uint32_t k = 0; k < 1 + restirNeighbours(P); k++) {
            bool need = false;
            V3 wi = v3(0.0f, 0.0f, 1.0f), cc = v3(0.0f), x2 = v3(0.0f), n2 = v3(0.0f), l2 = v3(0.0f);
            float dR = 0.0f, phat = 0.0f, rW = 0.0f, mq = 0.0f, jac = 0.0f;
            do {
                if (!act)
                    break;
                int nx = px, ny = py;
                if (k > 0) {
                    nx += (int)(rnext(rng) * (float)(2 * kRestirRadius + 1)) - kRestirRadius;
                    ny += (int)(rnext(rng) * (float)(2 * kRestirRadius + 1)) - kRestirRadius;
                    if (nx < 0 || ny < 0 || nx >= (int)w || ny >= (int)P.w[1])
                        break;
                }
                const float* R = c.F.restir + restirBase((uint32_t)ny * w + (uint32_t)nx, cur ^ 1u);
                if (asu(R[14]) == 0 || asu(R[27]) == 0)
                    break;
                V3 pq = v3(R[8], R[9], R[10]);
                V3 nq = v3(R[11], R[12], R[13]);
                if (dot(nq, n) < 0.9f || length(pq - p) > 0.05f * fmaxf(site.pHitT, 0.01f))
                    break;
                x2 = v3(R[16], R[17], R[18]);
                n2 = v3(R[19], R[20], R[21]);
                l2 = v3(R[22], R[23], R[24]);
                mq = fminf(R[26], cap);
                mTotal += mq;
                V3 toX = x2 - p;
                dR = length(toX);
                wi = toX / fmaxf(dR, 1.0e-6f);
                float dQ = length(x2 - pq);
                float cosT = dot(n, wi);
                float cosR = fabsf(dot(n2, wi));
                float cosQ = fabsf(dot(n2, normalize(pq - x2)));
                if (cosT <= 0.0f || dR < 1.0e-3f || cosR < 1.0e-3f || cosQ < 1.0e-3f)
                    break;
                jac = (cosR / cosQ) * (dQ * dQ) / (dR * dR);
                if (jac > kRestirMaxJacobian || jac < 1.0f / kRestirMaxJacobian)
                    break;
                cc = evalReflective(c, site.surf, site.woL, v3(dot(wi, site.T), dot(wi, site.B), cosT)) * l2;
                phat = lum(cc);
                rW = R[25];
                need = phat > 0.0f;
            } while (false);
            bool blocked;
            if (c.S.translucent) {
                blocked = need && occluded(c, p, wi, dR * (1.0f - 1.0e-3f));
            } else {
                Hit h = traceWave<true>(c.S.bvh, need ? p : v3(0.0f), wi, dR * (1.0f - 1.0e-3f), need, sh);
                blocked = need && h.hit;
            }
            if (need && !blocked) {
                float wgt = phat * rW * mq * jac;
                wsum += wgt;
                if (rnext(rng) * wsum < wgt) {
                    pSel = phat;
                    cSel = cc;
                    xSel = x2;
                    nSel = n2;
                    lSel = l2;
                }
            }
        }
    }
    if (!act)
        return;
    float* Sc = c.F.restir + restirScratch(pixel);
    Sc[0] = wsum;
    Sc[1] = mTotal;
    Sc[2] = pSel;
    Sc[3] = cSel.x;
    Sc[4] = cSel.y;
    Sc[5] = cSel.z;
    Sc[6] = xSel.x;
    Sc[7] = xSel.y;
    Sc[8] = xSel.z;
    Sc[9] = nSel.x;
    Sc[10] = nSel.y;
    Sc[11] = nSel.z;
    Sc[12] = lSel.x;
    Sc[13] = lSel.y;
    Sc[14] = lSel.z;
    Sc[22] = asf(0u);
    Sc[23] = asf(rng);
}

// giFinish (pt_core.slang:1566-1617): the path's own candidate (the surface its first bounce
// reached, with what the rest of the path gathered divided by the first bounce's throughput) joins
// the resampled ones; the pixel gets what the first hit gathered itself plus the pick's indirect
// light times W = wsum / (M pSel), and the new reservoir goes to this frame's half. The
// candidate's random number hashes the pixel with the sample index and the render's seed mix
// (zero for the reference's own stream).
__device__ V3 giFinish(const Ctx& c, V3 pathTotal, uint32_t pixel)
{
    float* C = c.F.restir + restirBase(pixel, c.P.w[6] & 1u);
    const float* Sc = c.F.restir + restirScratch(pixel);
    float wsum = Sc[0];
    float m = Sc[1] + 1.0f;
    float pSel = Sc[2];
    V3 cSel = v3(Sc[3], Sc[4], Sc[5]);
    V3 xSel = v3(Sc[6], Sc[7], Sc[8]);
    V3 nSel = v3(Sc[9], Sc[10], Sc[11]);
    V3 lSel = v3(Sc[12], Sc[13], Sc[14]);
    V3 thr1 = v3(Sc[15], Sc[16], Sc[17]);
    V3 d0 = v3(Sc[18], Sc[19], Sc[20]);
    float qPrev = Sc[21];
    bool own = asu(Sc[22]) != 0 && qPrev > 0.0f;
    V3 giTotal = pathTotal - d0;
    V3 result = d0;
    if (own) {
        V3 cOwn = giTotal * qPrev;
        float pOwn = lum(cOwn);
        if (pOwn > 0.0f) {
            float wOwn = pOwn / qPrev;
            wsum += wOwn;
            uint32_t r = pcg(pixel ^ pcg(c.P.w[6] * 0x85EBCA6Bu + 0x1234u) ^ c.P.seedMix);
            if (rnext(r) * wsum < wOwn) {
                pSel = pOwn;
                cSel = cOwn;
                xSel = v3(C[16], C[17], C[18]);
                nSel = v3(C[19], C[20], C[21]);
                lSel = v3(thr1.x > 1.0e-6f ? giTotal.x / thr1.x : 0.0f, thr1.y > 1.0e-6f ? giTotal.y / thr1.y : 0.0f,
                          thr1.z > 1.0e-6f ? giTotal.z / thr1.z : 0.0f);
            }
        }
    } else {
        result = result + giTotal;
    }
    float W = pSel > 0.0f ? wsum / (m * pSel) : 0.0f;
    result = result + cSel * W;
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

// ---------------------------------------------------------------- neural radiance cache hook

// What shadePath exchanges with the radiance cache. NRC = 1 (training paths): every opaque vertex
// records, after its own emission is added, the cache's inputs (position, outgoing direction,
// shading normal, roughness) with the path throughput and radiance total at that point; the
// radiance the rest of the path adds, divided by that throughput, is the vertex's training target.
// NRC = 2 (render paths): the first opaque vertex at `terminateDepth` hands the same inputs, the
// throughput and the total to the cache (a query record) and the path ends there.
constexpr uint32_t kNrcVertexWords = 5; // float4 per recorded vertex
constexpr uint32_t kNrcQueryWords = 5;  // float4 per query
struct NrcIo {
    float4* vertices; // NRC 1: [path][kMaxDepthNrc][kNrcVertexWords]
    float4* queries;  // NRC 2: [slot][kNrcQueryWords]
    uint32_t path;    // NRC 1: training path id
    uint32_t slot;    // NRC 2: the path's queue slot at this depth
    uint32_t terminateDepth;
    bool terminated;  // out (NRC 2): the path was handed to the cache
};
constexpr uint32_t kMaxDepthNrc = 8;

// Deferred shadow rays (plain wavefront, ReSTIR DI). With DEFER, shadePath does not trace its NEE
// shadow rays (nor the ReSTIR DI winner's): it records each one with the contribution it adds to the path total when unoccluded, in
// the order the fused code adds them; a later kernel traces the records and adds the visible
// contributions in that order, so the path total is the same bit for bit.
// Record k of slot i, word w (origin xyz, direction xyz, tmax, contribution rgb) lives at
// rec[(w * K + k) * n + i].
constexpr uint32_t kMaxShadow = 8;  // records per path at most (delta lights + one emitter/env sample + the ReSTIR DI winner)
constexpr uint32_t kShadowWords = 10;
constexpr uint32_t kPendingGI = 0x80000000u;    // pending pixel id bit: finish through giFinish
constexpr uint32_t kPendingFirst = 0x40000000u; // pending pixel id bit: a GI first hit's own light
struct ShadowSink {
    float* rec;
    uint32_t n, K, slot;
    uint32_t count;
    uint32_t restirRec = 0; // 1 + the record of the ReSTIR DI winner, 0 when none
    __device__ __forceinline__ void push(V3 o, V3 d, float tmax, V3 c)
    {
        if (count >= K) // never: K covers every delta light plus one sample (checked on the host)
            return;
        const size_t stride = (size_t)K * n;
        float* r = rec + (size_t)count * n + slot;
        const float v[kShadowWords] = {o.x, o.y, o.z, d.x, d.y, d.z, tmax, c.x, c.y, c.z};
#pragma unroll
        for (uint32_t w = 0; w < kShadowWords; w++)
            r[w * stride] = v[w];
        count++;
    }
};

// ---------------------------------------------------------------- shadePath (pt_core.slang:992-1543)

// GIP (ReSTIR GI's reuse at the first hit): 0 traces it inline (restirGIPrepare); 1 is the probe
// of k_gi_prepare, which stops at the reuse and hands its inputs to `site` (no delta lights, no
// writes); 2 takes the reuse k_gi_prepare made (restirGIPrepareWave), i.e. its scratch words and
// the random state after it.
template <bool RESTIR, bool SPLIT, int NRC = 0, bool DEFER = false, int GIP = 0>
__device__ bool shadePath(const Ctx& c, PathState& st, const Hit& hq, uint32_t depth, uint32_t pixel, NrcIo* io = nullptr,
                          ShadowSink* sink = nullptr, GiSite* site = nullptr)
{
    static_assert(!DEFER || (!SPLIT && NRC != 2), "deferred shadow rays are for the plain integrator, ReSTIR DI and training paths");
    const Scene& S = c.S;
    const Params& P = c.P;
    V3 sky = v3(P.f(20), P.f(21), P.f(22));
    const bool lightingOn = SPLIT;
    uint32_t lightCount = P.w[5];
    uint32_t emitterCount = P.w[23];
    bool hasEnv = P.w[24] > 0;
    float pEnv = hasEnv ? (emitterCount > 0 ? 0.5f : 1.0f) : 0.0f;
    uint32_t rng = st.rng;
    V3 org = st.org, dir = st.dir, thr = st.thr;
    uint32_t cls = st.cls;
    int chan = st.chan;
    bool alive = true;
    do {
        if (!hq.hit) {
            if (GIP == 1)
                return false;
            if (RESTIR && depth == 0)
                restirClear(c, pixel);
            if (depth == 0 && P.w[42] != 0) {
                alive = false;
                break;
            }
            partsAdd(st.parts, depth, cls, thr * sky, SPLIT);
            if (hasEnv) {
                float wB = 1.0f;
                if (st.misPrev) {
                    float pl = (st.restir && depth == 1 ? 1.0f : pEnv) * envPdf(c, dir);
                    float den = st.qPrev * st.qPrev + pl * pl;
                    wB = den > 0.0f ? st.qPrev * st.qPrev / den : 1.0f;
                }
                partsAdd(st.parts, depth, cls, thr * envRadiance(c, dir) * wB, SPLIT);
            }
            alive = false;
            break;
        }
        // the instance record (vertex offset, first index, material, emitter base, then the 3 x 4
        // transform), then the triangle's indices, then its vertices and vertex normals: each
        // group's loads are issued together (a scheduling barrier after them), so the chain is three
        // memory round trips, not one per index and vertex
        const uint32_t inst = hq.inst * 16;
        const uint4* IR = reinterpret_cast<const uint4*>(S.instances + inst);
        const uint4 I0 = IR[0], I1 = IR[1], I2 = IR[2], I3 = IR[3];
        __builtin_amdgcn_sched_barrier(0);
        const uint32_t tri = I0.y + 3 * hq.prim, vo = I0.x;
        const uint32_t j0 = S.indices[tri], j1 = S.indices[tri + 1], j2 = S.indices[tri + 2];
        __builtin_amdgcn_sched_barrier(0);
        const uint32_t i0 = j0 + vo, i1 = j1 + vo, i2 = j2 + vo;
        const V3 q0 = vtx(S.positions, i0), q1 = vtx(S.positions, i1), q2 = vtx(S.positions, i2);
        const V3 vn0 = vtx(S.vnormals, i0), vn1 = vtx(S.vnormals, i1), vn2 = vtx(S.vnormals, i2);
        __builtin_amdgcn_sched_barrier(0);
        const float X[12] = {asf(I1.x), asf(I1.y), asf(I1.z), asf(I1.w), asf(I2.x), asf(I2.y),
                             asf(I2.z), asf(I2.w), asf(I3.x), asf(I3.y), asf(I3.z), asf(I3.w)};
        V3 p0 = transformPoint(X, q0);
        V3 p1 = transformPoint(X, q1);
        V3 p2 = transformPoint(X, q2);
        V3 nOut = normalize(cross(p1 - p0, p2 - p0));
        V3 wo = -dir;
        bool outside = dot(nOut, wo) > 0.0f;
        V3 ng = outside ? nOut : -nOut;
        float pHitT = hq.t;
        st.pathLen += pHitT;
        V3 hit = org + dir * pHitT;
        V3 e1 = p1 - p0, e2 = p2 - p0, e3 = hit - p0;
        float d11 = dot(e1, e1), d12 = dot(e1, e2), d22 = dot(e2, e2);
        float den = fmaxf(d11 * d22 - d12 * d12, 1.0e-20f);
        float b1 = (d22 * dot(e3, e1) - d12 * dot(e3, e2)) / den;
        float b2 = (d11 * dot(e3, e2) - d12 * dot(e3, e1)) / den;

        // (every opacity is 1 in a scene without translucent materials)
        float opacity = S.translucent ? opacityAt(c, hq.inst, hq.prim, b1, b2) : 1.0f;
        if (opacity < 1.0f && rnext(rng) >= opacity) {
            org = hit + dir * 1.0e-4f;
            break;
        }
        V3 p = hit + ng * 1.0e-4f;
        if (RESTIR && depth == 1 && st.gi) {
            // the first bounce reached a surface: ReSTIR GI's candidate point and facing normal (pt_core.slang:1072-1082)
            float* C = c.F.restir + restirBase(pixel, P.w[6] & 1u);
            C[16] = hit.x;
            C[17] = hit.y;
            C[18] = hit.z;
            C[19] = ng.x;
            C[20] = ng.y;
            C[21] = ng.z;
            c.F.restir[restirScratch(pixel) + 22] = asf(1u);
        }

        V3 nsObj = (1.0f - b1 - b2) * vn0 + b1 * vn1 + b2 * vn2;
        V3 ns = nOut;
        if (dot(nsObj, nsObj) > 1.0e-12f) {
            ns = normalize(transformNormal(X, nsObj));
            if (dot(ns, nOut) < 0.0f)
                ns = -ns;
        }
        V3 n = outside ? ns : -ns;

        const float* M = S.materials + I0.z * 24;
        V3 albedo = v3(M[0], M[1], M[2]);
        uint32_t texBase = f2u(M[10]), texMR = f2u(M[11]), texEm = f2u(M[12]), texNm = f2u(M[13]);
        float texRMx = 1.0f, texRMy = 1.0f;
        V3 texEmission = v3(1.0f);
        if (texBase + texMR + texEm + texNm > 0) {
            float u0x = S.uvs[2 * i0], u0y = S.uvs[2 * i0 + 1];
            float u1x = S.uvs[2 * i1], u1y = S.uvs[2 * i1 + 1];
            float u2x = S.uvs[2 * i2], u2y = S.uvs[2 * i2 + 1];
            float ux = u0x + b1 * (u1x - u0x) + b2 * (u2x - u0x);
            float uy = u0y + b1 * (u1y - u0y) + b2 * (u2y - u0y);
            float d1x = u1x - u0x, d1y = u1y - u0y, d2x = u2x - u0x, d2y = u2y - u0y;
            float uvPerWorld = sqrtf(fabsf(d1x * d2y - d1y * d2x) / fmaxf(length(cross(e1, e2)), 1.0e-20f));
            float footprint = st.pathLen * P.f(27) / fmaxf(fabsf(dot(nOut, dir)), 0.1f) * uvPerWorld;
            if (texBase > 0)
                albedo = albedo * sampleTexture(c, texBase - 1, ux, uy, footprint);
            if (texMR > 0) {
                V3 mr = sampleTexture(c, texMR - 1, ux, uy, footprint);
                texRMx = mr.y;
                texRMy = mr.z;
            }
            if (texEm > 0)
                texEmission = sampleTexture(c, texEm - 1, ux, uy, footprint);
            if (texNm > 0) {
                float det = d1x * d2y - d2x * d1y;
                V3 dPdu = (e1 * d2y - e2 * d1y) / det;
                V3 dPdv = (e2 * d1x - e1 * d2x) / det;
                V3 tg = dPdu - ns * dot(ns, dPdu);
                if (fabsf(det) > 1.0e-20f && dot(tg, tg) > 1.0e-20f) {
                    tg = normalize(tg);
                    V3 bt = cross(ns, tg);
                    if (dot(bt, dPdv) > 0.0f)
                        bt = -bt;
                    V3 nt = sampleTexture(c, texNm - 1, ux, uy, footprint) * 2.0f - 1.0f;
                    ns = normalize(nt.x * M[14] * tg + nt.y * M[14] * bt + nt.z * ns);
                    n = outside ? ns : -ns;
                }
            }
        }
        float woN = dot(n, wo);
        if (woN < 0.05f)
            n = normalize(n + wo * (0.05f - woN));

        V3 le = v3(M[3], M[4], M[5]) * texEmission;
        if (st.prevInst != 0 && (emitMask(c, hq.inst) & recvMask(c, st.prevInst - 1)) == 0)
            le = v3(0.0f);
        if (le.x > 0.0f || le.y > 0.0f || le.z > 0.0f) {
            float wB = 1.0f;
            uint32_t eb = I0.w;
            if (st.misPrev && eb != 0) {
                const float* E = S.emitters + (size_t)(eb - 1 + hq.prim) * 16;
                float cosl = fmaxf(fabsf(dot(nOut, dir)), 1.0e-4f);
                float pl = (st.restir && depth == 1 ? 1.0f : 1.0f - pEnv) * E[13] * pHitT * pHitT / (E[12] * cosl);
                wB = st.qPrev * st.qPrev / (st.qPrev * st.qPrev + pl * pl);
            }
            partsAdd(st.parts, depth, cls, thr * le * wB, SPLIT);
            if (RESTIR && depth == 1 && st.gi) {
                // emission the first bounce sees stays with the first hit's own light (pt_core.slang:1162-1168)
                float* Sd = c.F.restir + restirScratch(pixel) + 18;
                V3 a = thr * le * wB;
                Sd[0] += a.x;
                Sd[1] += a.y;
                Sd[2] += a.z;
            }
        }
        if constexpr (NRC == 1) {
            if (depth < kMaxDepthNrc) {
                float4* v = io->vertices + ((size_t)io->path * kMaxDepthNrc + depth) * kNrcVertexWords;
                v[0] = make_float4(hit.x, hit.y, hit.z, M[7] * texRMx);
                v[1] = make_float4(wo.x, wo.y, wo.z, 0.0f);
                v[2] = make_float4(n.x, n.y, n.z, 0.0f);
                v[3] = make_float4(thr.x, thr.y, thr.z, 0.0f);
                v[4] = make_float4(st.parts.total.x, st.parts.total.y, st.parts.total.z, 1.0f); // w: recorded
            }
        }
        if constexpr (NRC == 2) {
            if (depth >= io->terminateDepth) {
                float4* q = io->queries + (size_t)io->slot * kNrcQueryWords;
                q[0] = make_float4(hit.x, hit.y, hit.z, M[7] * texRMx);
                q[1] = make_float4(wo.x, wo.y, wo.z, asf(pixel));
                q[2] = make_float4(n.x, n.y, n.z, 0.0f);
                q[3] = make_float4(thr.x, thr.y, thr.z, 0.0f);
                q[4] = make_float4(st.parts.total.x, st.parts.total.y, st.parts.total.z, 1.0f); // w: a query
                io->terminated = true;
                alive = false;
                break;
            }
        }
        float metallic = M[6] * texRMy;
        float rough = M[7] * texRMx;
        float transmission = M[8];
        float ior = M[9];
        float dispersion = M[21];
        if (dispersion > 0.0f && transmission > 0.0f) {
            if (chan < 0) {
                int ch = (int)(rnext(rng) * 3.0f);
                chan = ch < 2 ? ch : 2;
                thr = thr * (3.0f * v3(chan == 0 ? 1.0f : 0.0f, chan == 1 ? 1.0f : 0.0f, chan == 2 ? 1.0f : 0.0f));
            }
            ior = fmaxf(ior + (float)(chan - 1) * (ior - 1.0f) * 0.025f * dispersion, 1.0f);
        }
        // a glass second hit is no reconnection candidate (pt_core.slang:1187-1190)
        if (RESTIR && depth == 1 && st.gi && transmission > 0.0f)
            c.F.restir[restirScratch(pixel) + 22] = asf(0u);
        float alpha = fmaxf(rough * rough, 1.0e-3f);
        float eta = outside ? ior : 1.0f / ior;

        V3 T, B;
        basis(n, T, B);
        V3 woL = v3(dot(wo, T), dot(wo, B), fmaxf(dot(wo, n), 1.0e-4f));
        float fSel = fresnelDielectric(woL.z, ior);
        Surface surf;
        surf.albedo = albedo;
        surf.metallic = metallic;
        surf.alpha = alpha;
        surf.transmission = transmission;
        surf.ior = ior;
        surf.eta = eta;
        surf.fSel = fSel;
        surf.tComp = 1.0f;
        if (P.w[29] != 0 && transmission > 0.0f && rough > 0.0f)
            surf.tComp = 1.0f / fmaxf(dielectricAlbedo(c, woL.z, rough, ior, outside), 0.1f);
        surf.coat = outside ? M[15] : 0.0f;
        surf.coatAlpha = fmaxf(M[16] * M[16], 1.0e-3f);
        float coatF = fresnelDielectric(woL.z, 1.5f);
        surf.coatP = surf.coat * coatF;
        surf.sheenColor = outside ? v3(M[18], M[19], M[20]) : v3(0.0f);
        float sheenMax = fmaxf(surf.sheenColor.x, fmaxf(surf.sheenColor.y, surf.sheenColor.z));
        float sheenR = fmaxf(M[17], 0.07f);
        surf.sheenAlpha = sheenR * sheenR;
        surf.sheenP = sheenMax > 0.0f ? sheenMax * sheenAlbedo(c, woL.z, sheenR) : 0.0f;
        surf.ms = multiscatterFactor(c, surf, woL.z);
        surf.g1o = G1(woL.z, alpha);
        surf.g1oC = surf.coat > 0.0f ? G1(woL.z, surf.coatAlpha) : 0.0f;

        // delta lights (pt_core.slang:1224-1282); they draw no random numbers
        for (uint32_t li = 0; li < (GIP == 1 ? 0u : lightCount); li++) {
            const uint32_t* L = S.lights + li * 16;
            V3 col = v3(asf(L[1]), asf(L[2]), asf(L[3])) * asf(L[4]);
            V3 vec = v3(asf(L[5]), asf(L[6]), asf(L[7]));
            bool virt = L[0] == 3;
            if (virt && !(lightingOn && depth == 0))
                continue;
            if ((L[14] & recvMask(c, hq.inst)) == 0)
                continue;
            V3 wi, li_;
            float dist;
            if (L[0] == 1 || virt) {
                wi = v3(S.lightDirs[4 * li], S.lightDirs[4 * li + 1], S.lightDirs[4 * li + 2]); // -normalize(vec)
                dist = 1.0e30f;
                li_ = col;
            } else {
                V3 toL = vec - p;
                float d2 = dot(toL, toL);
                dist = sqrtf(d2);
                wi = toL / dist;
                li_ = col / d2;
                if (L[15] != 0)
                    li_ = li_ * iesFactor(c, L[15] - 1, -wi);
                if (L[0] == 2) {
                    V3 axis = v3(asf(L[8]), asf(L[9]), asf(L[10]));
                    float att = saturate(dot(axis, -wi) * asf(L[11]) + asf(L[12]));
                    li_ = li_ * (att * att);
                }
            }
            float cc = dot(n, wi);
            if (DEFER && cc > 0.0f) {
                // plain integrator: virtual lights were skipped above
                Lobes l = evalLobes(c, surf, woL, v3(dot(wi, T), dot(wi, B), cc));
                sink->push(p, wi, dist - 1.0e-3f, thr * (l.diffuse + l.specular) * li_);
            } else if (cc > 0.0f) {
                bool occ = occluded(c, p, wi, dist - 1.0e-3f);
                bool track = lightingOn && depth == 0;
                if (!occ || track) {
                    Lobes l = evalLobes(c, surf, woL, v3(dot(wi, T), dot(wi, B), cc));
                    V3 cd = thr * l.diffuse * li_;
                    V3 cs = thr * l.specular * li_;
                    if (!occ && !virt)
                        partsAddLight(st.parts, depth, cls, thr * (l.diffuse + l.specular) * li_, cd, cs, SPLIT);
                    if (SPLIT && track) {
                        float vv = virt ? 0.0f : lum(cd + cs);
                        st.visTotal += vv;
                        st.visLit += occ ? 0.0f : vv;
                        addSlot(c, pixel, L[13], cd / cc, cs / cc, wi, occ ? 0.0f : 1.0f, occ ? 0.0f : cc);
                    }
                }
            }
        }

        // ReSTIR DI at the first hit (pt_core.slang:1286-1332)
        bool reDI = false;
        if (RESTIR && depth == 0) {
            bool suits = P.f(3) == 0.0f && transmission == 0.0f && surf.alpha >= kRestirMinAlpha &&
                         (surf.coat == 0.0f || surf.coatAlpha >= kRestirMinAlpha);
            reDI = suits && (P.w[36] & 1u) != 0 && emitterCount > 0;
            // ReSTIR GI replaces the path's own indirect light, which the lighting split cannot follow
            bool reGI = suits && (P.w[36] & 2u) != 0 && !lightingOn;
            st.restir = reDI;
            st.gi = reGI;
            if constexpr (GIP == 1) {
                site->active = reGI;
                site->p = p;
                site->n = n;
                site->woL = woL;
                site->T = T;
                site->B = B;
                site->pHitT = pHitT;
                site->surf = surf;
                site->rng = rng;
                return false;
            }
            if (suits) {
                restirGbuffer(c, pixel, p, n);
                if (reGI) {
                    if (GIP == 2)
                        rng = asu(c.F.restir[restirScratch(pixel) + 23]);
                    else
                        restirGIPrepare(c, pixel, p, n, pHitT, surf, T, B, woL, rng);
                }
            } else
                restirClear(c, pixel);
        }
        if (RESTIR && reDI) {
            RestirPick pick = restirDI<DEFER>(c, pixel, recvMask(c, hq.inst), p, n, pHitT, surf, T, B, woL, rng);
            if (pick.valid) {
                V3 wi, g;
                float dist, cosi, plsa;
                restirLight(c, pick.e - 1, recvMask(c, hq.inst), pick.x, p, n, wi, dist, cosi, g, plsa);
                bool track = lightingOn;
                if (!pick.occ || track) {
                    V3 wiL = v3(dot(wi, T), dot(wi, B), cosi);
                    float wl = restirMis(surf, woL, wiL, plsa);
                    Lobes l = evalLobes(c, surf, woL, wiL);
                    V3 cd = thr * l.diffuse * g * (wl * pick.W);
                    V3 cs = thr * l.specular * g * (wl * pick.W);
                    if (DEFER) {
                        sink->push(p, wi, dist * (1.0f - 1.0e-3f), cd + cs);
                        sink->restirRec = sink->count;
                    } else if (!pick.occ) {
                        partsAddLight(st.parts, depth, cls, cd + cs, cd, cs, SPLIT);
                    }
                    if (SPLIT && track) {
                        float vv = lum(cd + cs);
                        st.visTotal += vv;
                        st.visLit += pick.occ ? 0.0f : vv;
                        addSlot(c, pixel, 3u, cd / cosi, cs / cosi, wi, pick.occ ? 0.0f : 1.0f, pick.occ ? 0.0f : cosi);
                    }
                }
            }
        }

        // emitter NEE (pt_core.slang:1335-1390)
        float chooseLight = rnext(rng);
        float pEnvHere = reDI ? (hasEnv ? 1.0f : 0.0f) : pEnv;
        bool chooseEnv = hasEnv && chooseLight < pEnvHere;
        if (emitterCount > 0 && !chooseEnv && !reDI) {
            float ue = rnext(rng), ua = rnext(rng), ub = rnext(rng);
            uint32_t lo = emitterIndex(S, ue);
            const float* E = S.emitters + (size_t)lo * 16;
            V3 a = v3(E[0], E[1], E[2]), b = v3(E[3], E[4], E[5]), ce = v3(E[6], E[7], E[8]);
            float su = sqrtf(ua);
            V3 xl = a * (1.0f - su) + b * (su * (1.0f - ub)) + ce * (su * ub);
            V3 toL = xl - p;
            float d2 = dot(toL, toL);
            float dist = sqrtf(d2);
            V3 wi = toL / dist;
            float cosl = fabsf(dot(normalize(cross(b - a, ce - a)), wi));
            float cosi = dot(n, wi);
            if (DEFER && cosi > 0.0f && cosl > 1.0e-5f && (asu(E[14]) & recvMask(c, hq.inst)) != 0) {
                float pl = (1.0f - pEnv) * E[13] * d2 / (E[12] * cosl);
                V3 wiL = v3(dot(wi, T), dot(wi, B), cosi);
                float qb = reflectivePdf(surf, woL, wiL);
                float wl = pl * pl / (pl * pl + qb * qb);
                V3 lev = v3(E[9], E[10], E[11]);
                Lobes l = evalLobes(c, surf, woL, wiL);
                sink->push(p, wi, dist * (1.0f - 1.0e-3f), thr * (l.specular + l.diffuse) * lev / pl * wl);
            } else if (cosi > 0.0f && cosl > 1.0e-5f && (asu(E[14]) & recvMask(c, hq.inst)) != 0) {
                bool occ = occluded(c, p, wi, dist * (1.0f - 1.0e-3f));
                bool track = lightingOn && depth == 0;
                if (!occ || track) {
                    float pl = (1.0f - pEnv) * E[13] * d2 / (E[12] * cosl);
                    V3 wiL = v3(dot(wi, T), dot(wi, B), cosi);
                    float qb = reflectivePdf(surf, woL, wiL);
                    float wl = pl * pl / (pl * pl + qb * qb);
                    V3 lev = v3(E[9], E[10], E[11]);
                    Lobes l = evalLobes(c, surf, woL, wiL);
                    V3 cd = thr * l.diffuse * lev / pl * wl;
                    V3 cs = thr * l.specular * lev / pl * wl;
                    if (!occ)
                        partsAddLight(st.parts, depth, cls, thr * (l.specular + l.diffuse) * lev / pl * wl, cd, cs, SPLIT);
                    if (SPLIT && track) {
                        float vv = lum(cd + cs);
                        st.visTotal += vv;
                        st.visLit += occ ? 0.0f : vv;
                        addSlot(c, pixel, 3u, cd / cosi, cs / cosi, wi, occ ? 0.0f : 1.0f, occ ? 0.0f : cosi);
                    }
                }
            }
        }

        // environment NEE (pt_core.slang:1392-1427)
        if (chooseEnv) {
            float u1e = rnext(rng), u2e = rnext(rng), u3e = rnext(rng), u4e = rnext(rng);
            float pdfE;
            V3 leE;
            V3 wi = envSample(c, u1e, u2e, u3e, u4e, pdfE, leE);
            float cosi = dot(n, wi);
            if (DEFER && pdfE > 0.0f && cosi > 0.0f) {
                float pl = pEnvHere * pdfE;
                V3 wiL = v3(dot(wi, T), dot(wi, B), cosi);
                float qb = reflectivePdf(surf, woL, wiL);
                float wl = pl * pl / (pl * pl + qb * qb);
                Lobes l = evalLobes(c, surf, woL, wiL);
                sink->push(p, wi, 1.0e30f, thr * (l.specular + l.diffuse) * leE / pl * wl);
            } else if (pdfE > 0.0f && cosi > 0.0f) {
                bool occ = occluded(c, p, wi, 1.0e30f);
                bool track = lightingOn && depth == 0;
                if (!occ || track) {
                    float pl = pEnvHere * pdfE;
                    V3 wiL = v3(dot(wi, T), dot(wi, B), cosi);
                    float qb = reflectivePdf(surf, woL, wiL);
                    float wl = pl * pl / (pl * pl + qb * qb);
                    Lobes l = evalLobes(c, surf, woL, wiL);
                    V3 cd = thr * l.diffuse * leE / pl * wl;
                    V3 cs = thr * l.specular * leE / pl * wl;
                    if (!occ)
                        partsAddLight(st.parts, depth, cls, thr * (l.specular + l.diffuse) * leE / pl * wl, cd, cs, SPLIT);
                    if (SPLIT && track) {
                        float vv = lum(cd + cs);
                        st.visTotal += vv;
                        st.visLit += occ ? 0.0f : vv;
                        if (P.w[41] == 0)
                            addSlot(c, pixel, 3u, cd / cosi, cs / cosi, wi, occ ? 0.0f : 1.0f, occ ? 0.0f : cosi);
                    }
                }
            }
        }

        // one-lobe BSDF sampling (pt_core.slang:1431-1523)
        float lobe = rnext(rng);
        float u1 = rnext(rng), u2 = rnext(rng), u3 = rnext(rng);
        V3 wiL = v3(0.0f, 0.0f, 1.0f);
        V3 wgt = v3(1.0f);
        bool ok = true, refracted = false;
        uint32_t lobeCls = 1;
        float uc = surf.coat > 0.0f ? rnext(rng) : 1.0f;
        float us = surf.sheenP > 0.0f ? rnext(rng) : 1.0f;
        if (uc < surf.coatP) {
            V3 h = sampleVndf(woL, surf.coatAlpha, u1, u2);
            float woh = dot(woL, h);
            wiL = 2.0f * woh * h - woL;
            ok = wiL.z > 0.0f;
            wgt = v3(fresnelDielectric(woh, 1.5f) / coatF * G1(fmaxf(wiL.z, 0.0f), surf.coatAlpha));
        } else if (us < surf.sheenP) {
            wiL = cosineLocal(u1, u2);
            V3 hh = normalize(woL + wiL);
            wgt = surf.sheenColor * sheenCore(woL.z, wiL.z, hh.z, surf.sheenAlpha) * kPi / (wiL.z * surf.sheenP);
        } else if (lobe < metallic) {
            V3 h = sampleVndf(woL, alpha, u1, u2);
            float woh = dot(woL, h);
            wiL = 2.0f * woh * h - woL;
            ok = wiL.z > 0.0f;
            wgt = schlick(albedo, woh) * surf.ms * G1(fmaxf(wiL.z, 0.0f), alpha);
        } else if (lobe < metallic + (1.0f - metallic) * transmission) {
            V3 h = sampleVndf(woL, alpha, u1, u2);
            float woh = dot(woL, h);
            if (u3 < fresnelDielectric(woh, eta)) {
                wiL = 2.0f * woh * h - woL;
                ok = wiL.z > 0.0f;
                wgt = v3(G1(fmaxf(wiL.z, 0.0f), alpha) * surf.tComp);
            } else {
                float r = 1.0f / eta;
                float k = 1.0f - r * r * (1.0f - woh * woh);
                wiL = -woL * r + (r * woh - sqrtf(fmaxf(k, 0.0f))) * h;
                ok = wiL.z < 0.0f && k > 0.0f;
                refracted = true;
                lobeCls = 2;
                wgt = albedo * G1(fmaxf(-wiL.z, 0.0f), alpha) * r * r * surf.tComp;
            }
        } else if (u3 < fSel) {
            V3 h = sampleVndf(woL, alpha, u1, u2);
            float woh = dot(woL, h);
            wiL = 2.0f * woh * h - woL;
            ok = wiL.z > 0.0f;
            wgt = v3(fresnelDielectric(woh, ior) / fSel * G1(fmaxf(wiL.z, 0.0f), alpha));
        } else {
            wiL = cosineLocal(u1, u2);
            wgt = albedo;
            lobeCls = 0;
        }
        if (!ok) {
            alive = false;
            break;
        }
        if (depth == 0)
            cls = lobeCls;
        st.misPrev = !refracted;
        st.qPrev = refracted ? 0.0f : reflectivePdf(surf, woL, wiL);
        st.prevInst = hq.inst + 1;

        dir = normalize(wiL.x * T + wiL.y * B + wiL.z * n);
        float side = dot(dir, ng);
        if (refracted ? side >= 0.0f : side <= 0.0f) {
            alive = false;
            break;
        }
        org = hit + ng * (side > 0.0f ? 1.0e-4f : -1.0e-4f);
        thr = thr * wgt;

        if (depth >= 2) {
            float pr = clampf(fmaxf(thr.x, fmaxf(thr.y, thr.z)), 0.05f, 0.95f);
            if (rnext(rng) > pr) {
                alive = false;
                break;
            }
            thr = thr / pr;
        }
    } while (false);
    st.rng = rng;
    st.org = org;
    st.dir = dir;
    st.thr = thr;
    st.cls = cls;
    st.chan = chan;
    if (RESTIR && depth == 0 && st.gi) {
        // the first hit's own gathered light, the first bounce's throughput and pdf (pt_core.slang:1531-1540)
        float* Sc = c.F.restir + restirScratch(pixel);
        Sc[15] = thr.x;
        Sc[16] = thr.y;
        Sc[17] = thr.z;
        Sc[18] = st.parts.total.x;
        Sc[19] = st.parts.total.y;
        Sc[20] = st.parts.total.z;
        Sc[21] = st.qPrev;
    }
    return alive;
}

// pixelActive (pt_core.slang:1551-1559) on the FP64 statistics
__device__ __forceinline__ bool pixelActive(const Ctx& c, uint32_t p)
{
    const double* S = c.F.stats + 3 * (size_t)p;
    double n = S[0];
    if (n < (double)c.P.w[34])
        return true;
    double mean = S[1] / n;
    double var = fmax(S[2] / n - mean * mean, 0.0) * n / (n - 1.0);
    return sqrt(var / n) > (double)c.P.f(33) * fmax(mean, 0.05);
}

// finishPath (pt_core.slang:1620-1667), FP64 film; only this path's thread touches the pixel
template <bool SPLIT, bool RESTIR = false>
__device__ void finishPath(const Ctx& c, const PathState& st, uint32_t pixel)
{
    Parts parts = st.parts;
    float visLit = st.visLit, visTotal = st.visTotal;
    V3 radiance = parts.total;
    if (RESTIR && !SPLIT && st.gi) {
        // a path through dispersive glass carries one tripled channel: not shared (pt_core.slang:1626-1632)
        if (st.chan >= 0)
            c.F.restir[restirScratch(pixel) + 22] = asf(0u);
        radiance = giFinish(c, radiance, pixel);
    }
    if (bad(radiance)) {
        radiance = v3(0.0f);
        parts = Parts{v3(0.0f), v3(0.0f), v3(0.0f), v3(0.0f), v3(0.0f), v3(0.0f)};
        visLit = 0.0f;
        visTotal = 0.0f;
    }
    double* F = c.F.film + 3 * (size_t)pixel;
    F[0] += (double)radiance.x;
    F[1] += (double)radiance.y;
    F[2] += (double)radiance.z;
    if (c.P.w[32] != 0) {
        double l = (double)lum(radiance);
        double* S = c.F.stats + 3 * (size_t)pixel;
        S[0] += 1.0;
        S[1] += l;
        S[2] += l * l;
    }
    if (SPLIT) {
        float* L = c.F.lighting + (size_t)kLightingStride * pixel;
        V3 split[5] = {parts.directDiffuse, parts.indirectDiffuse, parts.specular, parts.transmission, parts.emission};
#pragma unroll
        for (uint32_t k = 0; k < 5; k++) {
            L[3 * k] += split[k].x;
            L[3 * k + 1] += split[k].y;
            L[3 * k + 2] += split[k].z;
        }
        L[15] += visLit;
        L[16] += visTotal;
    }
}

} // namespace vh
