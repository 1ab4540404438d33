// vh_aov.hip - primary-hit AOVs: render_normals (shaders/normals.slang:39-110) and render_aovs
// (shaders/aov.slang:84-188), behind RayTracingBackend::render_normals / render_aovs
// (crates/verion-gpu/src/lib.rs:54-57).
//
// The reference runs one 8x8 thread group per 64 pixels and traces the pixel-centre ray and then
// the 16 coverage rays of a pixel one after another in one thread. Here the two halves are two
// launches shaped for a 64-wide wavefront:
//   - k_aov_primary: one lane per pixel for the centre ray (normal, albedo, depth, ids);
//   - k_aov_coverage: 16 lanes per pixel, one stratified coverage ray each, so a wave traverses
//     the 16 coherent rays of 4 pixels together instead of 16 rays in sequence per lane; lane 0 of
//     the group sums the 16 results in ray order with cross-lane reads (the reference's sequential
//     float sum, so the result is the same bit for bit).
// Outputs are written as whole per-pixel records, coalesced across the wave.
#include "vh_internal.h"

namespace vh {

namespace {

// pixelDir (aov.slang:73-81; normals.slang:47-57 inline): the camera-space unit direction through
// pixel position (pxx, pxy) and, in `dir`, its world-space direction (not yet normalised).
__device__ __forceinline__ V3 pixelDir(const Params& P, float pxx, float pxy, V3& dir)
{
    uint32_t w = P.w[0], h = P.w[1];
    float t = P.f(2);
    float aspect = (float)w / (float)h;
    float u = (2.0f * pxx / (float)w - 1.0f + P.f(39)) * aspect * t;
    float v = (1.0f - 2.0f * pxy / (float)h + P.f(40)) * t;
    V3 d = normalize(v3(u, v, -1.0f));
    float r3[3];
#pragma unroll
    for (uint32_t r = 0; r < 3; r++)
        r3[r] = P.f(8 + 4 * r) * d.x + P.f(9 + 4 * r) * d.y + P.f(10 + 4 * r) * d.z;
    dir = v3(r3[0], r3[1], r3[2]);
    return d;
}

__device__ __forceinline__ V3 cameraOrigin(const Params& P) { return v3(P.f(11), P.f(15), P.f(19)); }

// NORMALS: out3 gets normals.slang's normal. Otherwise albedo + depth (float4), object and material
// ids of aov.slang's centre ray.
template <bool NORMALS>
__global__ __launch_bounds__(256) void k_aov_primary(Scene S, Params P, float* out3, float4* albedoDepth, uint32_t* objectId,
                                                     uint32_t* materialId)
{
    Frame F{};
    Ctx c{S, F, P};
    const uint32_t w = P.w[0], n = P.w[0] * P.w[1];
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        uint32_t x = p % w, y = p / w;
        V3 dir;
        V3 d = pixelDir(P, (float)x + 0.5f, (float)y + 0.5f, dir);
        V3 org = cameraOrigin(P);
        dir = normalize(dir);
        Hit q = trace<false>(S.bvh, org, dir, 1.0e30f);
        V3 nrm = v3(0.0f), albedo = v3(0.0f);
        float depth = 0.0f;
        uint32_t id = 0, matId = 0;
        if (q.hit) {
            uint32_t inst = q.inst * 16;
            uint32_t tri = S.instances[inst + 1] + 3 * q.prim;
            uint32_t vo = S.instances[inst];
            uint32_t i0 = S.indices[tri] + vo, i1 = S.indices[tri + 1] + vo, i2 = S.indices[tri + 2] + vo;
            const float* M = S.materials + S.instances[inst + 2] * 24;
            uint32_t texBase = f2u(M[10]);
            float b1 = 0.0f, b2 = 0.0f;
            V3 p0 = v3(0.0f), p1 = v3(0.0f), p2 = v3(0.0f);
            if (NORMALS || texBase > 0) {
                p0 = transformPoint(S.instances, inst + 4, vtx(S.positions, i0));
                p1 = transformPoint(S.instances, inst + 4, vtx(S.positions, i1));
                p2 = transformPoint(S.instances, inst + 4, vtx(S.positions, i2));
                V3 e1 = p1 - p0, e2 = p2 - p0, e3 = org + dir * q.t - p0;
                float d11 = dot(e1, e1), d12 = dot(e1, e2), d22 = dot(e2, e2);
                float den = fmaxf(d11 * d22 - d12 * d12, 1.0e-20f);
                b1 = (d22 * dot(e3, e1) - d12 * dot(e3, e2)) / den;
                b2 = (d11 * dot(e3, e2) - d12 * dot(e3, e1)) / den;
            }
            if (NORMALS) {
                V3 ng = normalize(cross(p1 - p0, p2 - p0));
                nrm = ng;
                V3 nsObj = (1.0f - b1 - b2) * vtx(S.vnormals, i0) + b1 * vtx(S.vnormals, i1) + b2 * vtx(S.vnormals, i2);
                if (dot(nsObj, nsObj) > 1.0e-12f) {
                    nrm = normalize(transformNormal(S.instances, inst + 4, nsObj));
                    if (dot(nrm, ng) < 0.0f)
                        nrm = -nrm;
                }
                if (dot(ng, dir) > 0.0f)
                    nrm = -nrm;
            } else {
                id = q.inst + 1;
                matId = S.instances[inst + 2] + 1;
                depth = q.t * -d.z;
                albedo = v3(M[0], M[1], M[2]);
                if (texBase > 0) {
                    float u0x = S.uvs[2 * i0], u0y = S.uvs[2 * i0 + 1];
                    float u1x = S.uvs[2 * i1], u1y = S.uvs[2 * i1 + 1];
                    float u2x = S.uvs[2 * i2], u2y = S.uvs[2 * i2 + 1];
                    float ux = u0x + b1 * (u1x - u0x) + b2 * (u2x - u0x);
                    float uy = u0y + b1 * (u1y - u0y) + b2 * (u2y - u0y);
                    albedo = albedo * sampleTexture(c, texBase - 1, ux, uy, 0.0f);
                }
            }
        }
        if (NORMALS) {
            out3[3 * (size_t)p] = nrm.x;
            out3[3 * (size_t)p + 1] = nrm.y;
            out3[3 * (size_t)p + 2] = nrm.z;
        } else {
            albedoDepth[p] = make_float4(albedo.x, albedo.y, albedo.z, depth);
            objectId[p] = id;
            materialId[p] = matId;
        }
    }
}

// Coverage (aov.slang:160-179): lane k of a 16-lane group traces stratified ray k of its pixel
// through up to 8 translucent layers; the group's first lane sums the 16 in ray order.
__global__ __launch_bounds__(256) void k_aov_coverage(Scene S, Params P, float* coverage)
{
    Frame F{};
    Ctx c{S, F, P};
    const uint32_t w = P.w[0], n = P.w[0] * P.w[1];
    const uint32_t lane = threadIdx.x & 63u, k = lane & 15u, first = lane & ~15u;
    // every lane of a wave runs the same number of iterations (the cross-lane reads below need
    // the whole wave): the loop bound is per wave, pixels past the end trace nothing
    const uint32_t groupsPerGrid = gridDim.x * (blockDim.x / 16);
    for (uint32_t g0 = blockIdx.x * (blockDim.x / 16) + (threadIdx.x & ~63u) / 16; g0 < n; g0 += groupsPerGrid) {
        uint32_t p = g0 + lane / 16;
        float seen = 0.0f;
        if (p < n) {
            uint32_t x = p % w, y = p / w;
            V3 sdir;
            pixelDir(P, (float)x + ((float)(k % 4) + 0.5f) / 4.0f, (float)y + ((float)(k / 4) + 0.5f) / 4.0f, sdir);
            V3 so = cameraOrigin(P), sd = normalize(sdir);
            float through = 1.0f;
            for (uint32_t layer = 0; layer < 8 && through > 0.0f; layer++) {
                Hit sq = trace<false>(S.bvh, so, sd, 1.0e30f);
                if (!sq.hit)
                    break;
                float op = opacityAt(c, sq.inst, sq.prim, sq.b1, sq.b2);
                seen += through * op;
                through *= 1.0f - op;
                so = so + sd * (sq.t + 1.0e-4f);
            }
        }
        float cover = 0.0f;
#pragma unroll
        for (uint32_t j = 0; j < 16; j++)
            cover += __shfl(seen, (int)(first + j), 64);
        if (k == 0 && p < n)
            coverage[p] = cover / 16.0f;
    }
}

} // namespace

int launch_normals(const Scene& S, const Params& P, float* out3, hipStream_t s)
{
    uint64_t n = (uint64_t)P.w[0] * P.w[1];
    k_aov_primary<true><<<blocks_for(n, 256, 8192), 256, 0, s>>>(S, P, out3, nullptr, nullptr, nullptr);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? VH_OK : fail(VH_ERR_KERNEL, std::string("normals launch: ") + hipGetErrorString(e));
}

int launch_aovs(const Scene& S, const Params& P, float4* albedoDepth, uint32_t* objectId, uint32_t* materialId, float* coverage,
                hipStream_t s)
{
    uint64_t n = (uint64_t)P.w[0] * P.w[1];
    k_aov_primary<false><<<blocks_for(n, 256, 8192), 256, 0, s>>>(S, P, nullptr, albedoDepth, objectId, materialId);
    // 16 pixels per 256-thread block
    k_aov_coverage<<<blocks_for(n * 16, 256, 16384), 256, 0, s>>>(S, P, coverage);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? VH_OK : fail(VH_ERR_KERNEL, std::string("AOV launch: ") + hipGetErrorString(e));
}

} // namespace vh
