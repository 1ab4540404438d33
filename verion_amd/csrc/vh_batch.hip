// vh_batch.hip - many render jobs in one persistent wavefront, hipGraph-replayed sample passes
// and an online autotuner.
//
// The reference renders one job at a time: the session loop calls accumulate in 8-spp chunks
// (crates/verion-render/src/session.rs:323-336) and a small job leaves most of the GPU idle. Here
// every job of a batch is resident at once:
//   - one slot-indexed wavefront for all jobs; job j owns the wave-aligned slot segment
//     [slotBase_j, slotBase_j + roundup(pixels_j, 64)), and keeps its own queue counter per depth;
//   - the kernels walk 64-slot wave tasks over the concatenated live queues (an exclusive prefix of
//     the jobs' wave counts, built per block in LDS from the device counts); the job of a wave is
//     wave-uniform (readfirstlane), so its scene, film and parameter block are read with scalar
//     loads and cost no vector registers;
//   - survivors are appended per wave (ballot + popcount + one atomic) to the job's next queue;
//   - a sample pass is begin, select (adaptive jobs), 8 x (extend, shade), end. The pass index is
//     device state advanced by the end kernel, so the same launches replay every pass and one
//     captured hipGraph serves the whole run;
//   - the pixel order of depth 0 (tiles of tw x th pixels, or scanlines) and the grid size of the
//     grid-stride kernels are tuned online: the first passes of a run try one candidate each
//     (real work: the image does not depend on them), timed with HIP events and normalised by
//     the paths each pass traced; the fastest runs the rest of the batch.
// Each job's film is the same bit for bit as rendering it alone with vh_accumulate (wavefront):
// a pixel takes one path per pass, and a path's random stream depends only on (pixel, sample).
#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "vh_internal.h"
#include "vh_wavefront.h"

namespace vh {

namespace {

constexpr uint32_t kBxBlock = 256;   // 4 waves
constexpr uint32_t kBxWaves = kBxBlock / 64;
// The traversal kernels (extend, shadow) run one wave per block, so each wave retires on its own;
// their grid is launched kBxTraceScale times larger to cover the same waves.
constexpr uint32_t kBxTraceScale = kBxWaves;
constexpr uint32_t kBxMaxJobs = 255; // one job per thread of the prefix, one more thread for its total
constexpr uint32_t kCountStride = 2 * kMaxDepth + 1; // per job: live paths per depth, then shadow records per depth
constexpr uint32_t kShadowOff = kMaxDepth + 1;
constexpr uint32_t kNotDone = 0xffffffffu;

struct BatchJob {
    Scene S;
    Frame F;
    Params P;          // the job's parameter block; P.w[6] is its first sample
    uint32_t slotBase; // first slot of the job's segment
    uint32_t pixels;
    uint32_t spp;
    uint32_t pad;
};

struct BatchDev {
    uint32_t* state[2];
    uint4* hits;
    uint32_t* queue0;
    uint32_t* counts;              // [job][kMaxDepth + 1]
    uint32_t* pass;                // the pass being rendered
    uint32_t* donePass;            // per job: first pass with nothing to do (kNotDone until then)
    unsigned long long* traced;    // per job: depth-0 paths so far
    unsigned long long* passPaths; // per pass: depth-0 paths of all jobs
    uint32_t maxPasses;            // length of passPaths
    uint32_t n;                    // slots
    uint32_t njobs;
    uint32_t tileW, tileH;         // depth-0 pixel order; tileW = 0: scanlines
    uint32_t traceWaves;           // waves a short queue is spread over (raysPerWave)
    // deferred shadow rays (shK = 0: shade traces them itself), as in the single-accumulation
    // wavefront; job j's shadow queue is the segment [shK slotBase_j, shK (slotBase_j + cap_j))
    uint32_t shK;
    float* shRec;
    uint32_t* shVis;
    uint32_t* shQueue;
    uint32_t* shNRec;
    uint32_t* shOut;
    uint32_t* shPixel;
    float* shTotal;
};

// Exclusive prefix over jobs of the 64-slot waves of work `work` (one value per job, thread t for
// job t) into s_prefix[0 .. njobs]; returns the total. Call with the whole block.
__device__ uint32_t buildPrefix(uint32_t work, uint32_t njobs, uint32_t* s_prefix)
{
    __shared__ uint32_t s_wsum[kBxWaves];
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    uint32_t x = t < njobs ? (work + 63u) / 64u : 0u;
    uint32_t incl = x;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(incl, off, 64);
        if (lane >= off)
            incl += y;
    }
    if (lane == 63)
        s_wsum[wave] = incl;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t w = 0; w < wave; w++)
        before += s_wsum[w];
    uint32_t excl = before + incl - x;
    if (t <= njobs)
        s_prefix[t] = excl; // t == njobs (x = 0) writes the total
    __syncthreads();
    return s_prefix[njobs];
}

// buildPrefix for a block of one wave: each lane takes kPer consecutive jobs. work(j) is job j's
// live count; the prefix is over its waves of `per` rays, with per = raysPerWave(the live count of
// all jobs, traceWaves) (wave-uniform), which it also returns in *perOut.
template <class Work>
__device__ uint32_t buildPrefixWave(Work work, uint32_t njobs, uint32_t traceWaves, uint32_t* s_prefix, uint32_t* perOut)
{
    constexpr uint32_t kPer = (kBxMaxJobs + 1 + 63) / 64;
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t x[kPer];
    uint32_t rays = 0;
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++) {
        uint32_t j = lane * kPer + k;
        x[k] = j < njobs ? work(j) : 0u;
        rays += x[k];
    }
#pragma unroll
    for (uint32_t off = 32; off > 0; off >>= 1)
        rays += (uint32_t)__shfl_xor((int)rays, off, 64);
    const uint32_t per = raysPerWave(rays, traceWaves);
    *perOut = per;
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++) {
        x[k] = (x[k] + per - 1u) / per;
        sum += x[k];
    }
    uint32_t incl = sum;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(incl, off, 64);
        if (lane >= off)
            incl += y;
    }
    uint32_t run = incl - sum;
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++) {
        uint32_t j = lane * kPer + k;
        if (j <= njobs) // j == njobs (x = 0) writes the total
            s_prefix[j] = run;
        run += x[k];
    }
    __syncthreads();
    return s_prefix[njobs];
}

// The job of global wave task gw: the largest j with s_prefix[j] <= gw (jobs without work share
// their successor's start and are skipped). Wave-uniform.
__device__ __forceinline__ uint32_t findJob(const uint32_t* s_prefix, uint32_t njobs, uint32_t gw)
{
    uint32_t lo = 0, hi = njobs - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) / 2;
        if (s_prefix[mid] <= gw)
            lo = mid;
        else
            hi = mid - 1;
    }
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)lo);
}

// Appends the wave's `keep` lanes to *counter; returns this lane's index. Whole wave, converged.
__device__ __forceinline__ uint32_t waveAppend(bool keep, uint32_t* counter)
{
    uint64_t m = __ballot(keep);
    uint32_t lane = threadIdx.x & 63u;
    uint64_t lt = lane == 0 ? 0ull : ((~0ull) >> (64u - lane));
    uint32_t rank = (uint32_t)__popcll(m & lt);
    uint32_t base = 0;
    if (lane == 0 && m != 0)
        base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = (uint32_t)__shfl((int)base, 0, 64);
    return base + rank;
}

// Appends k entries per lane for the whole wave; returns this lane's first index. Whole wave,
// converged.
__device__ __forceinline__ uint32_t waveAppendN(uint32_t k, uint32_t* counter)
{
    uint32_t lane = threadIdx.x & 63u;
    uint32_t x = k;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        uint32_t y = (uint32_t)__shfl_up((int)x, off, 64);
        if (lane >= off)
            x += y;
    }
    uint32_t total = (uint32_t)__shfl((int)x, 63, 64);
    uint32_t base = 0;
    if (lane == 63 && total != 0)
        base = atomicAdd(counter, total);
    base = (uint32_t)__shfl((int)base, 63, 64);
    return base + x - k;
}

__device__ __forceinline__ uint32_t jobCapacity(const BatchJob& J) { return (J.pixels + 63u) & ~63u; }

// Start of a pass: each job's depth-0 queue (all its pixels, none yet for adaptive jobs, none
// once its samples are taken) and empty deeper queues.
__global__ void k_bx_begin(const BatchJob* __restrict__ jobs, BatchDev B)
{
    const uint32_t pass = *B.pass;
    for (uint32_t j = threadIdx.x; j < B.njobs; j += blockDim.x) {
        const BatchJob& J = jobs[j];
        uint32_t* c = B.counts + j * kCountStride;
        c[0] = pass < J.spp ? (J.P.w[32] != 0 ? 0u : J.pixels) : 0u;
        for (uint32_t d = 1; d < kCountStride; d++)
            c[d] = 0;
    }
}

// wf_select (wf_select.slang:7-17) for the adaptive jobs still sampling: active pixels, in the
// tuned depth-0 order, into the job's depth-0 queue.
__global__ __launch_bounds__(kBxBlock) void k_bx_select(const BatchJob* __restrict__ jobs, BatchDev B)
{
    __shared__ uint32_t s_prefix[kBxMaxJobs + 1];
    const uint32_t pass = *B.pass;
    uint32_t work = 0;
    if (threadIdx.x < B.njobs) {
        const BatchJob& J = jobs[threadIdx.x];
        work = (J.P.w[32] != 0 && pass < J.spp) ? J.pixels : 0u;
    }
    const uint32_t total = buildPrefix(work, B.njobs, s_prefix);
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t gw = blockIdx.x * kBxWaves + (threadIdx.x >> 6); gw < total; gw += gridDim.x * kBxWaves) {
        const uint32_t j = findJob(s_prefix, B.njobs, gw);
        const BatchJob& J = jobs[j];
        const Ctx c{J.S, J.F, J.P};
        uint32_t i = (gw - s_prefix[j]) * 64 + lane;
        uint32_t p = 0;
        bool active = false;
        if (i < J.pixels) {
            p = tilePixel(i, J.P.w[0], J.P.w[1], B.tileW, B.tileH);
            active = pixelActive(c, p);
        }
        uint32_t slot = waveAppend(active, B.counts + j * kCountStride);
        if (active && slot < J.pixels)
            B.queue0[J.slotBase + slot] = p;
    }
}

// wf_extend (wf_extend.slang:6-29) for every job's live queue at depth d
__global__ __launch_bounds__(64) void k_bx_extend(const BatchJob* __restrict__ jobs, BatchDev B, uint32_t d)
{
    __shared__ uint32_t s_prefix[kBxMaxJobs + 1];
    const uint32_t pass = *B.pass;
    uint32_t per;
    const uint32_t total =
        buildPrefixWave([&](uint32_t j) { return B.counts[j * kCountStride + d]; }, B.njobs, B.traceWaves, s_prefix, &per);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t* st = B.state[d & 1u];
    __shared__ WaveTris sh;
    for (uint32_t gw = blockIdx.x; gw < total; gw += gridDim.x) {
        const uint32_t j = findJob(s_prefix, B.njobs, gw);
        const BatchJob& J = jobs[j];
        const uint32_t count = B.counts[j * kCountStride + d];
        const uint32_t i = (gw - s_prefix[j]) * per + lane;
        const bool act = lane < per && i < count; // the wave's lanes trace together (traceWave); idle lanes help
        const uint32_t slot = J.slotBase + i;
        V3 org = v3(0.0f), dir = v3(0.0f, 0.0f, 1.0f);
        if (act && d == 0) {
            uint32_t p = J.P.w[32] != 0 ? B.queue0[slot] : tilePixel(i, J.P.w[0], J.P.w[1], B.tileW, B.tileH);
            PathState s;
            generatePath(J.P, p % J.P.w[0], p / J.P.w[0], J.P.w[6] + pass, s);
            org = s.org;
            dir = s.dir;
        } else if (act) {
            org = get3(st, B.n, 0, slot);
            dir = get3(st, B.n, 3, slot);
        }
        Hit h = traceWave<false>(J.S.bvh, org, dir, 1.0e30f, act, sh);
        if (act)
            B.hits[slot] = make_uint4(h.hit ? 1u : 0u, h.inst, h.prim, asu(h.t));
    }
}

// wf_shade (wf_shade.slang:7-28) for every job's live queue at depth d: shade, then append the
// survivor to the job's next queue or finish it into the job's film. DEFER: the plain
// integrator's NEE shadow rays are recorded into the job's shadow queue and resolved later.
template <bool DEFER>
__global__ __launch_bounds__(kBxBlock) void k_bx_shade(const BatchJob* __restrict__ jobs, BatchDev B, uint32_t d)
{
    __shared__ uint32_t s_prefix[kBxMaxJobs + 1];
    const uint32_t pass = *B.pass;
    const uint32_t total = buildPrefix(threadIdx.x < B.njobs ? B.counts[threadIdx.x * kCountStride + d] : 0u, B.njobs, s_prefix);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t* in = B.state[d & 1u];
    uint32_t* out = B.state[(d + 1) & 1u];
    for (uint32_t gw = blockIdx.x * kBxWaves + (threadIdx.x >> 6); gw < total; gw += gridDim.x * kBxWaves) {
        const uint32_t j = findJob(s_prefix, B.njobs, gw);
        const BatchJob& J = jobs[j];
        const Ctx c{J.S, J.F, J.P};
        const uint32_t count = B.counts[j * kCountStride + d];
        const uint32_t i = (gw - s_prefix[j]) * 64 + lane;
        const uint32_t slot = J.slotBase + i;
        bool keep = false;
        PathState st;
        uint32_t pixel = 0, nrec = 0;
        if (i < count) {
            if (d == 0) {
                pixel = J.P.w[32] != 0 ? B.queue0[slot] : tilePixel(i, J.P.w[0], J.P.w[1], B.tileW, B.tileH);
                generatePath(J.P, pixel % J.P.w[0], pixel / J.P.w[0], J.P.w[6] + pass, st);
            } else {
                loadPath<false>(in, B.n, slot, st, pixel);
            }
            uint4 hw = B.hits[slot];
            Hit h;
            h.hit = hw.x != 0;
            h.inst = hw.y;
            h.prim = hw.z;
            h.t = asf(hw.w);
            h.b1 = h.b2 = 0.0f;
            h.gid = 0;
            if (DEFER) {
                ShadowSink sink{B.shRec, B.n, B.shK, slot, 0u};
                keep = shadePath<false, false, 0, true>(c, st, h, d, pixel, nullptr, &sink) && d + 1 < J.P.w[7];
                nrec = sink.count;
                B.shNRec[slot] = nrec;
            } else {
                keep = shadePath<false, false>(c, st, h, d, pixel) && d + 1 < J.P.w[7];
            }
            if (!keep && nrec == 0)
                finishPath<false>(c, st, pixel);
        }
        uint32_t k = waveAppend(keep, B.counts + j * kCountStride + d + 1);
        if (keep && k < count) // survivors of depth d never outnumber its count
            storePath<false>(out, B.n, J.slotBase + k, st, pixel);
        if (DEFER) {
            uint32_t q = waveAppendN(nrec, B.counts + j * kCountStride + kShadowOff + d);
            if (nrec > 0) {
                B.shOut[slot] = keep ? J.slotBase + k : 0xffffffffu;
                B.shPixel[slot] = pixel;
                put3(reinterpret_cast<uint32_t*>(B.shTotal), B.n, 0, slot, st.parts.total);
                const uint32_t cap = B.shK * jobCapacity(J);
                for (uint32_t r = 0; r < nrec; r++)
                    if (q + r < cap)
                        B.shQueue[B.shK * J.slotBase + q + r] = r * B.n + slot;
            }
        }
    }
}

// The shadow rays every job recorded at depth d (occluded, pt_core.slang:580-613), one per lane.
__global__ __launch_bounds__(64) void k_bx_shadow(const BatchJob* __restrict__ jobs, BatchDev B, uint32_t d)
{
    __shared__ uint32_t s_prefix[kBxMaxJobs + 1];
    uint32_t per;
    const uint32_t total = buildPrefixWave([&](uint32_t j) { return B.counts[j * kCountStride + kShadowOff + d]; }, B.njobs,
                                           B.traceWaves, s_prefix, &per);
    const uint32_t lane = threadIdx.x & 63u;
    const size_t stride = (size_t)B.shK * B.n;
    const Frame F{};
    __shared__ WaveTris sh;
    for (uint32_t gw = blockIdx.x; gw < total; gw += gridDim.x) {
        const uint32_t j = findJob(s_prefix, B.njobs, gw);
        const BatchJob& J = jobs[j];
        const uint32_t count = B.counts[j * kCountStride + kShadowOff + d];
        const uint32_t i = (gw - s_prefix[j]) * per + lane;
        const bool act = lane < per && i < count;
        const Ctx c{J.S, F, J.P};
        uint32_t rid = act ? B.shQueue[B.shK * J.slotBase + i] : 0u;
        const float* r = B.shRec + rid;
        V3 o = v3(0.0f), dir = v3(0.0f, 0.0f, 1.0f);
        float tmax = 1.0e30f;
        if (act) {
            o = v3(r[0], r[stride], r[2 * stride]);
            dir = v3(r[3 * stride], r[4 * stride], r[5 * stride]);
            tmax = r[6 * stride];
        }
        if (!J.S.translucent) { // wave-uniform: the job is the wave's
            Hit h = traceWave<true>(J.S.bvh, o, dir, tmax, act, sh);
            if (act)
                B.shVis[rid] = h.hit ? 1u : 0u;
        } else if (act) {
            B.shVis[rid] = occluded(c, o, dir, tmax) ? 1u : 0u;
        }
    }
}

// Adds each pending path's visible shadow contributions in record order, then stores the total in
// its next slot or finishes the path into its job's film.
__global__ __launch_bounds__(kBxBlock) void k_bx_resolve(const BatchJob* __restrict__ jobs, BatchDev B, uint32_t d)
{
    __shared__ uint32_t s_prefix[kBxMaxJobs + 1];
    const uint32_t total = buildPrefix(threadIdx.x < B.njobs ? B.counts[threadIdx.x * kCountStride + d] : 0u, B.njobs, s_prefix);
    const uint32_t lane = threadIdx.x & 63u;
    const size_t stride = (size_t)B.shK * B.n;
    uint32_t* out = B.state[(d + 1) & 1u];
    for (uint32_t gw = blockIdx.x * kBxWaves + (threadIdx.x >> 6); gw < total; gw += gridDim.x * kBxWaves) {
        const uint32_t j = findJob(s_prefix, B.njobs, gw);
        const BatchJob& J = jobs[j];
        const uint32_t count = B.counts[j * kCountStride + d];
        const uint32_t i = (gw - s_prefix[j]) * 64 + lane;
        if (i >= count)
            continue;
        const uint32_t slot = J.slotBase + i;
        const uint32_t nrec = B.shNRec[slot];
        if (nrec == 0)
            continue;
        V3 sum = get3(reinterpret_cast<const uint32_t*>(B.shTotal), B.n, 0, slot);
        for (uint32_t k = 0; k < nrec; k++) {
            uint32_t rid = k * B.n + slot;
            if (B.shVis[rid] == 0) {
                const float* r = B.shRec + rid;
                sum = sum + v3(r[7 * stride], r[8 * stride], r[9 * stride]);
            }
        }
        uint32_t os = B.shOut[slot];
        if (os != 0xffffffffu) {
            put3(out, B.n, 13, os, sum);
        } else {
            const Ctx c{J.S, J.F, J.P};
            PathState st;
            st.parts = Parts{sum, v3(0.0f), v3(0.0f), v3(0.0f), v3(0.0f), v3(0.0f)};
            st.visLit = 0.0f;
            st.visTotal = 0.0f;
            finishPath<false>(c, st, B.shPixel[slot]);
        }
    }
}

// End of a pass: per job, the paths traced and (first time its depth-0 queue is empty) the pass
// at which it had nothing left to do; then the next pass.
__global__ void k_bx_end(const BatchJob* __restrict__ jobs, BatchDev B)
{
    __shared__ unsigned long long s_sum;
    const uint32_t pass = *B.pass;
    if (threadIdx.x == 0)
        s_sum = 0;
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < B.njobs; j += blockDim.x) {
        uint32_t c0 = B.counts[j * kCountStride];
        B.traced[j] += c0;
        if (c0 == 0 && B.donePass[j] == kNotDone)
            B.donePass[j] = pass;
        atomicAdd(&s_sum, (unsigned long long)c0);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (pass < B.maxPasses)
            B.passPaths[pass] = s_sum;
        *B.pass = pass + 1;
    }
    (void)jobs;
}

} // namespace

} // namespace vh

using namespace vh;

struct vh_batch {
    vh_context* ctx = nullptr;
    std::vector<vh_batch_job> jobs;
    std::vector<uint32_t> slotBase;
    std::vector<DevBuf> films, stats;
    DevBuf jobTable, state0, state1, hits, queue0, counts, pass, donePass, traced, passPaths;
    DevBuf shRec, shVis, shQueue, shNRec, shOut, shPixel, shTotal; // deferred shadow rays
    uint32_t shK = 0;
    uint32_t n = 0;
    uint32_t maxSpp = 0;
    bool anyAdaptive = false;
    vh_batch_tuning tuning{1024, 8, 8};
    bool autotune = true;
    std::vector<vh_batch_tuning> tried;
    std::vector<double> triedMsPerMpath;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    vh_batch_tuning graphTuning{};
    uint32_t passesRun = 0;
    double elapsed = 0.0;                  // GPU seconds of all runs so far
    std::vector<double> passEnd;           // end of pass k, seconds since the first run started
    uint32_t passesHost = 0;

    BatchDev dev() const
    {
        BatchDev B;
        B.state[0] = state0.as<uint32_t>();
        B.state[1] = state1.as<uint32_t>();
        B.hits = hits.as<uint4>();
        B.queue0 = queue0.as<uint32_t>();
        B.counts = counts.as<uint32_t>();
        B.pass = pass.as<uint32_t>();
        B.donePass = donePass.as<uint32_t>();
        B.traced = traced.as<unsigned long long>();
        B.passPaths = passPaths.as<unsigned long long>();
        B.maxPasses = maxSpp;
        B.n = n;
        B.njobs = (uint32_t)jobs.size();
        B.tileW = tuning.tile_width;
        B.tileH = tuning.tile_height;
        B.traceWaves = ctx->cus * kTraceWavesPerCu;
        B.shK = shK;
        B.shRec = shRec.as<float>();
        B.shVis = shVis.as<uint32_t>();
        B.shQueue = shQueue.as<uint32_t>();
        B.shNRec = shNRec.as<uint32_t>();
        B.shOut = shOut.as<uint32_t>();
        B.shPixel = shPixel.as<uint32_t>();
        B.shTotal = shTotal.as<float>();
        return B;
    }
    void dropGraph()
    {
        if (graphExec)
            (void)hipGraphExecDestroy(graphExec);
        if (graph)
            (void)hipGraphDestroy(graph);
        graphExec = nullptr;
        graph = nullptr;
    }
    ~vh_batch()
    {
        if (ctx)
            (void)hipStreamSynchronize(ctx->stream);
        dropGraph();
    }
};

namespace {

#define TRY(x)            \
    do {                  \
        int rc_ = (x);    \
        if (rc_ != VH_OK) \
            return rc_;   \
    } while (0)
#define HIPTRY(x) TRY(check_hip((x), #x))

// One sample pass of every job, as launches on the context stream.
int issue_pass(vh_batch* b)
{
    hipStream_t s = b->ctx->stream;
    const BatchJob* jobs = b->jobTable.as<BatchJob>();
    BatchDev B = b->dev();
    uint32_t grid = b->tuning.grid_blocks;
    k_bx_begin<<<1, kBxBlock, 0, s>>>(jobs, B);
    if (b->anyAdaptive)
        k_bx_select<<<grid, kBxBlock, 0, s>>>(jobs, B);
    const uint32_t tgrid = std::max(grid * kBxTraceScale, B.traceWaves); // one-wave blocks
    for (uint32_t d = 0; d < kMaxDepth; d++) {
        k_bx_extend<<<tgrid, 64, 0, s>>>(jobs, B, d);
        if (B.shK > 0) {
            k_bx_shade<true><<<grid, kBxBlock, 0, s>>>(jobs, B, d);
            k_bx_shadow<<<tgrid, 64, 0, s>>>(jobs, B, d);
            k_bx_resolve<<<grid, kBxBlock, 0, s>>>(jobs, B, d);
        } else {
            k_bx_shade<false><<<grid, kBxBlock, 0, s>>>(jobs, B, d);
        }
    }
    k_bx_end<<<1, kBxBlock, 0, s>>>(jobs, B);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? VH_OK : fail(VH_ERR_KERNEL, std::string("batch pass launch: ") + hipGetErrorString(e));
}

// Captures one pass with the current tuning into an executable graph.
int capture_pass(vh_batch* b)
{
    hipStream_t s = b->ctx->stream;
    b->dropGraph();
    HIPTRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = issue_pass(b);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    TRY(rc);
    HIPTRY(e);
    b->graph = g;
    HIPTRY(hipGraphInstantiate(&b->graphExec, g, nullptr, nullptr, 0));
    b->graphTuning = b->tuning;
    return VH_OK;
}

bool same_tuning(const vh_batch_tuning& a, const vh_batch_tuning& c)
{
    return a.grid_blocks == c.grid_blocks && a.tile_width == c.tile_width && a.tile_height == c.tile_height;
}

int check_tuning(const vh_batch_tuning& t)
{
    if (t.grid_blocks == 0 || t.grid_blocks > 65536)
        return fail(VH_ERR_INVALID_ARGUMENT, "grid_blocks must be in [1, 65536]");
    if ((t.tile_width == 0) != (t.tile_height == 0) || t.tile_width > 4096 || t.tile_height > 4096)
        return fail(VH_ERR_INVALID_ARGUMENT, "tiles are w x h with both in [1, 4096], or 0 x 0 for scanlines");
    return VH_OK;
}

// Candidates of the online tuning, most promising first (only as many are tried as the batch has
// passes in which every job still takes every pixel).
std::vector<vh_batch_tuning> candidates()
{
    return {{1024, 8, 8}, {2048, 8, 8}, {1024, 0, 0}, {4096, 8, 8}, {512, 8, 8}, {2048, 32, 2},
            {1024, 16, 4}, {2048, 0, 0}, {4096, 0, 0}, {1024, 32, 2}, {512, 0, 0}, {4096, 32, 2}};
}

} // namespace

extern "C" {

int vh_batch_create(vh_context* ctx, const vh_batch_job* jobs, uint32_t count, vh_batch** out)
{
    if (!ctx || !out || (count && !jobs))
        return fail(VH_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (count == 0 || count > kBxMaxJobs)
        return fail(VH_ERR_INVALID_ARGUMENT, "a batch holds 1 to 255 jobs");
    std::unique_ptr<vh_batch> b(new vh_batch());
    b->ctx = ctx;
    b->jobs.assign(jobs, jobs + count);
    uint64_t slots = 0, filmBytes = 0;
    for (uint32_t j = 0; j < count; j++) {
        const vh_batch_job& J = jobs[j];
        if (!J.renderer || J.renderer->ctx != ctx)
            return fail(VH_ERR_INVALID_ARGUMENT, "job " + std::to_string(j) + ": its renderer belongs to another context");
        if (J.width == 0 || J.height == 0)
            return fail(VH_ERR_SCENE, "job " + std::to_string(j) + ": image size must be non-zero");
        if (J.spp == 0)
            return fail(VH_ERR_INVALID_ARGUMENT, "job " + std::to_string(j) + ": spp must be positive");
        if (J.adaptive.threshold < 0.0f || !(J.adaptive.threshold == J.adaptive.threshold))
            return fail(VH_ERR_INVALID_ARGUMENT, "job " + std::to_string(j) + ": bad adaptive threshold");
        uint64_t px = (uint64_t)J.width * J.height;
        b->slotBase.push_back((uint32_t)slots);
        slots += (px + 63) / 64 * 64;
        filmBytes += 24 * px * (J.adaptive.threshold > 0.0f ? 2 : 1);
        b->maxSpp = std::max(b->maxSpp, J.spp);
        b->anyAdaptive |= J.adaptive.threshold > 0.0f;
    }
    if (slots >= (1ull << 31))
        return fail(VH_ERR_INVALID_ARGUMENT, "a batch holds fewer than 2^31 pixels");
    b->n = (uint32_t)slots;
    const uint64_t n = slots;
    // deferred shadow rays: one record per delta light plus one sample, for the job with most lights
    uint32_t K = 0;
    for (uint32_t j = 0; j < count; j++)
        K = std::max(K, jobs[j].renderer->info.light_count + 1);
    const bool defer = K <= kMaxShadow && (uint64_t)K * n < (1ull << 31);
    uint64_t shBytes = defer ? (4ull * (kShadowWords + 2) * K * n + 4ull * 6 * n) : 0;
    uint64_t bytes = filmBytes + n * (2 * 17 * 4 + 16 + 4) + count * (sizeof(BatchJob) + 4 * kCountStride + 16) + 8ull * b->maxSpp + shBytes;
    TRY(check_fits(bytes));
    if (defer) {
        HIPTRY(b->shRec.alloc(4ull * kShadowWords * K * n));
        HIPTRY(b->shVis.alloc(4ull * K * n));
        HIPTRY(b->shQueue.alloc(4ull * K * n));
        HIPTRY(b->shNRec.alloc(4 * n));
        HIPTRY(b->shOut.alloc(4 * n));
        HIPTRY(b->shPixel.alloc(4 * n));
        HIPTRY(b->shTotal.alloc(12 * n));
        b->shK = K;
    }
    HIPTRY(b->state0.alloc(4 * 17 * n));
    HIPTRY(b->state1.alloc(4 * 17 * n));
    HIPTRY(b->hits.alloc(16 * n));
    HIPTRY(b->queue0.alloc(4 * n));
    HIPTRY(b->counts.alloc(4ull * kCountStride * count));
    HIPTRY(b->pass.alloc(4));
    HIPTRY(b->donePass.alloc(4ull * count));
    HIPTRY(b->traced.alloc(8ull * count));
    HIPTRY(b->passPaths.alloc(8ull * b->maxSpp));
    std::vector<BatchJob> table(count);
    b->films.resize(count);
    b->stats.resize(count);
    for (uint32_t j = 0; j < count; j++) {
        const vh_batch_job& J = jobs[j];
        uint64_t px = (uint64_t)J.width * J.height;
        bool adaptive = J.adaptive.threshold > 0.0f;
        HIPTRY(b->films[j].alloc(24 * px));
        if (adaptive)
            HIPTRY(b->stats[j].alloc(24 * px));
        BatchJob& T = table[j];
        T.S = J.renderer->scene();
        T.S.bvh.counters = nullptr;
        T.F = Frame{b->films[j].as<double>(), adaptive ? b->stats[j].as<double>() : nullptr, nullptr, nullptr};
        T.P = base_params(J.renderer, J.width, J.height, false, adaptive, J.adaptive, false, vh_restir{});
        T.P.seedMix = J.seed == 0 ? 0u : pcg(J.seed ^ 0x5851F42Du);
        T.P.w[6] = 0;
        T.slotBase = b->slotBase[j];
        T.pixels = (uint32_t)px;
        T.spp = J.spp;
        T.pad = 0;
    }
    HIPTRY(b->jobTable.alloc(sizeof(BatchJob) * count));
    HIPTRY(hipMemcpy(b->jobTable.p, table.data(), sizeof(BatchJob) * count, hipMemcpyHostToDevice));
    *out = b.release();
    return vh_batch_reset(*out);
}

void vh_batch_destroy(vh_batch* b) { delete b; }

int vh_batch_reset(vh_batch* b)
{
    if (!b)
        return fail(VH_ERR_INVALID_ARGUMENT, "null batch");
    hipStream_t s = b->ctx->stream;
    for (size_t j = 0; j < b->jobs.size(); j++) {
        HIPTRY(hipMemsetAsync(b->films[j].p, 0, b->films[j].bytes, s));
        if (b->stats[j].p)
            HIPTRY(hipMemsetAsync(b->stats[j].p, 0, b->stats[j].bytes, s));
    }
    HIPTRY(hipMemsetAsync(b->pass.p, 0, 4, s));
    HIPTRY(hipMemsetAsync(b->donePass.p, 0xff, 4 * b->jobs.size(), s));
    HIPTRY(hipMemsetAsync(b->traced.p, 0, 8 * b->jobs.size(), s));
    HIPTRY(hipMemsetAsync(b->passPaths.p, 0, 8ull * b->maxSpp, s));
    HIPTRY(hipMemsetAsync(b->counts.p, 0, b->counts.bytes, s));
    HIPTRY(hipStreamSynchronize(s));
    b->passesRun = 0;
    b->elapsed = 0.0;
    b->passEnd.clear();
    b->tried.clear();
    b->triedMsPerMpath.clear();
    return VH_OK;
}

int vh_batch_set_tuning(vh_batch* b, const vh_batch_tuning* t)
{
    if (!b)
        return fail(VH_ERR_INVALID_ARGUMENT, "null batch");
    if (!t) {
        b->autotune = true;
        return VH_OK;
    }
    TRY(check_tuning(*t));
    b->tuning = *t;
    b->autotune = false;
    return VH_OK;
}

int vh_batch_get_tuning(const vh_batch* b, vh_batch_tuning* t)
{
    if (!b || !t)
        return fail(VH_ERR_INVALID_ARGUMENT, "null argument");
    *t = b->tuning;
    return VH_OK;
}

int vh_batch_run(vh_batch* b, int graph, uint32_t max_passes)
{
    if (!b)
        return fail(VH_ERR_INVALID_ARGUMENT, "null batch");
    hipStream_t s = b->ctx->stream;
    uint32_t left = b->maxSpp > b->passesRun ? b->maxSpp - b->passesRun : 0u;
    if (max_passes)
        left = std::min(left, max_passes);
    if (left == 0)
        return VH_OK;
    // candidates of the online tuning: one per pass while every job is still full (the first
    // min_spp passes of an adaptive job take every pixel), compared per path traced
    std::vector<vh_batch_tuning> cand;
    if (b->autotune && b->passesRun == 0)
        cand = candidates();
    // passes in which every job still takes every pixel: the same work for each candidate
    uint32_t full = left > 1 ? left - 1 : 0;
    for (const vh_batch_job& J : b->jobs)
        full = std::min(full, J.adaptive.threshold > 0.0f ? std::min(J.spp, J.adaptive.min_spp) : J.spp);
    const uint32_t tuneN = (uint32_t)std::min<size_t>(cand.size(), full);
    std::vector<hipEvent_t> ev(left + 1, nullptr);
    for (auto& e : ev)
        HIPTRY(hipEventCreate(&e));
    auto cleanup = [&]() {
        for (auto e : ev)
            (void)hipEventDestroy(e);
    };
    int rc = check_hip(hipEventRecord(ev[0], s), "hipEventRecord");
    uint32_t k = 0;
    for (; k < tuneN && rc == VH_OK; k++) {
        b->tuning = cand[k];
        rc = issue_pass(b);
        if (rc == VH_OK)
            rc = check_hip(hipEventRecord(ev[k + 1], s), "hipEventRecord");
    }
    if (tuneN > 0 && rc == VH_OK) {
        // pick the fastest per traced path
        rc = check_hip(hipEventSynchronize(ev[tuneN]), "hipEventSynchronize");
        std::vector<unsigned long long> paths(tuneN);
        if (rc == VH_OK)
            rc = check_hip(hipMemcpy(paths.data(), b->passPaths.p, 8ull * tuneN, hipMemcpyDeviceToHost), "hipMemcpy");
        double best = 1e300;
        for (uint32_t i = 0; i < tuneN && rc == VH_OK; i++) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, ev[i], ev[i + 1]);
            double perM = paths[i] ? (double)ms / ((double)paths[i] * 1.0e-6) : 1e300;
            b->tried.push_back(cand[i]);
            b->triedMsPerMpath.push_back(perM);
            if (perM < best) {
                best = perM;
                b->tuning = cand[i];
            }
        }
        b->autotune = false;
    }
    const bool useGraph = graph != 0;
    if (rc == VH_OK && useGraph && (!b->graphExec || !same_tuning(b->graphTuning, b->tuning)))
        rc = capture_pass(b);
    for (; k < left && rc == VH_OK; k++) {
        if (useGraph)
            rc = check_hip(hipGraphLaunch(b->graphExec, s), "hipGraphLaunch");
        else
            rc = issue_pass(b);
        if (rc == VH_OK)
            rc = check_hip(hipEventRecord(ev[k + 1], s), "hipEventRecord");
    }
    if (rc == VH_OK)
        rc = check_hip(hipEventSynchronize(ev[left]), "hipEventSynchronize");
    if (rc != VH_OK) {
        (void)hipStreamSynchronize(s);
        cleanup();
        return rc;
    }
    double base = b->elapsed;
    for (uint32_t i = 1; i <= left; i++) {
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, ev[0], ev[i]);
        b->passEnd.push_back(base + (double)ms * 1.0e-3);
    }
    b->elapsed = b->passEnd.back();
    b->passesRun += left;
    cleanup();
    return VH_OK;
}

int vh_batch_job_stats(const vh_batch* b, uint32_t job, double out[4])
{
    if (!b || !out || job >= b->jobs.size())
        return fail(VH_ERR_INVALID_ARGUMENT, "bad argument");
    uint32_t done = 0;
    unsigned long long traced = 0;
    HIPTRY(hipMemcpy(&done, b->donePass.as<uint32_t>() + job, 4, hipMemcpyDeviceToHost));
    HIPTRY(hipMemcpy(&traced, b->traced.as<unsigned long long>() + job, 8, hipMemcpyDeviceToHost));
    // a job whose last pass was the run's last is done at that pass's end
    uint32_t passes = done != kNotDone ? done : (b->passesRun >= b->jobs[job].spp ? b->jobs[job].spp : kNotDone);
    out[0] = (passes != kNotDone && passes >= 1 && passes <= b->passEnd.size()) ? b->passEnd[passes - 1] : -1.0;
    out[1] = passes != kNotDone ? (double)passes : (double)b->passesRun;
    out[2] = (double)traced;
    out[3] = b->elapsed;
    return VH_OK;
}

int vh_batch_tuning_report(const vh_batch* b, vh_batch_tuning* tried, double* ms_per_mpath, uint32_t cap, uint32_t* count)
{
    if (!b || !count)
        return fail(VH_ERR_INVALID_ARGUMENT, "null argument");
    uint32_t n = (uint32_t)b->tried.size();
    *count = n;
    for (uint32_t i = 0; i < n && i < cap; i++) {
        if (tried)
            tried[i] = b->tried[i];
        if (ms_per_mpath)
            ms_per_mpath[i] = b->triedMsPerMpath[i];
    }
    return VH_OK;
}

int vh_batch_read_film(const vh_batch* b, uint32_t job, double* film)
{
    if (!b || !film || job >= b->jobs.size())
        return fail(VH_ERR_INVALID_ARGUMENT, "bad argument");
    HIPTRY(hipStreamSynchronize(b->ctx->stream));
    HIPTRY(hipMemcpy(film, b->films[job].p, b->films[job].bytes, hipMemcpyDeviceToHost));
    return VH_OK;
}

int vh_batch_read_stats(const vh_batch* b, uint32_t job, double* stats)
{
    if (!b || !stats || job >= b->jobs.size())
        return fail(VH_ERR_INVALID_ARGUMENT, "bad argument");
    if (!b->stats[job].p)
        return fail(VH_ERR_UNSUPPORTED, "the job has no adaptive sampling");
    HIPTRY(hipStreamSynchronize(b->ctx->stream));
    HIPTRY(hipMemcpy(stats, b->stats[job].p, b->stats[job].bytes, hipMemcpyDeviceToHost));
    return VH_OK;
}

int vh_batch_mean(const vh_batch* b, uint32_t job, float* out)
{
    if (!b || !out || job >= b->jobs.size())
        return fail(VH_ERR_INVALID_ARGUMENT, "bad argument");
    uint64_t n = (uint64_t)b->jobs[job].width * b->jobs[job].height;
    std::vector<double> film(3 * n), st;
    TRY(vh_batch_read_film(b, job, film.data()));
    if (b->stats[job].p) {
        st.resize(3 * n);
        TRY(vh_batch_read_stats(b, job, st.data()));
    }
    double stats[4];
    TRY(vh_batch_job_stats(b, job, stats));
    double samples = std::max(stats[1], 1.0);
    for (uint64_t p = 0; p < n; p++) {
        double inv = st.empty() ? 1.0 / samples : 1.0 / std::max(st[3 * p], 1.0);
        for (int k = 0; k < 3; k++)
            out[3 * p + k] = (float)(film[3 * p + k] * inv);
    }
    return VH_OK;
}

} // extern "C"
