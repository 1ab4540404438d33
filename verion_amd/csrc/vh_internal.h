// vh_internal.h - host-side types shared by the translation units of libverion_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/verion_hip.h"
#include "vh_shade.h"

namespace vh {

// Records the error of the calling thread and returns `status`.
int fail(int status, const std::string& msg);
int fail_oom(uint64_t needed, uint64_t available);

// An owned device allocation.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr, o.bytes = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void release()
    {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t alloc(size_t n)
    {
        release();
        if (n == 0)
            n = 16;
        hipError_t e = hipMalloc(&p, n);
        if (e == hipSuccess)
            bytes = n;
        else
            p = nullptr;
        return e;
    }
    template <class T>
    T* as() const
    {
        return static_cast<T*>(p);
    }
};

// The scene tables after upload, with what the BVH build needs.
struct DeviceTables {
    const float* positions;
    const uint32_t* indices;
    const uint32_t* instances;
    const uint32_t* instStart; // first global triangle of each instance
    uint32_t instanceCount;
    uint32_t triangleCount;
};

struct BvhBuild {
    DevBuf nodes_buf;
    DevBuf tris;
    DevBuf perm;                  // leaf slot -> global triangle id
    DevBuf ids;                   // global triangle id -> (instance, primitive)
    std::vector<uint32_t> levels; // wide node ids of level j: levels[j] .. levels[j + 1] - 1
    double sah = 0.0;             // surface-area cost of the tree as built
    uint32_t nodes = 0;
    uint32_t leaves = 0;
    uint32_t depth = 0;
    uint32_t ntri = 0;
    float pad = 0.0f;
    float lo[3] = {0.0f, 0.0f, 0.0f}, hi[3] = {1.0f, 1.0f, 1.0f}; // scene box of all vertices
};

// A refit whose surface-area cost exceeds the built tree's by this factor is rebuilt instead.
constexpr double kRefitLimit = 1.5;

int build_bvh(BvhBuild& out, const DeviceTables& T, hipStream_t s);
// New world-space triangles in the built leaf order and new boxes, bottom-up; ratio = the
// refit tree's surface-area cost over the built one's. The triangle count must not change.
int refit_bvh(BvhBuild& b, const DeviceTables& T, hipStream_t s, double& ratio);

// Wavefront buffers of one accumulation (slot-indexed structure of arrays, two copies for the
// ping-pong between depths).
struct Wavefront {
    uint32_t* state[2]; // words * n each
    uint4* hits;        // n
    uint32_t* queue0;   // n, adaptive sampling's depth-0 queue
    uint32_t* counts;   // 9
    uint32_t n;
    uint32_t words;
    uint32_t tileW = 8, tileH = 8; // depth-0 pixel order: tiles, or 0 x 0 for scanlines
    // resident waves of the trace kernels: a queue of fewer than 64 x traceWaves rays is spread
    // over traceWaves waves (0: 64 rays per wave)
    uint32_t traceWaves = 0;
    // deferred shadow rays of the plain integrator (shK = 0: shade traces them itself)
    uint32_t shK = 0;         // records per slot
    float* shRec = nullptr;   // kShadowWords x shK x n
    uint32_t* shVis = nullptr;   // shK x n: 1 = occluded
    uint32_t* shQueue = nullptr; // shK x n: record ids (k n + i) to trace, per depth
    uint32_t* shNRec = nullptr;  // n: records of the path in slot i (0: nothing pending)
    uint32_t* shOut = nullptr;   // n: the path's slot at the next depth, ~0 if it ends
    uint32_t* shPixel = nullptr; // n
    float* shTotal = nullptr;    // 3 x n: path total before its shadow contributions
};

// The device buffers behind one Wavefront: the state of one sample pass in flight.
struct WfBuffers {
    DevBuf state0, state1, hits, queue0, counts;
    DevBuf shRec, shVis, shQueue, shNRec, shOut, shPixel, shTotal; // deferred shadow rays
    uint32_t shK = 0;
    bool ready = false;
    uint64_t bytes = 0;
};
// Sample passes of one accumulate call that the plain wavefront keeps in flight at once, each on
// its own stream with its own wavefront buffers and FP64 film, added to the accumulation's film in
// sample order (vh_set_wavefront_pipeline). A pass's deepest depths hold a few thousand paths
// whose launches last as long as their slowest chain of dependent fetches; the next passes fill
// the GPU meanwhile.
constexpr uint32_t kMaxPipeline = 4;
// counts[] layout (kCountWords words): live paths per depth at 0..kMaxDepth, shadow records per
// depth at kShadowCounts + d
constexpr uint32_t kShadowCounts = 16;
constexpr uint32_t kCountWords = 32;

// Kernel launchers (vh_render.hip)
struct LaunchCfg {
    hipStream_t stream;
    uint32_t wfGrid;   // blocks of the grid-stride wavefront kernels
};

constexpr uint32_t kMaxDepth = 8;           // pack.rs:13
constexpr uint32_t kSamplesPerDispatch = 8; // vulkan renderer.rs:29
constexpr uint32_t kCounterBanks = 64;      // banks of the traversal counters (vh_trace.h)

// Stage timing of the wavefront launches: an event after each launch, labelled with the stage the
// time since the previous event belongs to. kStageTrace: closest and shadow rays in one launch.
enum Stage : uint8_t { kStageExtend, kStageShade, kStageShadow, kStageResolve, kStageTrace, kStages };
// A pool of timing events reused from call to call: a call's first marks create events, later
// calls only record them, so no event is created between the launches of a timed pass.
struct StageMarks {
    std::vector<hipEvent_t> ev;
    std::vector<uint8_t> label; // label[k]: stage of ev[k-1] -> ev[k]; kStages: no interval ends at ev[k]
    size_t used = 0;            // events recorded since reset()
    StageMarks() = default;
    StageMarks(const StageMarks&) = delete;
    StageMarks& operator=(const StageMarks&) = delete;
    void reset()
    {
        used = 0;
        label.clear();
    }
    // Events a call will record, created ahead of its launches.
    hipError_t reserve(size_t n)
    {
        while (ev.size() < n) {
            hipEvent_t e;
            hipError_t err = hipEventCreate(&e);
            if (err != hipSuccess)
                return err;
            ev.push_back(e);
        }
        return hipSuccess;
    }
    hipError_t mark(hipStream_t s, uint8_t stage)
    {
        if (used == ev.size()) {
            hipError_t err = reserve(used + 1);
            if (err != hipSuccess)
                return err;
        }
        label.push_back(stage);
        return hipEventRecord(ev[used++], s);
    }
    ~StageMarks()
    {
        for (hipEvent_t e : ev)
            (void)hipEventDestroy(e);
    }
};

// blocks of `block` threads for n threads of work, at least 1 and at most `cap`
inline uint32_t blocks_for(uint64_t n, uint32_t block, uint32_t cap)
{
    uint64_t g = (n + block - 1) / block;
    return (uint32_t)(g < 1 ? 1 : (g > cap ? cap : g));
}

int launch_megakernel(const Scene& S, const Frame& F, const Params& P, uint32_t samples, bool restir, bool split, hipStream_t s);
// Plain megakernel with `samples` samples in parallel on persistent waves (kPersistWavesPerCu per
// CU; staged: 3 floats per pixel and sample; next: a work counter word; shRec: kShadowWords x shK
// floats per resident lane for the deferred shadow rays, or null to trace them in the shading).
constexpr uint32_t kPersistWavesPerCu = 16; // 4 per SIMD, the kernel's occupancy target
// Waves per CU a trace kernel spreads a short queue over (Wavefront::traceWaves): at 1080p on the
// procedural scene 8, 16 and 20 per CU trace within 1% of each other, 4 per CU and 32 per CU are
// 2% and 7% slower, and 64 rays to every wave (no spreading) 10% slower.
#ifndef VH_TRACE_WAVES_PER_CU
#define VH_TRACE_WAVES_PER_CU 16
#endif
constexpr uint32_t kTraceWavesPerCu = VH_TRACE_WAVES_PER_CU;
int launch_megakernel_samples(const Scene& S, const Frame& F, const Params& P, uint32_t samples, float* staged, uint32_t* next,
                              float* shRec, uint32_t shK, uint32_t cus, hipStream_t s);
// Paths of one persistent-megakernel dispatch: a call's samples run up to this many (pixel, sample)
// paths per dispatch in parallel (12 bytes of staged radiance each: 384 MB at most), so each
// dispatch's tail (its last, longest paths) and its film addition are paid once per many samples.
// The reference's 5M-triangle scene at 1024x768, 1,024 spp: 0.413 s with 2M paths per dispatch,
// 0.201 s with 8M, 0.174 s with 16M, 0.162 s with 32M.
#ifndef VH_PARALLEL_PATHS
#define VH_PARALLEL_PATHS (1u << 25)
#endif
constexpr uint64_t kParallelPaths = VH_PARALLEL_PATHS;
int launch_wf_sample(const Scene& S, const Frame& F, const Params& P, const Wavefront& W, bool restir, bool split, const LaunchCfg& cfg,
                     StageMarks* marks);
// film[i] += pass[i] and pass[i] = 0 over n doubles: a pass in flight's FP64 film (each pixel's
// 0 + radiance, exact) added in the order the passes were issued, which gives the film's bits of
// passes run one at a time.
int launch_film_add_pass(double* film, double* pass, uint64_t n, hipStream_t s);
// lightDirs[4 li .. 4 li + 2] = -normalize(direction words of light li), for the count delta lights
int launch_light_dirs(const uint32_t* lights, uint32_t count, float* out, hipStream_t s);
int launch_primary_hits(const Scene& S, const Params& P, uint4* out, hipStream_t s);
int launch_nrc_train_paths(const Scene& S, const Params& P, const Wavefront& W, uint32_t* trainPix, uint32_t nTrain, uint32_t npix,
                           uint32_t frame, float4* vertices, float4* samples, uint32_t* sampleCount, uint32_t sampleCap,
                           const LaunchCfg& cfg);
int launch_wf_reset(const Params& P, const Wavefront& W, hipStream_t s);
// One depth of a pass driven depth by depth: extend, then (shade) the shading with its shadow
// rays and resolve; leaveShadows: the deferred shadow rays and resolve are left to
// launch_wf_trace_resolve of depth d + 1.
int launch_wf_depth(const Scene& S, const Frame& F, const Params& P, const Wavefront& W, uint32_t d, bool shade, const LaunchCfg& cfg,
                    bool leaveShadows = false);
// Depth d's closest rays with depth d - 1's deferred shadow rays in one launch, then depth d - 1's
// resolve (d >= 1, opaque scene, depth d - 1 shaded with leaveShadows).
int launch_wf_trace_resolve(const Scene& S, const Frame& F, const Params& P, const Wavefront& W, uint32_t d, const LaunchCfg& cfg);
int launch_nrc_query_shade(const Scene& S, const Frame& F, const Params& P, const Wavefront& W, float4* queries, uint32_t d,
                           const LaunchCfg& cfg);
int launch_trace_rays(const Scene& S, const float* rays, uint64_t n, float tmax, uint4* out, hipStream_t s);
int launch_normals(const Scene& S, const Params& P, float* out3, hipStream_t s);
int launch_aovs(const Scene& S, const Params& P, float4* albedoDepth, uint32_t* objectId, uint32_t* materialId, float* coverage,
                hipStream_t s);

} // namespace vh

// ---------------------------------------------------------------- the handles of the C ABI

struct vh_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t pipe[vh::kMaxPipeline] = {}; // the wavefront's passes in flight (created on first use)
    uint32_t cus = 0;
    uint64_t uploaded = 0;
};

struct vh_renderer {
    vh_context* ctx = nullptr;
    vh::DevBuf positions, indices, uvs, vnormals, instances, materials, lights, emitters, emitCdf, envPixels, envMarginal,
        envCond, texels, texInfo, instStart, emitGuide, envGuide, lightDirs;
    std::vector<uint32_t> emitGuideHost, envGuideHost; // kept until the upload has been read
    uint32_t emitGuideBits = 0;
    uint32_t envGuideBits = 0; // marginal | row << 8
    vh::BvhBuild bvh;
    vh_scene_info info{};
    vh_view view{};
    uint32_t instanceCount = 0;
    uint32_t translucent = 0;
    bool multiscatter = true;
    int integrator = VH_MEGAKERNEL;
    bool stageTiming = false;
    bool adaptiveOn = false;
    vh_adaptive adaptive{};
    bool restirOn = false;
    vh_restir restir{};
    uint32_t seed = 0;
    uint32_t wfBlocks = 0;
    uint32_t tileW = 8, tileH = 8; // depth-0 pixel order of the wavefront
    uint32_t pipeline = 3;         // plain wavefront: sample passes in flight (1 = one at a time)
    vh::DevBuf counters; // traversal counter banks while counting is on
    // what a partial update must keep: the table lengths of positions, indices, uvs, vertex
    // normals, instances and meshes; hashes of the texture and environment tables
    uint64_t geomLens[6] = {0, 0, 0, 0, 0, 0};
    uint64_t texHash = 0, envHash = 0;

    vh::Scene scene() const
    {
        vh::Scene s;
        s.positions = positions.as<float>();
        s.indices = indices.as<uint32_t>();
        s.uvs = uvs.as<float>();
        s.vnormals = vnormals.as<float>();
        s.instances = instances.as<uint32_t>();
        s.materials = materials.as<float>();
        s.lights = lights.as<uint32_t>();
        s.lightDirs = lightDirs.as<float>();
        s.emitters = emitters.as<float>();
        s.emitCdf = emitCdf.as<float>();
        s.emitGuide = emitGuide.as<uint32_t>();
        s.emitGuideBits = emitGuideBits;
        s.envPixels = envPixels.as<float>();
        s.envMarginal = envMarginal.as<float>();
        s.envCond = envCond.as<float>();
        s.envGuide = envGuide.as<uint32_t>();
        s.envGuideBits = envGuideBits;
        s.texels = texels.as<float>();
        s.texInfo = texInfo.as<uint32_t>();
        s.bvh.nodes = bvh.nodes_buf.as<float4>();
        s.bvh.tris = bvh.tris.as<float4>();
        s.bvh.ids = bvh.ids.as<uint2>();
        s.bvh.ntri = bvh.ntri;
        s.bvh.counters = counters.as<unsigned long long>();
        s.translucent = translucent;
        return s;
    }
    uint32_t wavefrontBlocks() const { return wfBlocks ? wfBlocks : (ctx->cus * 8 > 256 ? ctx->cus * 8 : 256); }
};

struct vh_accumulation {
    vh_renderer* r = nullptr;
    uint32_t width = 0, height = 0;
    bool lighting = false;
    vh::DevBuf film, stats, light, restirBuf;
    vh::WfBuffers wf[vh::kMaxPipeline];      // wf[0]: every wavefront pass; wf[1..]: passes in flight beside it
    vh::DevBuf pipeFilm[vh::kMaxPipeline];   // FP64 film of each pass in flight, zero between passes
    hipEvent_t pipeEv[2 * vh::kMaxPipeline + 1] = {}; // pass j done, film j added, and the fork
    vh::DevBuf staged; // the megakernel's per-sample radiance when samples run in parallel
    vh::DevBuf persistRec; // the persistent megakernel's shadow records (per resident lane)
    vh::StageMarks marks;  // stage-timing events, reused across calls
    hipEvent_t span[2] = {nullptr, nullptr}; // start and end of a call
    uint32_t samples = 0;
    double gpuSeconds = 0.0;
    double stage[vh::kStages] = {}; // GPU seconds per vh::Stage
    bool stageValid = false;
    bool adaptiveOn = false;
    vh_adaptive adaptive{};
    bool restirOn = false;
    vh_restir restir{};
    uint64_t pixels() const { return (uint64_t)width * height; }
    // device bytes of the wavefront passes: every set of buffers allocated, and the films of passes in flight
    uint64_t wavefrontBytes() const
    {
        uint64_t b = 0;
        for (uint32_t j = 0; j < vh::kMaxPipeline; j++)
            b += (wf[j].ready ? wf[j].bytes : 0) + pipeFilm[j].bytes;
        return b;
    }
    ~vh_accumulation()
    {
        for (hipEvent_t e : span)
            if (e)
                (void)hipEventDestroy(e);
        for (hipEvent_t e : pipeEv)
            if (e)
                (void)hipEventDestroy(e);
    }
    // shadow records a deferred path may make: one per delta light, one emitter or environment
    // sample, and ReSTIR DI's winner
    uint32_t shadowRecords() const { return r->info.light_count + 1 + (restirOn && restir.di ? 1u : 0u); }
    uint32_t traceWaves() const { return r->ctx->cus * vh::kTraceWavesPerCu; }
    vh::Wavefront wavefront(bool split, uint32_t j = 0) const
    {
        const vh::WfBuffers& b = wf[j];
        vh::Wavefront W;
        W.state[0] = b.state0.as<uint32_t>();
        W.state[1] = b.state1.as<uint32_t>();
        W.hits = b.hits.as<uint4>();
        W.queue0 = b.queue0.as<uint32_t>();
        W.counts = b.counts.as<uint32_t>();
        W.n = (uint32_t)pixels();
        W.words = 17 + (split ? 17 : 0);
        W.tileW = r->tileW;
        W.tileH = r->tileH;
        W.traceWaves = traceWaves();
        if (!split && b.shK > 0 && shadowRecords() <= b.shK) { // (a scene update may add lights)
            W.shK = b.shK;
            W.shRec = b.shRec.as<float>();
            W.shVis = b.shVis.as<uint32_t>();
            W.shQueue = b.shQueue.as<uint32_t>();
            W.shNRec = b.shNRec.as<uint32_t>();
            W.shOut = b.shOut.as<uint32_t>();
            W.shPixel = b.shPixel.as<uint32_t>();
            W.shTotal = b.shTotal.as<float>();
        }
        return W;
    }
};

namespace vh {
// shared host helpers (vh_api.hip)
int check_hip(hipError_t e, const char* what);
int check_fits(uint64_t needed);
int ensure_wavefront(vh_accumulation* a, uint32_t j = 0);
Params base_params(const vh_renderer* r, uint32_t w, uint32_t h, bool lighting, bool adaptiveOn, const vh_adaptive& ad, bool restirOn,
                   const vh_restir& re);
} // namespace vh
