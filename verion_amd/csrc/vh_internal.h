// vh_internal.h - host-side types shared by the translation units of libverion_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>

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
    uint32_t nodes = 0;
    uint32_t leaves = 0;
    uint32_t depth = 0;
    uint32_t ntri = 0;
    float pad = 0.0f;
};

int build_bvh(BvhBuild& out, const DeviceTables& T, hipStream_t s);

// Wavefront buffers of one accumulation (slot-indexed structure of arrays, two copies for the
// ping-pong between depths).
struct Wavefront {
    uint32_t* state[2]; // words * n each
    uint4* hits;        // n
    uint32_t* queue0;   // n, adaptive sampling's depth-0 queue
    uint32_t* counts;   // 9
    uint32_t n;
    uint32_t words;
};

// Kernel launchers (vh_render.hip)
struct LaunchCfg {
    hipStream_t stream;
    uint32_t wfGrid;   // blocks of the grid-stride wavefront kernels
};

int launch_megakernel(const Scene& S, const Frame& F, const Params& P, uint32_t samples, bool restir, bool split, hipStream_t s);
int launch_wf_sample(const Scene& S, const Frame& F, const Params& P, const Wavefront& W, bool restir, bool split, const LaunchCfg& cfg,
                     hipEvent_t* stageEvents);
int launch_primary_hits(const Scene& S, const Params& P, uint4* out, hipStream_t s);
int launch_trace_rays(const Scene& S, const float* rays, uint64_t n, float tmax, uint4* out, hipStream_t s);

} // namespace vh
