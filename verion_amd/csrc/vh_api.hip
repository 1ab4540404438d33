// vh_api.hip - the C ABI of include/verion_hip.h: context, renderer, accumulation, readback.
//
// Host side of the reference backends (crates/verion-gpu-vulkan/src/renderer.rs:100-919,
// crates/verion-gpu-metal/src/renderer.rs) for HIP: the tables of pack::Packed are uploaded once,
// the LBVH is built on the device (vh_bvh.hip), and `vh_accumulate` plans the dispatches the way
// renderer.rs:538-852 does (megakernel blocks of up to 8 samples, 1 under ReSTIR; the wavefront's
// per-sample, per-depth stages), launches them on one stream and waits.
#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "vh_internal.h"

namespace vh {

namespace {
thread_local std::string g_error;
thread_local uint64_t g_needed = 0;
thread_local uint64_t g_available = 0;
} // namespace

int fail(int status, const std::string& msg)
{
    g_error = msg;
    return status;
}

int fail_oom(uint64_t needed, uint64_t available)
{
    g_needed = needed;
    g_available = available;
    g_error = "out of GPU memory: " + std::to_string(needed) + " bytes needed, " + std::to_string(available) + " available";
    return VH_ERR_OUT_OF_MEMORY;
}

int check_hip(hipError_t e, const char* what)
{
    if (e == hipSuccess)
        return VH_OK;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        return fail(VH_ERR_OUT_OF_MEMORY, std::string(what) + ": " + hipGetErrorString(e));
    return fail(VH_ERR_KERNEL, std::string(what) + ": " + hipGetErrorString(e));
}

#define TRY(x)                   \
    do {                         \
        int rc_ = (x);           \
        if (rc_ != VH_OK)        \
            return rc_;          \
    } while (0)
#define HIPTRY(x) TRY(check_hip((x), #x))

uint64_t free_bytes()
{
    size_t f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) != hipSuccess)
        return 0;
    return f;
}

// Error::OutOfMemory is raised before any allocation (vulkan context.rs:371-383).
int check_fits(uint64_t needed)
{
    uint64_t avail = free_bytes();
    if (needed > avail)
        return fail_oom(needed, avail);
    return VH_OK;
}

template <class T>
int upload(vh_context* ctx, DevBuf& b, const T* src, uint64_t n)
{
    HIPTRY(b.alloc(sizeof(T) * (n ? n : 1)));
    if (n) {
        HIPTRY(hipMemcpyAsync(b.p, src, sizeof(T) * n, hipMemcpyHostToDevice, ctx->stream));
        ctx->uploaded += sizeof(T) * n;
    } else {
        HIPTRY(hipMemsetAsync(b.p, 0, b.bytes, ctx->stream));
    }
    return VH_OK;
}

// Rewrites a table in place when its size is unchanged (a partial update), else reallocates.
template <class T>
int reupload(vh_context* ctx, DevBuf& b, const T* src, uint64_t n)
{
    if (b.p && b.bytes == sizeof(T) * n && n) {
        HIPTRY(hipMemcpyAsync(b.p, src, sizeof(T) * n, hipMemcpyHostToDevice, ctx->stream));
        ctx->uploaded += sizeof(T) * n;
        return VH_OK;
    }
    return upload(ctx, b, src, n);
}

// FNV-1a over 32-bit words (with the length), to tell whether a table changed
uint64_t hash_words(uint64_t h, const void* p, uint64_t words)
{
    const uint32_t* w = static_cast<const uint32_t*>(p);
    h = (h ^ words) * 0x100000001B3ull;
    for (uint64_t i = 0; i < words; i++)
        h = (h ^ w[i]) * 0x100000001B3ull;
    return h;
}

uint64_t texture_hash(const vh_scene_tables* t)
{
    uint64_t h = hash_words(0xCBF29CE484222325ull, t->texels, t->texels_len);
    return hash_words(h, t->tex_info, t->tex_info_len);
}

uint64_t env_hash(const vh_scene_tables* t)
{
    uint32_t dims[3] = {t->info.env_width, t->info.env_height, asu(t->info.env_pdf_scale)};
    uint64_t h = hash_words(0xCBF29CE484222325ull, dims, 3);
    h = hash_words(h, t->env_pixels, t->env_pixels_len);
    h = hash_words(h, t->env_marginal, t->env_marginal_len);
    return hash_words(h, t->env_cond, t->env_cond_len);
}

// Consistency of the tables against the layouts of pt_core.slang:12-49 (the kernels trust them).
int validate(const vh_scene_tables* t, std::vector<uint32_t>& instStart, uint64_t& ntri, uint32_t& translucent)
{
    if (!t)
        return fail(VH_ERR_INVALID_ARGUMENT, "null scene tables");
    if (t->instance_count == 0)
        return fail(VH_ERR_SCENE, "scene has no instances");
    if (t->instances_len != 16 * t->instance_count)
        return fail(VH_ERR_SCENE, "instance table is not 16 words per instance");
    if (t->positions_len % 3 || t->indices_len % 3 || t->materials_len % 24 || t->materials_len == 0)
        return fail(VH_ERR_SCENE, "positions, indices or materials have a partial record");
    uint64_t nverts = t->positions_len / 3;
    if (t->uvs_len != 2 * nverts || t->vnormals_len != 3 * nverts)
        return fail(VH_ERR_SCENE, "uvs or vertex normals do not match the vertex count");
    uint64_t nmat = t->materials_len / 24;
    uint64_t ntex = t->tex_info_len / 4;
    for (uint64_t m = 0; m < nmat; m++) {
        for (int w : {10, 11, 12, 13, 23}) {
            float f = t->materials[24 * m + w];
            if (!(f >= 0.0f) || f > (float)ntex)
                return fail(VH_ERR_SCENE, "material refers to a missing texture");
        }
    }
    for (uint64_t k = 0; k < ntex; k++) {
        uint64_t base = t->tex_info[4 * k], w = t->tex_info[4 * k + 1], h = t->tex_info[4 * k + 2], lv = t->tex_info[4 * k + 3];
        uint64_t texels = 0, lw = w, lh = h;
        for (uint64_t l = 0; l < lv; l++) {
            texels += lw * lh;
            lw = std::max<uint64_t>(lw / 2, 1);
            lh = std::max<uint64_t>(lh / 2, 1);
        }
        if (w == 0 || h == 0 || lv == 0 || 3 * (base + texels) > t->texels_len)
            return fail(VH_ERR_SCENE, "texture table out of range");
    }
    uint64_t need_lights = (uint64_t)t->info.light_count * 16 + kAlbedoWords + 2 * t->instance_count;
    if (t->lights_len < need_lights)
        return fail(VH_ERR_SCENE, "light table shorter than lights + albedo tables + link masks");
    // IES profiles (word 15 = offset + 1, pack.rs:679-686): the frame and a 361 x 181 table past the masks
    for (uint64_t k = 0; k < t->info.light_count; k++) {
        uint64_t w15 = t->lights[16 * k + 15];
        if (w15 != 0 && (w15 - 1 < need_lights || w15 - 1 + 9 + 361ull * 181 > t->lights_len))
            return fail(VH_ERR_SCENE, "IES profile out of the light table");
    }
    if (t->emitters_len != 16ull * t->info.emitter_count || t->emit_cdf_len != t->info.emitter_count)
        return fail(VH_ERR_SCENE, "emitter table does not match the emitter count");
    uint64_t ew = t->info.env_width, eh = t->info.env_height;
    if ((ew == 0) != (eh == 0) || t->env_pixels_len != 3 * ew * eh || t->env_marginal_len != eh || t->env_cond_len != ew * eh)
        return fail(VH_ERR_SCENE, "environment tables do not match the map size");
    instStart.assign(t->instance_count, 0);
    ntri = 0;
    translucent = 0;
    for (uint64_t i = 0; i < t->instance_count; i++) {
        uint32_t mesh = t->instance_mesh[i];
        if (mesh >= t->mesh_count)
            return fail(VH_ERR_SCENE, "instance refers to a missing mesh");
        const uint32_t* mr = t->meshes + 4 * mesh;
        const uint32_t* rec = t->instances + 16 * i;
        if (rec[0] != mr[0] || rec[1] != mr[2])
            return fail(VH_ERR_SCENE, "instance offsets differ from its mesh range");
        if ((uint64_t)mr[0] + mr[1] > nverts || (uint64_t)mr[2] + mr[3] > t->indices_len || mr[3] % 3)
            return fail(VH_ERR_SCENE, "mesh range out of the vertex or index table");
        if (rec[2] >= nmat)
            return fail(VH_ERR_SCENE, "instance refers to a missing material");
        if (rec[3] != 0 && (uint64_t)rec[3] - 1 + mr[3] / 3 > t->info.emitter_count)
            return fail(VH_ERR_SCENE, "emitter base out of the emitter table");
        for (uint64_t k = 0; k < mr[3]; k++)
            if (t->indices[mr[2] + k] >= mr[1])
                return fail(VH_ERR_SCENE, "triangle index out of its mesh");
        const float* M = t->materials + 24 * rec[2];
        if (M[22] < 1.0f || M[23] > 0.0f)
            translucent = 1;
        instStart[i] = (uint32_t)ntri;
        ntri += mr[3] / 3;
    }
    if (ntri >= 0x7fffffffull)
        return fail(VH_ERR_SCENE, "more than 2^31 triangles");
    return VH_OK;
}

// pack::base_params (pack.rs:828-874)
Params base_params(const vh_renderer* r, uint32_t w, uint32_t h, bool lighting, bool adaptiveOn, const vh_adaptive& ad, bool restirOn,
                   const vh_restir& re)
{
    Params p;
    std::memset(&p, 0, sizeof(p));
    auto bits = [](float f) { return asu(f); };
    p.w[0] = w;
    p.w[1] = h;
    p.w[2] = bits(r->view.camera[2]);
    p.w[3] = bits(r->view.aperture);
    p.w[4] = bits(r->view.focus);
    p.w[5] = r->info.light_count;
    p.w[7] = kMaxDepth;
    for (int k = 0; k < 12; k++)
        p.w[8 + k] = bits(r->view.camera[4 + k]);
    for (int k = 0; k < 3; k++)
        p.w[20 + k] = bits(r->info.sky[k]);
    p.w[23] = r->info.emitter_count;
    p.w[24] = r->info.env_width;
    p.w[25] = r->info.env_height;
    p.w[26] = bits(r->info.env_pdf_scale);
    p.w[27] = bits(2.0f * r->view.camera[2] / (float)h);
    p.w[41] = r->info.virtual_light ? 1u : 0u;
    p.w[42] = r->info.transparent_background ? 1u : 0u;
    p.w[39] = bits(r->view.shift[0]);
    p.w[40] = bits(r->view.shift[1]);
    p.w[28] = lighting ? 1u : 0u;
    p.w[29] = r->multiscatter ? 1u : 0u;
    if (adaptiveOn) {
        p.w[32] = 1;
        p.w[33] = bits(ad.threshold);
        p.w[34] = ad.min_spp;
    }
    if (restirOn) {
        p.w[36] = (re.di ? 1u : 0u) | (re.gi ? 2u : 0u);
        p.w[37] = re.candidates;
        p.w[38] = re.history;
        if (p.w[36] != 0)
            p.w[36] |= re.neighbours << 8; // 0 = the reference's 2
    }
    p.seedMix = r->seed == 0 ? 0u : pcg(r->seed ^ 0x5851F42Du);
    return p;
}

// Cutpoints of a power CDF for emitterIndex: guide[k] = the first index whose value reaches
// k / 2^bits (n - 1 when none does), k = 0 .. 2^bits. A CDF that is not non-decreasing (never
// made by the packer) gets bits = 0, i.e. the full binary search.
uint32_t cdf_guide_bits(uint32_t n)
{
    uint32_t bits = 0;
    while (bits < 16 && (1u << bits) < n)
        bits++;
    return bits;
}

bool cdf_guide(const float* cdf, uint32_t n, uint32_t bits, uint32_t* guide)
{
    for (uint32_t i = 0; i + 1 < n; i++)
        if (!(cdf[i] <= cdf[i + 1]))
            return false;
    uint32_t G = 1u << bits, i = 0;
    for (uint32_t k = 0; k <= G; k++) {
        float t = (float)k / (float)G; // exact: k <= 2^16, G a power of two
        while (i + 1 < n && !(cdf[i] >= t))
            i++;
        guide[k] = i;
    }
    return true;
}

int upload_emitter_guide(vh_renderer* r, const vh_scene_tables* t)
{
    uint32_t n = (uint32_t)t->info.emitter_count;
    uint32_t bits = n > 1 ? cdf_guide_bits(n) : 0;
    r->emitGuideHost.assign((1u << bits) + 1, 0u);
    if (n > 1 && !cdf_guide(t->emit_cdf, n, bits, r->emitGuideHost.data())) {
        bits = 0;
        r->emitGuideHost.assign({0u, n - 1});
    }
    r->emitGuideBits = bits;
    return reupload(r->ctx, r->emitGuide, r->emitGuideHost.data(), r->emitGuideHost.size());
}

// Cutpoints for envSample: the marginal CDF's guide (2^bm + 1 entries), then one guide per row of
// the conditional CDFs (2^bc + 1 entries each). CDFs that are not non-decreasing (never made by the
// packer) get guides spanning the whole table, i.e. the full binary searches.
int upload_env_guide(vh_renderer* r, const vh_scene_tables* t)
{
    const uint32_t w = t->info.env_width, h = t->info.env_height;
    std::vector<uint32_t>& g = r->envGuideHost;
    if (w == 0 || h == 0) {
        r->envGuideBits = 0;
        g.assign(2, 0u);
        return reupload(r->ctx, r->envGuide, g.data(), g.size());
    }
    uint32_t bm = cdf_guide_bits(h), bc = cdf_guide_bits(w);
    auto fill = [&]() {
        const uint32_t gm = (1u << bm) + 1, gc = (1u << bc) + 1;
        g.assign(gm + (size_t)h * gc, 0u);
        bool ok = cdf_guide(t->env_marginal, h, bm, g.data());
        for (uint32_t j = 0; j < h && ok; j++)
            ok = cdf_guide(t->env_cond + (size_t)j * w, w, bc, g.data() + gm + (size_t)j * gc);
        return ok;
    };
    if (!fill()) {
        bm = bc = 0;
        g.assign(2 + 2 * (size_t)h, 0u);
        g[1] = h - 1;
        for (uint32_t j = 0; j < h; j++)
            g[2 + 2 * (size_t)j + 1] = w - 1;
    }
    r->envGuideBits = bm | (bc << 8);
    return reupload(r->ctx, r->envGuide, g.data(), g.size());
}

int upload_tables(vh_renderer* r, const vh_scene_tables* t)
{
    std::vector<uint32_t> instStart;
    uint64_t ntri = 0;
    uint32_t translucent = 0;
    TRY(validate(t, instStart, ntri, translucent));
    // tables + world triangles (twice during the build) + sort buffers + binary and wide nodes
    uint64_t tables = 4 * (t->positions_len + t->indices_len + t->uvs_len + t->vnormals_len + t->instances_len + t->materials_len +
                           t->lights_len + t->emitters_len + t->emit_cdf_len + t->env_pixels_len + t->env_marginal_len +
                           t->env_cond_len + t->texels_len + t->tex_info_len + t->instance_count);
    uint64_t build = ntri * (2 * 48 + 16 + 64 + 4 * 8 + 128 + 48);
    TRY(check_fits(tables + build));
    vh_context* c = r->ctx;
    TRY(upload(c, r->positions, t->positions, t->positions_len));
    TRY(upload(c, r->indices, t->indices, t->indices_len));
    TRY(upload(c, r->uvs, t->uvs, t->uvs_len));
    TRY(upload(c, r->vnormals, t->vnormals, t->vnormals_len));
    TRY(upload(c, r->instances, t->instances, t->instances_len));
    TRY(upload(c, r->materials, t->materials, t->materials_len));
    TRY(upload(c, r->lights, t->lights, t->lights_len));
    HIPTRY(r->lightDirs.alloc(16 * (size_t)std::max<uint64_t>(t->info.light_count, 1)));
    TRY(launch_light_dirs(r->lights.as<uint32_t>(), (uint32_t)t->info.light_count, r->lightDirs.as<float>(), c->stream));
    TRY(upload(c, r->emitters, t->emitters, t->emitters_len));
    TRY(upload(c, r->emitCdf, t->emit_cdf, t->emit_cdf_len));
    TRY(upload_emitter_guide(r, t));
    TRY(upload(c, r->envPixels, t->env_pixels, t->env_pixels_len));
    TRY(upload(c, r->envMarginal, t->env_marginal, t->env_marginal_len));
    TRY(upload(c, r->envCond, t->env_cond, t->env_cond_len));
    TRY(upload_env_guide(r, t));
    TRY(upload(c, r->texels, t->texels, t->texels_len));
    TRY(upload(c, r->texInfo, t->tex_info, t->tex_info_len));
    TRY(upload(c, r->instStart, instStart.data(), instStart.size()));
    HIPTRY(hipStreamSynchronize(c->stream));
    r->info = t->info;
    r->view = t->view;
    r->instanceCount = (uint32_t)t->instance_count;
    r->translucent = translucent;
    uint64_t lens[6] = {t->positions_len, t->indices_len, t->uvs_len, t->vnormals_len, t->instances_len, t->mesh_count};
    std::memcpy(r->geomLens, lens, sizeof(lens));
    r->texHash = texture_hash(t);
    r->envHash = env_hash(t);
    DeviceTables dt{r->positions.as<float>(), r->indices.as<uint32_t>(), r->instances.as<uint32_t>(), r->instStart.as<uint32_t>(),
                    r->instanceCount, (uint32_t)ntri};
    TRY(build_bvh(r->bvh, dt, c->stream));
    if (r->bvh.depth > kMaxTreeDepth)
        return fail(VH_ERR_SCENE, "LBVH deeper than the traversal stack holds (depth " + std::to_string(r->bvh.depth) + ")");
    return VH_OK;
}

int ensure_wavefront(vh_accumulation* a, uint32_t j)
{
    WfBuffers& b = a->wf[j];
    if (b.ready)
        return VH_OK;
    uint64_t n = a->pixels();
    uint32_t words = 17 + (a->lighting ? 17 : 0);
    // deferred shadow rays (the plain integrator, ReSTIR DI, ReSTIR GI past its first bounce): one
    // record per delta light, one for the emitter or environment sample and one for ReSTIR DI's
    // winner; scenes with more lights than that leaves room for, and the lighting split, shade fused
    uint32_t K = a->shadowRecords();
    bool defer = !a->lighting && K <= kMaxShadow;
    uint64_t shBytes = defer ? (4ull * (kShadowWords + 2) * K * n + 4ull * 6 * n) : 0;
    uint64_t bytes = 2 * 4 * words * n + 16 * n + 4 * n + 4 * kCountWords + shBytes;
    TRY(check_fits(bytes));
    HIPTRY(b.state0.alloc(4 * words * n));
    HIPTRY(b.state1.alloc(4 * words * n));
    HIPTRY(b.hits.alloc(16 * n));
    HIPTRY(b.queue0.alloc(4 * n));
    HIPTRY(b.counts.alloc(4 * kCountWords));
    HIPTRY(hipMemsetAsync(b.counts.p, 0, 4 * kCountWords, a->r->ctx->stream));
    if (defer) {
        HIPTRY(b.shRec.alloc(4ull * kShadowWords * K * n));
        HIPTRY(b.shVis.alloc(4ull * K * n));
        HIPTRY(b.shQueue.alloc(4ull * K * n));
        HIPTRY(b.shNRec.alloc(4 * n));
        HIPTRY(b.shOut.alloc(4 * n));
        HIPTRY(b.shPixel.alloc(4 * n));
        HIPTRY(b.shTotal.alloc(12 * n));
        b.shK = K;
    }
    b.bytes = bytes;
    b.ready = true;
    return VH_OK;
}

// Passes in flight j = 1 .. depth - 1 beside wf[0]: their wavefront buffers, the FP64 films of all
// `depth` passes (zeroed once; the film add zeroes them again), the streams and the events.
int ensure_pipeline(vh_accumulation* a, uint32_t depth)
{
    vh_context* ctx = a->r->ctx;
    const uint64_t filmBytes = 24 * a->pixels();
    for (uint32_t j = 0; j < depth; j++) {
        if (j > 0)
            TRY(ensure_wavefront(a, j));
        if (!a->pipeFilm[j].p) {
            TRY(check_fits(filmBytes));
            HIPTRY(a->pipeFilm[j].alloc(filmBytes));
            HIPTRY(hipMemsetAsync(a->pipeFilm[j].p, 0, filmBytes, ctx->stream));
        }
        if (!ctx->pipe[j])
            HIPTRY(hipStreamCreateWithFlags(&ctx->pipe[j], hipStreamNonBlocking));
    }
    for (hipEvent_t& e : a->pipeEv)
        if (!e)
            HIPTRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return VH_OK;
}

} // namespace vh

using namespace vh;

extern "C" {

const char* vh_last_error(void) { return g_error.c_str(); }

int vh_trace_limits(uint32_t out[4])
{
    if (!out)
        return fail(VH_ERR_INVALID_ARGUMENT, "null argument");
    out[0] = (uint32_t)kStack;
    out[1] = (uint32_t)kShort;
    out[2] = kMaxTreeDepth;
    out[3] = kLeafMax;
    return VH_OK;
}

int vh_cdf_guide(const float* cdf, uint32_t n, uint32_t bits, uint32_t* guide)
{
    if (!cdf || !guide || n == 0 || bits > 16)
        return fail(VH_ERR_INVALID_ARGUMENT, "vh_cdf_guide: null table, empty CDF or more than 16 bits");
    if (!cdf_guide(cdf, n, bits, guide))
        return fail(VH_ERR_INVALID_ARGUMENT, "vh_cdf_guide: the CDF is not non-decreasing");
    return VH_OK;
}
uint64_t vh_last_error_needed(void) { return g_needed; }
uint64_t vh_last_error_available(void) { return g_available; }

int vh_context_create(vh_context** out)
{
    if (!out)
        return fail(VH_ERR_INVALID_ARGUMENT, "null output");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return fail(VH_ERR_NO_DEVICE, "no HIP device");
    int dev = 0;
    HIPTRY(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPTRY(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(VH_ERR_UNSUPPORTED, std::string("device is ") + prop.gcnArchName + ", this library is built for gfx950");
    auto* c = new vh_context();
    c->device = dev;
    c->cus = (uint32_t)prop.multiProcessorCount;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return check_hip(e, "hipStreamCreateWithFlags");
    }
    *out = c;
    return VH_OK;
}

void vh_context_destroy(vh_context* ctx)
{
    if (!ctx)
        return;
    if (ctx->stream)
        (void)hipStreamDestroy(ctx->stream);
    for (hipStream_t p : ctx->pipe)
        if (p)
            (void)hipStreamDestroy(p);
    delete ctx;
}

int vh_memory(const vh_context* ctx, vh_memory_usage* out)
{
    if (!ctx || !out)
        return fail(VH_ERR_INVALID_ARGUMENT, "null argument");
    size_t f = 0, t = 0;
    HIPTRY(hipMemGetInfo(&f, &t));
    out->budget = t;
    out->allocated = t - f;
    out->uploaded = ctx->uploaded;
    return VH_OK;
}

int vh_renderer_create(vh_context* ctx, const vh_scene_tables* tables, vh_renderer** out)
{
    if (!ctx || !out)
        return fail(VH_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    std::unique_ptr<vh_renderer> r(new vh_renderer());
    r->ctx = ctx;
    TRY(upload_tables(r.get(), tables));
    *out = r.release();
    return VH_OK;
}

void vh_renderer_destroy(vh_renderer* r)
{
    if (!r)
        return;
    (void)hipStreamSynchronize(r->ctx->stream);
    delete r;
}

// Renderer::update (vulkan renderer.rs:404-502) for a partial plan: each flag re-uploads its tables
// in place; moved instances and deformed meshes refit the LBVH (or rebuild it, see kRefitLimit).
static int update_tables(vh_renderer* r, const vh_scene_tables* t, const vh_update_plan& plan, vh_update_report& rep)
{
    std::vector<uint32_t> instStart;
    uint64_t ntri = 0;
    uint32_t translucent = 0;
    TRY(validate(t, instStart, ntri, translucent));
    uint64_t lens[6] = {t->positions_len, t->indices_len, t->uvs_len, t->vnormals_len, t->instances_len, t->mesh_count};
    if (std::memcmp(lens, r->geomLens, sizeof(lens)) != 0 || t->instance_count != r->instanceCount || ntri != r->bvh.ntri)
        return fail(VH_ERR_INVALID_ARGUMENT, "a partial update must keep the geometry table sizes (use a full plan)");
    if (plan.blas_count && !plan.blases)
        return fail(VH_ERR_INVALID_ARGUMENT, "null mesh list");
    for (uint64_t k = 0; k < plan.blas_count; k++)
        if (plan.blases[k] >= t->mesh_count)
            return fail(VH_ERR_INVALID_ARGUMENT, "plan names a missing mesh");
    vh_context* c = r->ctx;
    HIPTRY(hipDeviceSynchronize()); // NRC side streams and batches read these tables too
    if (plan.camera)
        r->view = t->view;
    if (plan.vertices) {
        TRY(reupload(c, r->positions, t->positions, t->positions_len));
        TRY(reupload(c, r->uvs, t->uvs, t->uvs_len));
        TRY(reupload(c, r->vnormals, t->vnormals, t->vnormals_len));
    }
    if (plan.materials) {
        TRY(reupload(c, r->materials, t->materials, t->materials_len));
        uint64_t h = texture_hash(t);
        if (h != r->texHash) {
            TRY(reupload(c, r->texels, t->texels, t->texels_len));
            TRY(reupload(c, r->texInfo, t->tex_info, t->tex_info_len));
            r->texHash = h;
        }
    }
    if (plan.lights) {
        TRY(reupload(c, r->lights, t->lights, t->lights_len));
        if (r->lightDirs.bytes < 16 * (size_t)std::max<uint64_t>(t->info.light_count, 1))
            HIPTRY(r->lightDirs.alloc(16 * (size_t)std::max<uint64_t>(t->info.light_count, 1)));
        TRY(launch_light_dirs(r->lights.as<uint32_t>(), (uint32_t)t->info.light_count, r->lightDirs.as<float>(), c->stream));
        r->info.light_count = t->info.light_count;
        std::memcpy(r->info.sky, t->info.sky, sizeof(r->info.sky));
        r->info.virtual_light = t->info.virtual_light;
        r->info.transparent_background = t->info.transparent_background;
        uint64_t h = env_hash(t);
        if (h != r->envHash) {
            TRY(reupload(c, r->envPixels, t->env_pixels, t->env_pixels_len));
            TRY(reupload(c, r->envMarginal, t->env_marginal, t->env_marginal_len));
            TRY(reupload(c, r->envCond, t->env_cond, t->env_cond_len));
            TRY(upload_env_guide(r, t));
            r->info.env_width = t->info.env_width;
            r->info.env_height = t->info.env_height;
            r->info.env_pdf_scale = t->info.env_pdf_scale;
            r->envHash = h;
        }
    }
    if (plan.emitters) {
        TRY(reupload(c, r->emitters, t->emitters, t->emitters_len));
        TRY(reupload(c, r->emitCdf, t->emit_cdf, t->emit_cdf_len));
        TRY(upload_emitter_guide(r, t));
        r->info.emitter_count = t->info.emitter_count;
    }
    if (plan.instances)
        TRY(reupload(c, r->instances, t->instances, t->instances_len));
    r->translucent = translucent;
    HIPTRY(hipStreamSynchronize(c->stream));
    if (plan.vertices || plan.tlas) {
        DeviceTables dt{r->positions.as<float>(), r->indices.as<uint32_t>(), r->instances.as<uint32_t>(), r->instStart.as<uint32_t>(),
                        r->instanceCount, (uint32_t)ntri};
        double ratio = 1.0;
        TRY(refit_bvh(r->bvh, dt, c->stream, ratio));
        rep.bvh_work = 1;
        if (ratio > kRefitLimit) {
            TRY(build_bvh(r->bvh, dt, c->stream));
            rep.bvh_work = 2;
            if (r->bvh.depth > kMaxTreeDepth)
                return fail(VH_ERR_SCENE, "LBVH deeper than the traversal stack holds (depth " + std::to_string(r->bvh.depth) + ")");
        }
        HIPTRY(hipStreamSynchronize(c->stream));
    }
    rep.blases_built = plan.vertices ? plan.blas_count : 0;
    rep.tlas_built = plan.instances && plan.tlas ? 1u : 0u;
    return VH_OK;
}

int vh_renderer_update(vh_renderer* r, const vh_scene_tables* tables, const vh_update_plan* plan, vh_update_report* report)
{
    if (!r || !tables)
        return fail(VH_ERR_INVALID_ARGUMENT, "null argument");
    uint64_t before = r->ctx->uploaded;
    vh_update_report rep{};
    if (!plan || plan->full) {
        HIPTRY(hipDeviceSynchronize());
        TRY(upload_tables(r, tables));
        rep.blases_built = tables->mesh_count;
        rep.tlas_built = 1;
        rep.bvh_work = 2;
    } else {
        TRY(update_tables(r, tables, *plan, rep));
    }
    rep.bytes_uploaded = r->ctx->uploaded - before;
    if (report)
        *report = rep;
    return VH_OK;
}

int vh_set_camera(vh_renderer* r, const vh_view* view)
{
    if (!r || !view)
        return fail(VH_ERR_INVALID_ARGUMENT, "null argument");
    r->view = *view;
    return VH_OK;
}

int vh_set_multiscatter(vh_renderer* r, int on)
{
    if (!r)
        return fail(VH_ERR_INVALID_ARGUMENT, "null renderer");
    r->multiscatter = on != 0;
    return VH_OK;
}

int vh_set_integrator(vh_renderer* r, int integrator)
{
    if (!r || (integrator != VH_MEGAKERNEL && integrator != VH_WAVEFRONT))
        return fail(VH_ERR_INVALID_ARGUMENT, "bad renderer or integrator");
    r->integrator = integrator;
    return VH_OK;
}

int vh_integrator(const vh_renderer* r) { return r ? r->integrator : -1; }

int vh_set_stage_timing(vh_renderer* r, int on)
{
    if (!r)
        return fail(VH_ERR_INVALID_ARGUMENT, "null renderer");
    r->stageTiming = on != 0;
    return VH_OK;
}

int vh_set_adaptive(vh_renderer* r, const vh_adaptive* adaptive)
{
    if (!r)
        return fail(VH_ERR_INVALID_ARGUMENT, "null renderer");
    r->adaptiveOn = adaptive != nullptr;
    if (adaptive)
        r->adaptive = *adaptive;
    return VH_OK;
}

int vh_set_restir(vh_renderer* r, const vh_restir* restir)
{
    if (!r)
        return fail(VH_ERR_INVALID_ARGUMENT, "null renderer");
    if (restir && restir->neighbours > kRestirMaxNeighbours)
        return fail(VH_ERR_INVALID_ARGUMENT, "ReSTIR reuses at most " + std::to_string(kRestirMaxNeighbours) + " neighbours");
    r->restirOn = restir != nullptr;
    if (restir)
        r->restir = *restir;
    return VH_OK;
}

int vh_set_seed(vh_renderer* r, uint32_t seed)
{
    if (!r)
        return fail(VH_ERR_INVALID_ARGUMENT, "null renderer");
    r->seed = seed;
    return VH_OK;
}

int vh_set_wavefront_blocks(vh_renderer* r, uint32_t blocks)
{
    if (!r)
        return fail(VH_ERR_INVALID_ARGUMENT, "null renderer");
    r->wfBlocks = blocks;
    return VH_OK;
}

int vh_set_wavefront_pipeline(vh_renderer* r, uint32_t passes)
{
    if (!r || passes < 1 || passes > kMaxPipeline)
        return fail(VH_ERR_INVALID_ARGUMENT, "passes in flight must be 1 to " + std::to_string(kMaxPipeline));
    r->pipeline = passes;
    return VH_OK;
}

int vh_set_wavefront_tiles(vh_renderer* r, uint32_t tile_width, uint32_t tile_height)
{
    if (!r)
        return fail(VH_ERR_INVALID_ARGUMENT, "null renderer");
    if ((tile_width == 0) != (tile_height == 0) || tile_width > 4096 || tile_height > 4096)
        return fail(VH_ERR_INVALID_ARGUMENT, "tiles are w x h with both in [1, 4096], or 0 x 0 for scanlines");
    r->tileW = tile_width;
    r->tileH = tile_height;
    return VH_OK;
}

int vh_accumulation_create(vh_renderer* r, uint32_t width, uint32_t height, int lighting, vh_accumulation** out)
{
    if (!r || !out)
        return fail(VH_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (width == 0 || height == 0)
        return fail(VH_ERR_INVALID_ARGUMENT, "zero image size");
    uint64_t n = (uint64_t)width * height;
    if (n >= (1ull << 30)) // pixel ids carry two flag bits in the deferred shading
        return fail(VH_ERR_INVALID_ARGUMENT, "image of 2^30 pixels or more");
    uint64_t bytes = 24 * n + (lighting ? 4 * kLightingStride * n : 0);
    TRY(check_fits(bytes));
    std::unique_ptr<vh_accumulation> a(new vh_accumulation());
    a->r = r;
    a->width = width;
    a->height = height;
    a->lighting = lighting != 0;
    hipStream_t s = r->ctx->stream;
    HIPTRY(a->film.alloc(24 * n));
    HIPTRY(hipMemsetAsync(a->film.p, 0, 24 * n, s));
    if (a->lighting) {
        HIPTRY(a->light.alloc(4 * kLightingStride * n));
        HIPTRY(hipMemsetAsync(a->light.p, 0, 4 * kLightingStride * n, s));
    }
    HIPTRY(hipStreamSynchronize(s));
    *out = a.release();
    return VH_OK;
}

void vh_accumulation_destroy(vh_accumulation* acc)
{
    if (!acc)
        return;
    (void)hipStreamSynchronize(acc->r->ctx->stream);
    delete acc;
}

int vh_accumulate(vh_renderer* r, vh_accumulation* a, uint32_t samples)
{
    if (!r || !a || a->r != r)
        return fail(VH_ERR_INVALID_ARGUMENT, "accumulation does not belong to this renderer");
    hipStream_t s = r->ctx->stream;
    uint64_t n = a->pixels();
    if (a->samples == 0) {
        a->adaptiveOn = r->adaptiveOn;
        a->adaptive = r->adaptive;
        a->restirOn = r->restirOn;
        a->restir = r->restir;
        if (a->restirOn && !a->restirBuf.p) {
            uint64_t bytes = 4ull * kRestirStride * n;
            TRY(check_fits(bytes));
            HIPTRY(a->restirBuf.alloc(bytes));
            HIPTRY(hipMemsetAsync(a->restirBuf.p, 0, bytes, s));
        }
    }
    if (a->restirOn != r->restirOn ||
        (a->restirOn && std::memcmp(&a->restir, &r->restir, sizeof(vh_restir)) != 0))
        return fail(VH_ERR_SCENE, "ReSTIR was switched mid-accumulation; start a new accumulation");
    if (a->adaptiveOn && !a->stats.p) {
        TRY(check_fits(24 * n));
        HIPTRY(a->stats.alloc(24 * n));
        HIPTRY(hipMemsetAsync(a->stats.p, 0, 24 * n, s));
    }
    if (a->adaptiveOn != r->adaptiveOn)
        return fail(VH_ERR_SCENE, "adaptive sampling was switched mid-accumulation; start a new accumulation");
    if (samples == 0)
        return VH_OK;
    bool wave = r->integrator == VH_WAVEFRONT;
    if (wave)
        TRY(ensure_wavefront(a));

    Scene S = r->scene();
    Frame F{a->film.as<double>(), a->stats.as<double>(), a->light.as<float>(), a->restirBuf.as<float>()};
    Params base = base_params(r, a->width, a->height, a->lighting, a->adaptiveOn, r->adaptive, a->restirOn, a->restir);
    bool restir = a->restirOn;
    bool split = a->lighting;

    for (hipEvent_t& e : a->span)
        if (!e)
            HIPTRY(hipEventCreate(&e));
    hipEvent_t e0 = a->span[0], e1 = a->span[1];
    StageMarks& marks = a->marks;
    marks.reset();
    bool timing = wave && r->stageTiming;
    if (timing) // the marks of one pass (about 3 per depth), created before the first launch
        HIPTRY(marks.reserve((size_t)samples * (4 * kMaxDepth + 4)));
    int rc = VH_OK;
    HIPTRY(hipEventRecord(e0, s));
    // the plain wavefront keeps several sample passes in flight (without stage timing, which
    // brackets launches on one stream)
    const uint32_t depth = std::min(r->pipeline, samples);
    const bool piped = wave && depth > 1 && !restir && !split && !a->adaptiveOn && !timing;
    if (piped)
        TRY(ensure_pipeline(a, depth));
    if (piped) {
        // pass k runs on stream pipe[k % depth] into film j = k % depth; the library stream adds
        // film j to the accumulation's film once pass k is done, in sample order, and pass k + depth
        // waits for that add before it reuses buffers and film j
        hipEvent_t fork = a->pipeEv[2 * kMaxPipeline];
        rc = check_hip(hipEventRecord(fork, s), "hipEventRecord");
        for (uint32_t j = 0; j < depth && rc == VH_OK; j++)
            rc = check_hip(hipStreamWaitEvent(r->ctx->pipe[j], fork, 0), "hipStreamWaitEvent");
        for (uint32_t k = 0; k < samples && rc == VH_OK; k++) {
            const uint32_t j = k % depth;
            hipStream_t ps = r->ctx->pipe[j];
            hipEvent_t done = a->pipeEv[j], added = a->pipeEv[kMaxPipeline + j];
            if (k >= depth)
                rc = check_hip(hipStreamWaitEvent(ps, added, 0), "hipStreamWaitEvent");
            Frame Fj = F;
            Fj.film = a->pipeFilm[j].as<double>();
            Params P = base;
            P.w[6] = a->samples + k;
            P.w[31] = (uint32_t)n;
            LaunchCfg cfg{ps, r->wfBlocks ? r->wfBlocks : std::max(256u, r->ctx->cus * 8)};
            if (rc == VH_OK)
                rc = launch_wf_sample(S, Fj, P, a->wavefront(false, j), false, false, cfg, nullptr);
            if (rc == VH_OK)
                rc = check_hip(hipEventRecord(done, ps), "hipEventRecord");
            if (rc == VH_OK)
                rc = check_hip(hipStreamWaitEvent(s, done, 0), "hipStreamWaitEvent");
            if (rc == VH_OK)
                rc = launch_film_add_pass(F.film, Fj.film, 3 * n, s);
            if (rc == VH_OK)
                rc = check_hip(hipEventRecord(added, s), "hipEventRecord");
        }
        if (rc != VH_OK) { // nothing of this call may still run once it returns, and no pass film
                           // may carry a pass that was never added into the next call
            for (uint32_t j = 0; j < depth; j++)
                (void)hipStreamSynchronize(r->ctx->pipe[j]);
            for (uint32_t j = 0; j < depth; j++)
                (void)hipMemsetAsync(a->pipeFilm[j].p, 0, 24 * n, s);
            (void)hipStreamSynchronize(s);
        }
    } else if (wave) {
        Wavefront W = a->wavefront(split);
        LaunchCfg cfg{s, r->wfBlocks ? r->wfBlocks : std::max(256u, r->ctx->cus * 8)};
        for (uint32_t k = 0; k < samples && rc == VH_OK; k++) {
            Params P = base;
            P.w[6] = a->samples + k;
            P.w[31] = (uint32_t)n;
            rc = launch_wf_sample(S, F, P, W, restir, split, cfg, timing ? &marks : nullptr);
        }
    } else if (!restir && !split && !a->adaptiveOn) {
        // the plain megakernel: the samples of a dispatch (up to kParallelPaths paths) in parallel
        // on persistent waves, added to the film in sample order (1080p procedural scene, 1 spp
        // per call: 359 Msamples/s against 338 for one thread per pixel)
        uint32_t chunk = (uint32_t)std::min<uint64_t>(std::min<uint64_t>(samples, 1024), std::max<uint64_t>(1, kParallelPaths / n));
        const uint64_t stagedBytes = (12ull * n * chunk + 255) & ~255ull;
        uint64_t bytes = stagedBytes + 256; // + the persistent megakernel's work counter
        if (a->staged.bytes < bytes) {
            TRY(check_fits(bytes));
            HIPTRY(a->staged.alloc(bytes));
        }
        uint32_t* next = reinterpret_cast<uint32_t*>(a->staged.as<char>() + stagedBytes);
        // deferred shadow rays, one record per delta light and one emitter or environment sample,
        // for every resident lane (scenes with more lights trace them in the shading)
        const uint32_t K = r->info.light_count + 1;
        float* shRec = nullptr;
        if (K <= kMaxShadow) {
            uint64_t recBytes = 4ull * kShadowWords * K * (uint64_t)r->ctx->cus * kPersistWavesPerCu * 64;
            if (a->persistRec.bytes < recBytes) {
                TRY(check_fits(recBytes));
                HIPTRY(a->persistRec.alloc(recBytes));
            }
            shRec = a->persistRec.as<float>();
        }
        uint32_t k = 0;
        while (k < samples && rc == VH_OK) {
            Params P = base;
            P.w[6] = a->samples + k;
            uint32_t m = std::min(samples - k, chunk);
            P.w[35] = m;
            rc = launch_megakernel_samples(S, F, P, m, a->staged.as<float>(), next, shRec, K, r->ctx->cus, s);
            k += m;
        }
    } else {
        uint32_t k = 0;
        while (k < samples && rc == VH_OK) {
            Params P = base;
            P.w[6] = a->samples + k;
            P.w[35] = std::min(samples - k, restir ? 1u : kSamplesPerDispatch);
            rc = launch_megakernel(S, F, P, P.w[35], restir, split, s);
            k += P.w[35];
        }
    }
    HIPTRY(hipEventRecord(e1, s));
    hipError_t se = hipEventSynchronize(e1);
    if (rc == VH_OK && se != hipSuccess)
        rc = check_hip(se, "accumulate");
    if (rc == VH_OK) {
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        a->gpuSeconds += (double)ms * 1.0e-3;
        a->samples += samples;
        if (timing) {
            for (size_t k = 1; k < marks.used; k++) {
                if (marks.label[k] >= kStages)
                    continue;
                float ms = 0.0f;
                hipError_t te = hipEventElapsedTime(&ms, marks.ev[k - 1], marks.ev[k]);
                if (te != hipSuccess) {
                    rc = check_hip(te, "stage timing");
                    break;
                }
                a->stage[marks.label[k]] += (double)ms * 1.0e-3;
            }
            a->stageValid = rc == VH_OK;
        }
    }
    return rc;
}

uint32_t vh_acc_samples(const vh_accumulation* a) { return a ? a->samples : 0; }
double vh_acc_gpu_seconds(const vh_accumulation* a) { return a ? a->gpuSeconds : 0.0; }

int vh_acc_stage_seconds(const vh_accumulation* a, double out[2])
{
    if (!a || !out)
        return fail(VH_ERR_INVALID_ARGUMENT, "null argument");
    if (!a->stageValid)
        return fail(VH_ERR_UNSUPPORTED, "stage timing was off");
    out[0] = a->stage[kStageExtend] + a->stage[kStageTrace];
    out[1] = a->stage[kStageShade] + a->stage[kStageShadow] + a->stage[kStageResolve];
    return VH_OK;
}

int vh_acc_kernel_seconds(const vh_accumulation* a, double out[5])
{
    if (!a || !out)
        return fail(VH_ERR_INVALID_ARGUMENT, "null argument");
    if (!a->stageValid)
        return fail(VH_ERR_UNSUPPORTED, "stage timing was off");
    for (int k = 0; k < kStages; k++)
        out[k] = a->stage[k];
    return VH_OK;
}

uint64_t vh_acc_wavefront_bytes(const vh_accumulation* a) { return a ? a->wavefrontBytes() : 0; }
uint64_t vh_acc_restir_bytes(const vh_accumulation* a) { return a && a->restirOn ? 4ull * kRestirStride * a->pixels() : 0; }

int vh_acc_read_film(const vh_accumulation* a, double* out)
{
    if (!a || !out)
        return fail(VH_ERR_INVALID_ARGUMENT, "null argument");
    HIPTRY(hipMemcpy(out, a->film.p, 24 * a->pixels(), hipMemcpyDeviceToHost));
    return VH_OK;
}

int vh_acc_read_stats(const vh_accumulation* a, double* out)
{
    if (!a || !out)
        return fail(VH_ERR_INVALID_ARGUMENT, "null argument");
    if (!a->stats.p)
        return fail(VH_ERR_UNSUPPORTED, "adaptive sampling is off for this accumulation");
    HIPTRY(hipMemcpy(out, a->stats.p, 24 * a->pixels(), hipMemcpyDeviceToHost));
    return VH_OK;
}

int vh_acc_read_lighting(const vh_accumulation* a, float* out)
{
    if (!a || !out)
        return fail(VH_ERR_INVALID_ARGUMENT, "null argument");
    if (!a->light.p)
        return fail(VH_ERR_UNSUPPORTED, "the accumulation was made without the lighting split");
    HIPTRY(hipMemcpy(out, a->light.p, 4ull * kLightingStride * a->pixels(), hipMemcpyDeviceToHost));
    return VH_OK;
}

// readback::active_pixels with Adaptive::active (lib.rs:154-161), on the FP64 statistics
uint64_t vh_acc_active_pixels(const vh_accumulation* a)
{
    if (!a)
        return 0;
    uint64_t n = a->pixels();
    if (!a->adaptiveOn || !a->stats.p)
        return n;
    std::vector<double> st(3 * n);
    if (hipMemcpy(st.data(), a->stats.p, 24 * n, hipMemcpyDeviceToHost) != hipSuccess)
        return n;
    uint64_t active = 0;
    for (uint64_t p = 0; p < n; p++) {
        double c = st[3 * p];
        if (c < (double)a->adaptive.min_spp) {
            active++;
            continue;
        }
        double mean = st[3 * p + 1] / c;
        double var = std::max(st[3 * p + 2] / c - mean * mean, 0.0) * c / (c - 1.0);
        if (std::sqrt(var / c) > (double)a->adaptive.threshold * std::max(mean, 0.05))
            active++;
    }
    return active;
}

// Accumulation::mean (readback.rs:18-38): the pixel's own sample count under adaptive sampling
int vh_acc_mean(const vh_accumulation* a, float* out)
{
    if (!a || !out)
        return fail(VH_ERR_INVALID_ARGUMENT, "null argument");
    uint64_t n = a->pixels();
    std::vector<double> film(3 * n), st;
    HIPTRY(hipMemcpy(film.data(), a->film.p, 24 * n, hipMemcpyDeviceToHost));
    if (a->stats.p) {
        st.resize(3 * n);
        HIPTRY(hipMemcpy(st.data(), a->stats.p, 24 * n, hipMemcpyDeviceToHost));
    }
    for (uint64_t p = 0; p < n; p++) {
        double inv = st.empty() ? 1.0 / std::max(a->samples, 1u) : 1.0 / std::max(st[3 * p], 1.0);
        for (int k = 0; k < 3; k++)
            out[3 * p + k] = (float)(film[3 * p + k] * inv);
    }
    return VH_OK;
}

int vh_primary_hits(vh_renderer* r, uint32_t width, uint32_t height, uint32_t sample, uint32_t* hits)
{
    if (!r || !hits || width == 0 || height == 0)
        return fail(VH_ERR_INVALID_ARGUMENT, "bad argument");
    uint64_t n = (uint64_t)width * height;
    DevBuf out;
    TRY(check_fits(16 * n));
    HIPTRY(out.alloc(16 * n));
    Params P = base_params(r, width, height, false, false, vh_adaptive{}, false, vh_restir{});
    P.w[6] = sample;
    TRY(launch_primary_hits(r->scene(), P, out.as<uint4>(), r->ctx->stream));
    HIPTRY(hipStreamSynchronize(r->ctx->stream));
    HIPTRY(hipMemcpy(hits, out.p, 16 * n, hipMemcpyDeviceToHost));
    return VH_OK;
}

int vh_trace_rays(vh_renderer* r, uint64_t n, const float* rays, float tmax, uint32_t* hits)
{
    if (!r || (n && (!rays || !hits)))
        return fail(VH_ERR_INVALID_ARGUMENT, "bad argument");
    if (n == 0)
        return VH_OK;
    DevBuf in, out;
    TRY(check_fits(40 * n));
    HIPTRY(in.alloc(24 * n));
    HIPTRY(out.alloc(16 * n));
    HIPTRY(hipMemcpyAsync(in.p, rays, 24 * n, hipMemcpyHostToDevice, r->ctx->stream));
    TRY(launch_trace_rays(r->scene(), in.as<float>(), n, tmax, out.as<uint4>(), r->ctx->stream));
    HIPTRY(hipStreamSynchronize(r->ctx->stream));
    HIPTRY(hipMemcpy(hits, out.p, 16 * n, hipMemcpyDeviceToHost));
    return VH_OK;
}

int vh_set_counters(vh_renderer* r, int on)
{
    if (!r)
        return fail(VH_ERR_INVALID_ARGUMENT, "null renderer");
    if (!on) {
        r->counters.release();
        return VH_OK;
    }
    if (!r->counters.p) {
        HIPTRY(r->counters.alloc(kCounterBanks * 8 * sizeof(unsigned long long)));
        HIPTRY(hipMemsetAsync(r->counters.p, 0, kCounterBanks * 8 * sizeof(unsigned long long), r->ctx->stream));
        HIPTRY(hipStreamSynchronize(r->ctx->stream));
    }
    return VH_OK;
}

int vh_read_counters(vh_renderer* r, uint64_t out[8])
{
    if (!r || !out)
        return fail(VH_ERR_INVALID_ARGUMENT, "null argument");
    if (!r->counters.p)
        return fail(VH_ERR_UNSUPPORTED, "counting is off");
    std::vector<uint64_t> banks(kCounterBanks * 8);
    HIPTRY(hipStreamSynchronize(r->ctx->stream));
    HIPTRY(hipMemcpy(banks.data(), r->counters.p, banks.size() * sizeof(uint64_t), hipMemcpyDeviceToHost));
    HIPTRY(hipMemset(r->counters.p, 0, banks.size() * sizeof(uint64_t)));
    for (int k = 0; k < 8; k++) {
        out[k] = 0;
        for (uint32_t b = 0; b < kCounterBanks; b++)
            out[k] += banks[8 * b + k];
    }
    return VH_OK;
}

int vh_bvh_stats(const vh_renderer* r, uint64_t out[4])
{
    if (!r || !out)
        return fail(VH_ERR_INVALID_ARGUMENT, "null argument");
    out[0] = r->bvh.nodes;
    out[1] = r->bvh.leaves;
    out[2] = r->bvh.depth;
    out[3] = r->bvh.ntri;
    return VH_OK;
}

// RayTracingBackend::render_normals (lib.rs:55; vulkan renderer.rs:855-877)
int vh_render_normals(vh_renderer* r, uint32_t width, uint32_t height, float* normals)
{
    if (!r || !normals)
        return fail(VH_ERR_INVALID_ARGUMENT, "null argument");
    if (width == 0 || height == 0)
        return fail(VH_ERR_SCENE, "image size must be non-zero");
    uint64_t n = (uint64_t)width * height;
    if (n > (1ull << 31))
        return fail(VH_ERR_INVALID_ARGUMENT, "image too large");
    DevBuf out;
    TRY(check_fits(12 * n));
    HIPTRY(out.alloc(12 * n));
    Params P = base_params(r, width, height, false, false, vh_adaptive{}, false, vh_restir{});
    TRY(launch_normals(r->scene(), P, out.as<float>(), r->ctx->stream));
    HIPTRY(hipStreamSynchronize(r->ctx->stream));
    HIPTRY(hipMemcpy(normals, out.p, 12 * n, hipMemcpyDeviceToHost));
    return VH_OK;
}

// RayTracingBackend::render_aovs (lib.rs:57; vulkan renderer.rs:879-924)
int vh_render_aovs(vh_renderer* r, uint32_t width, uint32_t height, float* albedo_depth, uint32_t* object_id, uint32_t* material_id,
                   float* coverage)
{
    if (!r || !albedo_depth || !object_id || !material_id || !coverage)
        return fail(VH_ERR_INVALID_ARGUMENT, "null argument");
    if (width == 0 || height == 0)
        return fail(VH_ERR_SCENE, "image size must be non-zero");
    uint64_t n = (uint64_t)width * height;
    if (n > (1ull << 27))
        return fail(VH_ERR_INVALID_ARGUMENT, "image too large");
    DevBuf ad, oid, mid, cov;
    TRY(check_fits(28 * n));
    HIPTRY(ad.alloc(16 * n));
    HIPTRY(oid.alloc(4 * n));
    HIPTRY(mid.alloc(4 * n));
    HIPTRY(cov.alloc(4 * n));
    Params P = base_params(r, width, height, false, false, vh_adaptive{}, false, vh_restir{});
    TRY(launch_aovs(r->scene(), P, ad.as<float4>(), oid.as<uint32_t>(), mid.as<uint32_t>(), cov.as<float>(), r->ctx->stream));
    HIPTRY(hipStreamSynchronize(r->ctx->stream));
    HIPTRY(hipMemcpy(albedo_depth, ad.p, 16 * n, hipMemcpyDeviceToHost));
    HIPTRY(hipMemcpy(object_id, oid.p, 4 * n, hipMemcpyDeviceToHost));
    HIPTRY(hipMemcpy(material_id, mid.p, 4 * n, hipMemcpyDeviceToHost));
    HIPTRY(hipMemcpy(coverage, cov.p, 4 * n, hipMemcpyDeviceToHost));
    return VH_OK;
}

} // extern "C"
