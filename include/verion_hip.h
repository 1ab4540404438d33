/*
 * verion_hip.h - C ABI of the MI355X path-tracing backend.
 *
 * One-to-one with the Rust backend contract of Verion:
 *   trait RayTracingBackend   crates/verion-gpu/src/lib.rs:12-58
 *   trait Accumulation        crates/verion-gpu/src/lib.rs:74-93
 * A `verion-gpu-hip` crate implements those traits with `extern "C"` calls to the functions below;
 * the scene reaches the library as the flat tables `pack::pack` already produces
 * (crates/verion-gpu/src/pack.rs:110-133, 726-756), so the Rust side keeps the one packer.
 *
 * Conventions
 *   - Every call returns a vh_status (0 = success) unless its return type says otherwise; the
 *     message of the last failure on the calling thread is vh_last_error().
 *   - Handles nest: an accumulation belongs to a renderer, a renderer to a context. Destroy in
 *     reverse order. All calls on one context come from one host thread (the reference session is
 *     not Send: crates/verion-render/src/session.rs:157-158). Nothing calls back.
 *   - Input tables are copied at vh_renderer_create; the caller may free them afterwards.
 *   - Output buffers are allocated by the caller; each call documents its size.
 *   - vh_accumulate blocks until the GPU has finished (lib.rs:50-52).
 */
#ifndef VERION_HIP_H
#define VERION_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error kinds of the Vulkan backend's `Error` enum (crates/verion-gpu-vulkan/src/lib.rs:23-43). */
typedef enum vh_status {
    VH_OK = 0,
    VH_ERR_NO_DEVICE = 1,        /* Error::NoDevice */
    VH_ERR_UNSUPPORTED = 2,      /* Error::Unsupported{missing} */
    VH_ERR_OUT_OF_MEMORY = 3,    /* Error::OutOfMemory{needed, available}, raised before allocating */
    VH_ERR_KERNEL = 4,           /* Error::Shader: a kernel failed to launch or the device reported a fault */
    VH_ERR_SCENE = 5,            /* Error::Scene: invalid tables, empty scene, mid-accumulation switches */
    VH_ERR_INVALID_ARGUMENT = 6, /* null handle, zero image size, bad enum value */
} vh_status;

/* Integrator (lib.rs:97-103). Both give the same image bit for bit. */
typedef enum vh_integrator_kind {
    VH_MEGAKERNEL = 0,
    VH_WAVEFRONT = 1,
} vh_integrator_kind;

/* pack::SceneInfo (pack.rs:55-69). */
typedef struct vh_scene_info {
    uint32_t light_count;  /* lights in the table, the virtual main light included */
    uint32_t emitter_count;
    uint32_t env_width;    /* 0 = no environment map */
    uint32_t env_height;
    float env_pdf_scale;
    float sky[3];
    uint32_t virtual_light;          /* bool */
    uint32_t transparent_background; /* bool */
} vh_scene_info;

/* pack::View (pack.rs:72-82): camera[2] = tan(fov/2), camera[4..16] = camera-to-world 3x4 rows. */
typedef struct vh_view {
    float camera[16];
    float aperture;
    float focus;
    float shift[2];
} vh_view;

/* pack::Packed (pack.rs:110-133). Lengths count elements (floats or words), not bytes.
 * Layouts are those of the bindings in shaders/include/pt_core.slang:12-49. */
typedef struct vh_scene_tables {
    const float* positions;    uint64_t positions_len;   /* xyz per vertex */
    const uint32_t* indices;   uint64_t indices_len;     /* mesh-local vertex indices */
    const float* uvs;          uint64_t uvs_len;         /* uv per vertex */
    const float* vnormals;     uint64_t vnormals_len;    /* xyz per vertex, zero when a mesh has none */
    const uint32_t* instances; uint64_t instances_len;   /* 16 words per instance */
    const float* materials;    uint64_t materials_len;   /* 24 words per material */
    const uint32_t* lights;    uint64_t lights_len;      /* 16 per light, albedo tables, link masks, IES */
    const float* emitters;     uint64_t emitters_len;    /* 16 per emissive triangle */
    const float* emit_cdf;     uint64_t emit_cdf_len;
    const float* env_pixels;   uint64_t env_pixels_len;
    const float* env_marginal; uint64_t env_marginal_len;
    const float* env_cond;     uint64_t env_cond_len;
    const float* texels;       uint64_t texels_len;
    const uint32_t* tex_info;  uint64_t tex_info_len;    /* 4 words per texture */
    const uint32_t* meshes;    uint64_t mesh_count;      /* pack::MeshRange, 4 words per mesh */
    const uint32_t* instance_mesh; uint64_t instance_count; /* the TLAS list: mesh index per instance */
    vh_scene_info info;
    vh_view view;
} vh_scene_tables;

/* Adaptive (lib.rs:110-161). */
typedef struct vh_adaptive {
    float threshold;
    uint32_t min_spp;
} vh_adaptive;

/* Restir (lib.rs:131-149). */
typedef struct vh_restir {
    uint32_t di;         /* bool */
    uint32_t gi;         /* bool; ReSTIR GI (pt_core.slang:901-980, 1566-1617): reconnection reuse of the
                            first bounce's hit; off while the lighting split is on, like the reference */
    uint32_t candidates;
    uint32_t history;
    uint32_t neighbours; /* random neighbours reused per frame besides the pixel's own reservoir, 0..16;
                            0 = 2, the reference's fixed RESTIR_NEIGHBOURS (pt_core.slang:706), whose
                            random stream is then unchanged. No counterpart in the reference. */
} vh_restir;

/* MemoryUsage (lib.rs:242-249). */
typedef struct vh_memory_usage {
    uint64_t budget;
    uint64_t allocated;
    uint64_t uploaded;
} vh_memory_usage;

/* UpdateReport (lib.rs:65-71). bvh_work has no counterpart in the reference: the LBVH is one tree
 * over world-space triangles, and a TLAS or BLAS rebuild of the plan becomes a refit of it (or a
 * rebuild when the refit's surface-area cost grows past 1.5 times the built tree's). */
typedef struct vh_update_report {
    uint64_t blases_built;   /* meshes whose triangles were redone (every mesh for a full update) */
    uint32_t tlas_built;
    uint64_t bytes_uploaded;
    uint32_t bvh_work;       /* 0: LBVH untouched, 1: refit in place, 2: built anew */
} vh_update_report;

/* pack::UpdatePlan (pack.rs:758-821): what a batch of scene changes makes the backend redo;
 * UpdatePlan::new(&Changes) derives it (pack.rs:785-815). */
typedef struct vh_update_plan {
    uint32_t full;       /* meshes or instances added or re-pointed: everything, as at creation */
    uint32_t materials;  /* material table; texels and texture info when they differ */
    uint32_t lights;     /* light table (albedo tables, link masks, IES), sky, environment map when it differs */
    uint32_t emitters;   /* emitter table and CDF */
    uint32_t instances;  /* instance table */
    uint32_t vertices;   /* positions, uvs and vertex normals rewritten at the same sizes */
    uint32_t tlas;       /* instance transforms changed */
    uint32_t camera;     /* the view */
    const uint32_t* blases; uint64_t blas_count; /* meshes whose vertices changed */
} vh_update_plan;

typedef struct vh_context vh_context;
typedef struct vh_renderer vh_renderer;
typedef struct vh_accumulation vh_accumulation;

/* --- errors --------------------------------------------------------------------------------- */

/* Message of the last failure on this thread ("" when none). Valid until the next failing call. */
const char* vh_last_error(void);
/* Error::OutOfMemory{needed, available} of the last VH_ERR_OUT_OF_MEMORY on this thread. */
uint64_t vh_last_error_needed(void);
uint64_t vh_last_error_available(void);

/* Host-only helper (no device work): the cutpoint table the renderer builds over the emitter power
 * CDF (`emit_cdf`, pack.rs) so that a draw u selects the same emitter as the binary search of
 * pt_core.slang:1343-1351 while searching only between two cutpoints: guide[k] (k = 0 .. 2^bits)
 * = the first index whose CDF value reaches k / 2^bits, or n - 1 when none does. `guide` holds
 * 2^bits + 1 words; bits <= 16; the CDF must be non-decreasing. No counterpart in the reference. */
int vh_cdf_guide(const float* cdf, uint32_t n, uint32_t bits, uint32_t* guide);

/* --- context (RayTracingBackend::context / memory, lib.rs:19-20) ------------------------------ */

/* Opens device 0 (the current HIP device). VH_ERR_NO_DEVICE without one, VH_ERR_UNSUPPORTED when
 * it is not gfx950. */
int vh_context_create(vh_context** out);
void vh_context_destroy(vh_context* ctx);
int vh_memory(const vh_context* ctx, vh_memory_usage* out);

/* --- renderer (RayTracingBackend::new / update / setters, lib.rs:21-45) ----------------------- */

/* Uploads the tables, computes the world-space triangles and builds the LBVH on the GPU
 * (this replaces the driver BLAS/TLAS of crates/verion-gpu-vulkan/src/accel.rs:153-233).
 * VH_ERR_SCENE when the scene has no instances (pack.rs:727-729) or a table is inconsistent. */
int vh_renderer_create(vh_context* ctx, const vh_scene_tables* tables, vh_renderer** out);
void vh_renderer_destroy(vh_renderer* r);
/* RayTracingBackend::update (lib.rs:27, vulkan renderer.rs:404-502): `tables` is the whole scene
 * after the edit; the plan says what changed and only that is uploaded and redone. NULL (or
 * plan->full) replaces every table and builds the LBVH anew. A partial plan must keep every table
 * size but the material, light, emitter and texture tables (VH_ERR_INVALID_ARGUMENT otherwise).
 * The call waits for the device first: no accumulation may be running. */
int vh_renderer_update(vh_renderer* r, const vh_scene_tables* tables, const vh_update_plan* plan, vh_update_report* report);
/* set_camera (lib.rs:30): the view of pack::View::new. */
int vh_set_camera(vh_renderer* r, const vh_view* view);
int vh_set_multiscatter(vh_renderer* r, int on);
int vh_set_integrator(vh_renderer* r, int integrator);
int vh_integrator(const vh_renderer* r);
int vh_set_stage_timing(vh_renderer* r, int on);
/* NULL switches adaptive sampling off. */
int vh_set_adaptive(vh_renderer* r, const vh_adaptive* adaptive);
/* NULL switches ReSTIR off (Renderer::set_restir, lib.rs:131-149). */
int vh_set_restir(vh_renderer* r, const vh_restir* restir);
/* Random-stream seed for accumulations started afterwards. 0 (the default) is the reference's own
 * stream, pcg(pixel ^ pcg(sample * 0x9E3779B9 + 0x7F4A7C15)) (pt_core.slang:651); any other seed
 * XORs pcg(seed ^ 0x5851F42D) into the pixel hash. No counterpart in the reference (it has no seed). */
int vh_set_seed(vh_renderer* r, uint32_t seed);
/* Grid of the wavefront's grid-stride extend/shade kernels, in 256-thread blocks; 0 = 8 per CU.
 * No counterpart in the reference (its dispatch size is the queue length, wf_args.slang:19); set
 * by the autotuner. Images do not depend on it. */
int vh_set_wavefront_blocks(vh_renderer* r, uint32_t blocks);
/* Order in which the wavefront takes pixels at depth 0: tiles of tile_width x tile_height pixels
 * (8 x 8 by default), or 0 x 0 for scanlines. No counterpart in the reference (it dispatches
 * pixels in 64-wide groups of a row, wavefront.rs:11). Images do not depend on it. */
int vh_set_wavefront_tiles(vh_renderer* r, uint32_t tile_width, uint32_t tile_height);
/* Sample passes of one vh_accumulate call that the plain wavefront (no ReSTIR, lighting split or
 * adaptive sampling, stage timing off) keeps in flight at once, 1 to 4 (default 3; 1 = one pass at
 * a time). Each pass in flight has its own stream, wavefront buffers and FP64 film, added to the
 * accumulation's film in sample order, so the film's bits do not depend on it. No counterpart in
 * the reference (its wavefront records one pass after another into one submission, vulkan renderer.rs:681-760). */
int vh_set_wavefront_pipeline(vh_renderer* r, uint32_t passes);

/* --- accumulation (new_accumulation / accumulate, lib.rs:47-52; trait Accumulation :74-93) ---- */

/* VH_ERR_INVALID_ARGUMENT for a zero size (vulkan renderer.rs:510-512). */
int vh_accumulation_create(vh_renderer* r, uint32_t width, uint32_t height, int lighting, vh_accumulation** out);
void vh_accumulation_destroy(vh_accumulation* acc);
/* Traces `samples` more paths per pixel (at most, under adaptive sampling) and waits for the GPU.
 * VH_ERR_SCENE when adaptive sampling or ReSTIR was switched since the first sample. */
int vh_accumulate(vh_renderer* r, vh_accumulation* acc, uint32_t samples);

uint32_t vh_acc_samples(const vh_accumulation* acc);
double vh_acc_gpu_seconds(const vh_accumulation* acc);
/* Extend and shade GPU seconds of the wavefront when stage timing was on; VH_ERR_UNSUPPORTED otherwise.
 * Where one launch traces closest and shadow rays (opaque scenes with deferred shadow rays), its
 * time counts as extend. */
int vh_acc_stage_seconds(const vh_accumulation* acc, double out[2]);
/* Debug export (no counterpart in the reference): the same GPU seconds split by kernel: extend,
 * shade, the deferred shadow rays, their resolve, and trace: depth d's closest rays and depth
 * d - 1's shadow rays in one launch, which opaque scenes with deferred shadow rays use in place of
 * extend and shadow (shadow and resolve are 0 when shade traces its shadow rays itself: the
 * lighting split, more delta lights than records per path). */
int vh_acc_kernel_seconds(const vh_accumulation* acc, double out[5]);
uint64_t vh_acc_wavefront_bytes(const vh_accumulation* acc);
uint64_t vh_acc_restir_bytes(const vh_accumulation* acc);
/* Pixels the next sample would trace (readback::active_pixels, readback.rs:51-61). */
uint64_t vh_acc_active_pixels(const vh_accumulation* acc);
/* Raw running sums, FP64: 3 doubles per pixel (radiance sum), row-major from the top-left. */
int vh_acc_read_film(const vh_accumulation* acc, double* out);
/* Adaptive statistics, FP64: samples, luminance sum, square sum per pixel. VH_ERR_UNSUPPORTED without adaptive. */
int vh_acc_read_stats(const vh_accumulation* acc, double* out);
/* Lighting split sums, 65 floats per pixel (pt_core.slang:39-47). VH_ERR_UNSUPPORTED without lighting. */
int vh_acc_read_lighting(const vh_accumulation* acc, float* out);
/* Accumulation::mean (readback.rs:18-38): 3 floats per pixel. */
int vh_acc_mean(const vh_accumulation* acc, float* out);

/* --- primary-hit AOVs (render_normals / render_aovs, lib.rs:54-57) ---------------------------- */

/* render_normals (lib.rs:55; shaders/normals.slang:39-110): the pixel-centre ray's interpolated
 * vertex normal (the geometric normal for meshes without normals; normal maps not applied), on the
 * geometric normal's side and facing the viewer; 0 where the ray misses. 3 floats per pixel,
 * row-major from the top-left. VH_ERR_SCENE for a zero image size (vulkan renderer.rs:856-858). */
int vh_render_normals(vh_renderer* r, uint32_t width, uint32_t height, float* normals);
/* render_aovs (lib.rs:57, struct Aovs lib.rs:165-176; shaders/aov.slang:84-188), per pixel from the
 * pixel-centre ray: albedo_depth = base colour x level-0 base texture (3 floats) and the distance
 * along the camera's forward axis; object_id = instance + 1, material_id = material + 1 (0 on a
 * miss); coverage = fraction of 16 stratified rays (4x4 grid) that hit, through up to 8
 * translucent layers. VH_ERR_SCENE for a zero image size. */
int vh_render_aovs(vh_renderer* r, uint32_t width, uint32_t height, float* albedo_depth,
                   uint32_t* object_id, uint32_t* material_id, float* coverage);

/* --- debug exports (no counterpart in the reference) --------------------------------------- */

/* Closest hit of the camera ray of sample `sample` for every pixel (generatePathFor,
 * pt_core.slang:645-682, then traceClosest :684-700): hit flag, instance, primitive, t bits per
 * pixel (4 words). Used by the bit-exact primary-hit gate. */
int vh_primary_hits(vh_renderer* r, uint32_t width, uint32_t height, uint32_t sample, uint32_t* hits);
/* Closest hits of `n` arbitrary rays (org xyz, dir xyz per ray; 4 words out per ray). */
int vh_trace_rays(vh_renderer* r, uint64_t n, const float* rays, float tmax, uint32_t* hits);
/* Host-only (no device work): the traversal's limits: stack entries per lane, of which the first
 * ones live in LDS, the deepest LBVH (levels of 4-wide nodes) a renderer accepts (deeper trees are
 * VH_ERR_SCENE at vh_renderer_create / vh_renderer_update, since a traversal of a tree of depth D may
 * hold 3 (D - 1) + 4 entries), and the most triangles per leaf. */
int vh_trace_limits(uint32_t out[4]);
/* LBVH shape: nodes, leaves, maximum depth, triangles. */
int vh_bvh_stats(const vh_renderer* r, uint64_t out[4]);
/* Traversal counting for the roofline's algorithmic bytes: while on, every ray query adds to
 * [closest rays, their nodes visited, their triangles tested, shadow rays, nodes, triangles]. */
int vh_set_counters(vh_renderer* r, int on);
/* Reads the counters (the six above, then closest rays that hit, then 0) and clears them. VH_ERR_UNSUPPORTED while counting is off. */
int vh_read_counters(vh_renderer* r, uint64_t out[8]);

/* --- neural radiance cache (no counterpart in the reference: its milestone M7 plans one,
 *     docs/implementation/M7-nrc-adaptive-integrator/README.md rows 2-5) ------------------------
 * A 96-64-64-64-64-64-16 MLP (ReLU hidden, linear output, first 3 outputs RGB) over a 12-frequency
 * position encoding, degree-3 real SH of the outgoing direction, the shading normal and the
 * roughness; BF16 MFMA with FP32 accumulation, FP32 master weights, Adam, relative L2 loss.
 * Parameters are 23,888 floats: W0 (96x64 row-major [k][n]), b0, W1..W4 (64x64), b1..b4, W5 (64x16), b5.
 * Raw inputs are 12 floats per query: position (in [0,1]^3 for the debug calls), roughness,
 * direction xyz, 0, normal xyz, 0. */
typedef struct vh_nrc_config {
    uint32_t train_paths;     /* training paths per frame (65,536) */
    uint32_t train_steps;     /* Adam steps per frame */
    float learning_rate;      /* 1e-2 */
    uint32_t terminate_depth; /* render paths end in the cache at their first opaque vertex at this depth (1..7) */
    float ema_decay;          /* inference uses the bias-corrected exponential moving average of the
                                 parameters over Adam steps with this decay; 0 = the parameters themselves */
    uint32_t warmup_frames;   /* training-only frames (paths + Adam steps, no film) before the first
                                 rendered frame of the cache; their time counts as the render's */
    uint32_t minibatch;       /* bool: Adam step k of a frame takes slice k of train_steps equal slices of
                                 the frame's samples (each sample once per frame) instead of all of them */
} vh_nrc_config;
typedef struct vh_nrc vh_nrc;

/* NULL config = {65536, 8, 1e-2, 1, 0.95, 16, 0}. The cache starts from vh_nrc_init_params(0). */
int vh_nrc_create(vh_renderer* r, const vh_nrc_config* config, vh_nrc** out);
void vh_nrc_destroy(vh_nrc* nrc);
/* Host-only (no device work): the cache's seeded initial parameters (n = 23888). He-uniform weights
 * (uniform in +-sqrt(6 / fan_in)), zero biases, the output layer scaled by 0.1 with its 13 unused
 * outputs zero; weight p draws u = (pcg(pcg(seed ^ 0xA511E9B3) + p) >> 8) * 2^-24 with the pcg of
 * pt_core.slang:69-74 and takes lim * (2u - 1). */
int vh_nrc_init_params(uint32_t seed, float* params, uint64_t n);
/* vh_nrc_set_params with vh_nrc_init_params(seed): a fresh cache (Adam and the average restart). */
int vh_nrc_init(vh_nrc* nrc, uint32_t seed);
/* Replaces the FP32 master parameters (n = 23888), their moving average, and resets Adam. */
int vh_nrc_set_params(vh_nrc* nrc, const float* params, uint64_t n);
int vh_nrc_get_params(vh_nrc* nrc, float* params, uint64_t n);
/* Fused inference of n raw queries with the parameters the integrator uses (the moving average):
 * 3 floats out per query. */
int vh_nrc_infer(vh_nrc* nrc, uint64_t n, const float* raw12, float* out3);
/* Gradient of the mean relative L2 loss over a batch (23888 floats) and the loss; no update. */
int vh_nrc_gradient(vh_nrc* nrc, uint64_t n, const float* raw12, const float* target3, float* grad, float* loss);
/* vh_nrc_gradient on the caller's device buffers, allocating nothing. No counterpart in the
 * reference: for callers that cannot allocate device memory (serving loops, harnesses that forbid
 * it). raw12 (12 floats per sample), target3 (3 per sample), grad (23888 floats) and loss (1 float,
 * may be NULL) are device pointers on the cache's device; n <= 65536 (VH_ERR_INVALID_ARGUMENT
 * beyond). The samples are staged in a 4 MiB device array of the library that every call on a
 * device shares and overwrites: calls that overlap on one device race. Blocks until done; the
 * result equals vh_nrc_gradient's bit for bit. */
int vh_nrc_gradient_noalloc(vh_nrc* nrc, uint64_t n, const float* raw12, const float* target3, float* grad, float* loss);
/* `steps` Adam steps on one batch; the loss of the last step's forward pass. */
int vh_nrc_train(vh_nrc* nrc, uint64_t n, const float* raw12, const float* target3, uint32_t steps, float* loss);
/* The NRC integrator: per frame, train_paths plain training paths on random pixels (every opaque
 * vertex becomes a sample), train_steps Adam steps, then one sample of every pixel that ends in the
 * cache at its first opaque vertex at terminate_depth, into `acc` (FP64 film). Not with the
 * lighting split, adaptive sampling or ReSTIR (VH_ERR_UNSUPPORTED). */
int vh_nrc_accumulate(vh_nrc* nrc, vh_accumulation* acc, uint32_t frames);
/* Last loss, last training-sample count, total training seconds (the warm-up, and the time the
 * render stream waits for a frame's Adam steps, which otherwise run beside its path tracing),
 * total rendering seconds; the two add up to the frames' GPU time. */
int vh_nrc_stats(const vh_nrc* nrc, double out[4]);

/* --- batched jobs (no counterpart in the reference: its session renders one job at a time,
 *     crates/verion-render/src/session.rs:323-336) ------------------------------------------------
 * Every job of a batch is resident in one shared wavefront and every sample pass advances all jobs
 * still sampling; a job's film equals rendering it alone with vh_accumulate (wavefront) bit for
 * bit. Plain integrator only (no lighting split, ReSTIR or radiance cache). The renderers of the
 * jobs must outlive the batch and must not be updated while it exists. */
typedef struct vh_batch vh_batch;
typedef struct vh_batch_job {
    vh_renderer* renderer; /* the scene */
    uint32_t width, height;
    uint32_t spp;          /* samples per pixel (the most, with adaptive sampling) */
    vh_adaptive adaptive;  /* threshold 0 = off */
    uint32_t seed;         /* random-stream seed, as vh_set_seed (0 = the reference stream) */
} vh_batch_job;
/* Launch shape of the batch kernels: blocks of the grid-stride kernels, and the depth-0 pixel
 * order (tiles of tile_width x tile_height pixels, or 0 x 0 for scanlines). */
typedef struct vh_batch_tuning {
    uint32_t grid_blocks;
    uint32_t tile_width;
    uint32_t tile_height;
} vh_batch_tuning;

/* 1 to 255 jobs; allocates every film and the shared wavefront. Tuning starts automatic. */
int vh_batch_create(vh_context* ctx, const vh_batch_job* jobs, uint32_t count, vh_batch** out);
void vh_batch_destroy(vh_batch* batch);
/* Clears films, statistics and the pass counter (the next run starts from sample 0). */
int vh_batch_reset(vh_batch* batch);
/* Fixes the tuning; NULL returns to automatic tuning (done during the next run from its start). */
int vh_batch_set_tuning(vh_batch* batch, const vh_batch_tuning* tuning);
int vh_batch_get_tuning(const vh_batch* batch, vh_batch_tuning* tuning);
/* Runs sample passes until every job has its samples (at most max_passes of them; 0 = no limit),
 * blocking until the GPU is done. graph = 1 replays one captured pass as a hipGraph. With
 * automatic tuning the first passes try one candidate each (the image does not depend on them)
 * and the fastest per traced path runs the rest. */
int vh_batch_run(vh_batch* batch, int graph, uint32_t max_passes);
/* out[0] = GPU seconds from the start of the first run to the end of the job's last pass with
 * work (-1 while it still has work), out[1] = its passes with work, out[2] = paths traced (depth-0
 * paths over all passes), out[3] = GPU seconds of all runs. */
int vh_batch_job_stats(const vh_batch* batch, uint32_t job, double out[4]);
/* The candidates the automatic tuning tried and their milliseconds per million paths. */
int vh_batch_tuning_report(const vh_batch* batch, vh_batch_tuning* tried, double* ms_per_mpath, uint32_t cap, uint32_t* count);
/* Running sums of a job: film (3 doubles per pixel), statistics (3 doubles, adaptive jobs only). */
int vh_batch_read_film(const vh_batch* batch, uint32_t job, double* film);
int vh_batch_read_stats(const vh_batch* batch, uint32_t job, double* stats);
/* Mean image of a job (readback.rs:18-38): 3 floats per pixel. */
int vh_batch_mean(const vh_batch* batch, uint32_t job, float* out);

#ifdef __cplusplus
}
#endif

#endif /* VERION_HIP_H */
