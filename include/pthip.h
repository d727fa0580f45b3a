/*
 * pthip.h - C ABI of the MI355X path-tracing back end (libpthip.so).
 *
 * This boundary replaces the reference's Vulkan ray-tracing back end: the C++ class
 * PathTracer (src/pathtrace/pathtrace.h:43-103) as driven by PathTraceDispatcher
 * (src/pathtrace/ptdispatch.h:24-55), its GPU contract (descriptor bindings
 * src/pathtrace/pathtrace.cpp:36-69, uniform block src/pathtrace/pathtrace.h:21-41) and the
 * driver-built acceleration structures (src/pathtrace/accstruct.cpp:69-286).
 *
 * Plain C types only. Every call returns PT_OK (0) or a negative PT_E* status; no C++
 * exception crosses this boundary. The message of the last failure on a context is
 * available from pt_last_error(ctx); failures of pt_create are reported by
 * pt_last_error(NULL) (thread-local).
 *
 * Threading: one context per host thread. A context owns one HIP stream and all device
 * memory it uses (scene copy, BLAS/TLAS, wavefront queues, accumulation image).
 */
#ifndef PTHIP_H_INCLUDED
#define PTHIP_H_INCLUDED

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PT_OK 0
#define PT_EINVAL (-1)  /* bad argument (e.g. spp == 0, mirrors pathtrace.cpp:400-401) */
#define PT_ENOMEM (-2)  /* device allocation failed */
#define PT_EDEVICE (-3) /* a HIP runtime call or kernel failed, or a traversal stack overflowed (the
                         * rays lost BVH nodes: pt_render, pt_read_*, pt_synchronize and pt_get_stats
                         * report it until the accumulation restarts with pt_set_camera,
                         * pt_set_max_bounces, pt_cache_configure or pt_build_light_tables; the
                         * pt_trace_* call that overflowed reports its own) */
#define PT_ESTATE (-4)  /* call not valid in the current state (e.g. no camera set) */
#define PT_ENODEV (-5)  /* no HIP device present */

/* Byte sizes of the packed scene records; identical to the reference's GPU layout. */
#define PT_VERTEX_BYTES 32u    /* src/scene/mesh.h:10-14: pos f32x3 | normal f32x3 | uv f32x2 */
#define PT_INSTANCE_BYTES 128u /* src/scene/scenebuffer.h:13-20 (InstanceData)               */
#define PT_MATERIAL_BYTES 96u  /* src/scene/material.h:24-29 (PbrMaterial)                   */
#define PT_LIGHT_BIN_WORDS 10u /* share/constants.h:27 */
#define PT_EMITTER_BIN_WORDS 6u /* share/constants.h:28 */
#define PT_NO_EMITTER 0xFFFFFFFFu

/* One mesh of the packed scene: the region of the vertex/index arrays it owns
 * (SceneBuffer::get_start_indices, src/scene/scenebuffer.h:53-54). One BLAS is built per mesh,
 * as AccelerationStructure::create_blas does (src/pathtrace/accstruct.cpp:69-164). */
typedef struct pt_mesh_desc {
    uint32_t vertex_start; /* in vertices */
    uint32_t index_start;  /* in u32 indices (3 per triangle) */
    uint32_t tri_count;
    uint32_t vertex_count;
} pt_mesh_desc;

/* One texture of the packed scene (SceneTextures, src/scene/texturedata.cpp:30-46): 8-bit
 * channels at texels + offset, rows top to bottom; channels 1 (R8), 2 (R8G8) or 4 (RGBA8);
 * srgb != 0 for base-colour and emission maps with four channels (VK_FORMAT_R8G8B8A8_SRGB),
 * decoded to linear per texel before filtering. Sampled bilinearly with REPEAT addressing at
 * mip level 0 (src/graphics/sampler.cpp:8-22). */
typedef struct pt_texture_desc {
    uint64_t offset;
    uint32_t width;
    uint32_t height;
    uint32_t channels;
    uint32_t srgb;
} pt_texture_desc;

/* The packed scene, exactly the regions of the reference's SceneBuffer
 * (src/scene/scenebuffer.cpp:65-168) plus the instance->mesh map the TLAS needs
 * (src/pathtrace/accstruct.cpp:197-216). Host memory; only read during pt_create. */
typedef struct pt_scene_blob {
    const void* vertices; /* vertex_count x 32 B */
    uint64_t vertex_count;
    const uint32_t* indices; /* index_count u32, mesh-relative vertex numbers */
    uint64_t index_count;
    const void* instances; /* instance_count x 128 B InstanceData */
    uint32_t instance_count;
    const void* materials; /* material_count x 96 B PbrMaterial */
    uint32_t material_count;
    const uint32_t* light_words; /* TableLightSampler::copy output (lightsample.cpp:209-221) */
    uint64_t light_word_count;
    uint32_t light_count; /* bins of the light alias table (SceneBuffer::get_light_count) */
    uint32_t mesh_count;
    const pt_mesh_desc* meshes;
    const uint32_t* instance_mesh; /* instance_count entries: mesh (BLAS) of each instance */
    float envcol[4];               /* linear environment colour (UBO environment_colour) */
    const pt_texture_desc* textures; /* texture_count entries (the materials' map indices) */
    uint32_t texture_count;
    uint32_t reserved;
    const uint8_t* texels;
    uint64_t texel_bytes;
} pt_scene_blob;

/* Camera block: the same content PathTracer::set_camera writes into the uniform block
 * (src/pathtrace/pathtrace.cpp:381-396): inverse view and inverse projection, row-major. */
typedef struct pt_camera {
    float inv_view[16];
    float inv_proj[16];
    float near_clip;
    float far_clip;
    float focus_dist;
    float lens_radius;
    float exposure;
} pt_camera;

/* A ray for pt_trace_rays (the TraceRay call sites: pathtrace.rg.hlsl:35,51,
 * pathtrace.ch.hlsl:199). */
typedef struct pt_ray {
    float origin[3];
    float tmin;
    float dir[3];
    float tmax;
} pt_ray;

/* A hit: InstanceID(), PrimitiveIndex(), RayTCurrent() and the two barycentrics the
 * closest-hit shader receives (pathtrace.ch.hlsl:228-240). instance == PT_NO_HIT on a miss. */
#define PT_NO_HIT 0xFFFFFFFFu
typedef struct pt_hit {
    float t;
    float u;
    float v;
    uint32_t instance;
    uint32_t primitive;
} pt_hit;

/* Counters and times accumulate over the pt_render calls since the last pt_set_camera,
 * pt_set_max_bounces, pt_set_option or pt_reset_stats. */
typedef struct pt_stats {
    uint64_t extension_rays; /* rays traced by closest-hit traversal */
    uint64_t shadow_rays;    /* rays traced by any-hit traversal */
    uint64_t paths;          /* pixel samples */
    double render_ms;        /* device time from the first to the last pt_render (HIP events) */
    double bvh_build_ms;     /* device time of the BLAS+TLAS build in pt_create */
    uint64_t bvh_bytes;      /* bytes of nodes + triangle records */
    uint32_t sample_index;   /* next sample index (PathTracer::get_sample_index) */
    uint32_t max_bounces;
    uint32_t width;
    uint32_t height;
    uint32_t node_count;
    uint32_t tri_record_count;
    /* With PT_OPT_COUNT_FETCHES: 128-byte node and 48-byte triangle-record fetches. */
    uint64_t ext_node_fetches;
    uint64_t ext_tri_fetches;
    uint64_t shadow_node_fetches;
    uint64_t shadow_tri_fetches;
    /* With PT_OPT_PROFILE_KERNELS: summed HIP-event durations and launch counts of the
     * extend (closest hit), shade, shadow (any hit) and refill (ended paths out, next paths
     * in) stages of the last pt_render. */
    double extend_ms;
    double shade_ms;
    double shadow_ms;
    uint32_t extend_launches;
    uint32_t shade_launches;
    uint32_t shadow_launches;
    uint32_t refill_launches;
    double refill_ms;
    /* With PT_OPT_COUNT_FETCHES: the node fetches of the top-level hierarchy among the node
     * fetches above, and the instances (BLAS) entered. */
    uint64_t ext_tlas_node_fetches;
    uint64_t ext_instance_entries;
    uint64_t shadow_tlas_node_fetches;
    uint64_t shadow_instance_entries;
} pt_stats;

/* Options for pt_set_option (instrumentation; no counterpart in the reference, whose only
 * instrumentation is the wall-clock print of src/pathtrace/ptdispatch.cpp:65-70). */
#define PT_OPT_PROFILE_KERNELS 1u /* value != 0: time every stage launch with HIP events */
#define PT_OPT_COUNT_FETCHES 2u   /* value != 0: count BVH node / triangle fetches per ray */
#define PT_OPT_GRAPHS 3u          /* value != 0 (default): replay each wavefront iteration from a
                                   * captured HIP graph (not while kernels are being timed) */
/* Launch and hierarchy configuration (no counterpart in the reference, whose traversal
 * belongs to the driver). Images do not depend on them, except that another hierarchy (leaf
 * size, node width, builder) can resolve exact ties between equally distant triangles
 * differently. */
#define PT_OPT_BLOCKS_PER_CU 4u   /* traversal grid: blocks of 256 threads per CU (1..64, default 6) */
#define PT_OPT_POOL_SLOTS 5u      /* path pool slots (1024..2^26; default min(16M, max(8 px, 64K)); a batch's
                                   * min(16M, max(4 x its pixels, 64K)) */
#define PT_OPT_LDS_STACK 6u       /* traversal stack entries per lane kept in LDS (8, 16, 32) */
#define PT_OPT_LEAF_SIZE 7u       /* max triangles per BLAS leaf (1..16, default 1): rebuilds the BVH */
#define PT_OPT_SORT_HITS 8u       /* value != 0: shade extension hits in (scene, material) order
                                   * after a radix sort (default 0; iterations then launch
                                   * without graphs) */
#define PT_OPT_NODE_WIDTH 9u      /* children per BVH node the SAH collapse may use (2..4, default 4;
                                   * every node is one 128-byte line): rebuilds the BVH */
#define PT_OPT_BVH_BUILDER 10u    /* binary tree under the SAH collapse: 0 PLOC clustering (default),
                                   * 1 Karras radix tree over the Morton codes: rebuilds the BVH */
#define PT_OPT_TLAS_OPEN 12u      /* levels of each instance's BLAS the TLAS may open (0..2, default 0):
                                   * an instance whose BLAS children have much tighter world boxes than
                                   * its own is entered through those subtrees (partial re-braiding);
                                   * rebuilds the BVH */
#define PT_OPT_TLAS_BOXES 13u     /* boxes of the TLAS leaves (0..3, default 3): 0 the transformed corners
                                   * of the BLAS node's box; 1, 2 the union of its transformed descendant
                                   * boxes that many levels down; 3 an instance's transformed vertices
                                   * (level 2 for entries at opened subtrees): rebuilds the BVH */
#define PT_OPT_TLAS_BUILDER 14u   /* TLAS builder: 0 the BLAS builder on the device, 1 (default) a full
                                   * SAH sweep with an SAH-optimal collapse on the host, for up to 16384
                                   * TLAS leaves (more use 0): rebuilds the BVH */
#define PT_OPT_SPILL_DEPTH 11u    /* traversal stack entries per lane spilled to global memory beyond the
                                   * LDS ones (0..112, default 112); a ray whose stack would grow deeper
                                   * loses nodes, which the calls below report as PT_EDEVICE */

typedef struct pt_ctx pt_ctx;

/* PathTracer::PathTracer + SceneBuffer upload + AccelerationStructure build
 * (src/pathtrace/pathtrace.cpp:81-116, accstruct.cpp:288-296). Builds one BVH per mesh and
 * a TLAS over the instances on the device, and the light sampler's alias tables
 * (pt_build_light_tables) when light_words has the packed layout. width/height (1..65535
 * each) size the accumulation image (PathTraceDispatcher ctor, src/pathtrace/ptdispatch.cpp:26-27). */
int pt_create(const pt_scene_blob* scene, uint32_t width, uint32_t height, pt_ctx** out);

/* PathTracer::set_camera (src/pathtrace/pathtrace.cpp:381-396). Resets sample_index to 0. */
int pt_set_camera(pt_ctx* ctx, const pt_camera* camera);

/* PathTracer::set_max_bounces (src/pathtrace/pathtrace.cpp:403-407). Resets sample_index. */
int pt_set_max_bounces(pt_ctx* ctx, uint32_t max_bounces);

/* PathTraceDispatcher::render (src/pathtrace/ptdispatch.cpp:97-134): traces spp further
 * samples per pixel into the accumulation image, one sample index after another exactly as
 * the reference's 1-spp submits do (running mean of correct_radiance(L*exposure)), and
 * advances sample_index by spp (PathTracer::update_uniforms, pathtrace.cpp:511-516).
 * spp == 0 -> PT_EINVAL (pathtrace.cpp:400-401). Asynchronous on the context's stream, and
 * returns once every path of the call has started: the last paths may still be in flight
 * while the next pt_render's paths enter the pool (two job slots), so a job's tail overlaps
 * the next job's start. Every call that reads results or changes what the paths use
 * (pt_read_*, pt_get_stats, pt_set_*, pt_trace_*, pt_cache_*, pt_autotune, pt_synchronize,
 * pt_destroy) first finishes them; the image equals one pt_render of all the samples. */
int pt_render(pt_ctx* ctx, uint32_t spp);

/* Readback of the accumulation image (ptdispatch.cpp:77-94 copy_image + vkMapMemory):
 * width*height RGBA32F rows of row_pitch bytes into caller memory. Synchronises. */
int pt_read_accumulation(pt_ctx* ctx, float* rgba, size_t row_pitch);

/* The PNG output path of the headless dispatcher: PathTraceDispatcher::put_result's ACES tone
 * mapper over the accumulation (shaders/compute/tonemap.cs.hlsl with src_is_srgb = false,
 * dst_is_srgb = true, src/pathtrace/ptdispatch.cpp:29,155-180) and write_png's 8-bit
 * conversion (uchar)clamp(x * 255, 0, 255) (src/io/imagewrite.cpp:39-62), into width x height
 * RGBA8 pixels, rows row_pitch bytes apart. The accumulation is not modified (the reference
 * tone-maps it in place before writing the file). Synchronises. */
int pt_read_tonemapped(pt_ctx* ctx, uint8_t* rgba8, size_t row_pitch);

/* Primary-ray hits of one sample index: per pixel, the first TraceRay of the raygen shader
 * (pathtrace.rg.hlsl:16-35) with its camera ray and [near, far] interval. hits has
 * width*height entries, row-major. Synchronises. */
int pt_trace_primary(pt_ctx* ctx, uint32_t sample_index, pt_hit* hits);

/* Arbitrary rays through the same traversal: closest hit (any_hit == 0, the TraceRay of
 * pathtrace.rg.hlsl:51) or first hit (any_hit != 0, the shadow TraceRay of
 * pathtrace.ch.hlsl:199 with ACCEPT_FIRST_HIT_AND_END_SEARCH). Synchronises. */
int pt_trace_rays(pt_ctx* ctx, const pt_ray* rays, uint32_t count, pt_hit* hits, int any_hit);

/* pt_trace_rays without any device allocation: no counterpart in the reference. For callers
 * that cannot allocate (tight serving loops, code that must not call hipMalloc): the rays and
 * hits are staged through a fixed device array of the library (one per GPU, PT_NOALLOC_MAX_RAYS
 * rays and hits, about 13 MiB, held until the process ends). count > PT_NOALLOC_MAX_RAYS ->
 * PT_EINVAL. Results equal pt_trace_rays'. Calls that overlap on one GPU (from several contexts
 * or threads) race for the array. Synchronises. */
#define PT_NOALLOC_MAX_RAYS 262144u
int pt_trace_rays_noalloc(pt_ctx* ctx, const pt_ray* rays, uint32_t count, pt_hit* hits, int any_hit);

/* Counters and timings (the reference prints its timing in ptdispatch.cpp:65-70). */
int pt_get_stats(pt_ctx* ctx, pt_stats* out);

/* Restarts the pt_stats counters at the next pt_render (no counterpart in the reference). */
int pt_reset_stats(pt_ctx* ctx);

/* Sets an instrumentation option (PT_OPT_*). */
int pt_set_option(pt_ctx* ctx, uint32_t option, int64_t value);

/* ---- Batches: many views of one or more scenes in one persistent wavefront run -----------
 * No counterpart in the reference, which renders one camera of one scene per PathTracer
 * (src/pathtrace/pathtrace.h:43-103). A batch takes the scenes of existing contexts (their
 * device-built BVHs are shared, not copied; the contexts must outlive the batch) and a list
 * of views; pt_batch_render traces every view's spp further samples with all views' paths
 * streaming through one path pool, each view accumulating exactly as pt_render would
 * (same sampler seeding per pixel and sample index, same running mean). */
typedef struct pt_view {
    uint32_t scene;       /* index into the batch's scene list */
    uint32_t width;
    uint32_t height;
    uint32_t spp;         /* samples per pixel added by each pt_batch_render (> 0) */
    uint32_t max_bounces; /* PathTracer::set_max_bounces */
    uint32_t reserved;
    pt_camera camera;     /* PathTracer::set_camera content */
} pt_view;

typedef struct pt_batch pt_batch;

int pt_batch_create(pt_ctx* const* scenes, uint32_t scene_count, const pt_view* views, uint32_t view_count,
                    pt_batch** out);
/* Replaces a view's camera, scene, spp or depth (not its size) and restarts its accumulation. */
int pt_batch_set_view(pt_batch* batch, uint32_t view, const pt_view* v);
/* Asynchronous on the batch's stream; like pt_render, returns once all its paths have started,
 * and the other pt_batch_* calls finish them first. */
int pt_batch_render(pt_batch* batch);
int pt_batch_read(pt_batch* batch, uint32_t view, float* rgba, size_t row_pitch);
/* pt_read_tonemapped for one view of the batch. */
int pt_batch_read_tonemapped(pt_batch* batch, uint32_t view, uint8_t* rgba8, size_t row_pitch);
/* Counters summed over all views (and BVH figures over all scenes); width/height are 0. */
int pt_batch_get_stats(pt_batch* batch, pt_stats* out);
int pt_batch_set_option(pt_batch* batch, uint32_t option, int64_t value);
int pt_batch_reset_stats(pt_batch* batch);
const char* pt_batch_last_error(const pt_batch* batch);
void pt_batch_destroy(pt_batch* batch);

/* ---- Learned radiance cache ---------------------------------------------------------------
 * No counterpart in the reference: an optional estimator on top of its path tracer. A small
 * MLP (96 encoded inputs, five hidden layers of 64 ReLU units, RGB out; BF16 matrix cores,
 * FP32 accumulation and FP32 Adam) learns, online, the radiance leaving the second (or
 * query_vertex-th) path vertex from the tracer's own training paths. With query != 0 every
 * other path ends at that vertex with the cache's prediction (a biased, lower-variance
 * estimate); with query == 0 the renderer only trains, and its images equal the plain
 * tracer's. */
#define PT_CACHE_FEATURES 96 /* encoded input width of pt_cache_forward / pt_cache_train_batch */

typedef struct pt_cache_config {
    uint32_t enable;       /* 0: free the cache, plain path tracing */
    uint32_t query;        /* != 0: end non-training paths at the 2nd vertex with the prediction */
    uint32_t train_stride; /* one path in train_stride is a training path (>= 1) */
    uint32_t train_steps;  /* optimiser steps after each frame (frame_spp > 0) or after each
                            * wavefront iteration (frame_spp == 0) */
    uint32_t batch;        /* samples per optimiser step */
    float learning_rate;   /* Adam step size */
    uint32_t seed;         /* weight initialisation (He-uniform from a Splitmix32 stream) */
    uint32_t frame_spp;    /* > 0: render jobs of at most frame_spp samples per pixel, training
                            * on each job's completed records between jobs. 0: jobs as large as
                            * without the cache, training in flight on the records completed so
                            * far, between wavefront iterations (no host synchronisation) */
    float continue_prob;   /* 0: a queried path ends with the prediction (biased). p in (0, 1]:
                            * it goes on with probability p and weight 1/p while the prediction
                            * is added with weight 1 - 1/p, a control variate (unbiased; per-sample
                            * sums may then be negative and are accumulated as they are) */
    uint32_t query_vertex; /* path vertex whose outgoing radiance the cache learns and answers
                            * (2..8; 0: the second vertex). Later vertices leave more of each
                            * path unbiased and weigh the cache's error by the throughput to them */
    float ema_decay;       /* 0: queries use the trained weights. d in (0, 1): queries use the
                            * exponential moving average e <- d e + (1 - d) w of the weights after
                            * each optimiser step, from e = 0, divided by 1 - d^steps (restarted
                            * by pt_cache_configure and pt_cache_set_params) */
    float query_max_pdf;   /* 0: every non-training path at the vertex ends with the prediction.
                            * q > 0: only those that arrived by a direction sampled with a pdf
                            * (solid angle) below q; paths from sharp glossy or specular lobes,
                            * whose radiance a blurred prediction would miss, go on unbiased */
} pt_cache_config;

typedef struct pt_cache_stats {
    uint32_t enabled;
    uint32_t query;
    uint32_t steps;         /* optimiser steps since pt_cache_configure */
    uint32_t last_records;  /* training records of the last job */
    uint64_t records_total;
    float last_loss;        /* relative L2 loss of the last optimiser step's batch */
    uint32_t param_count;
    float bounds_lo[3];     /* world box the positions are normalised to */
    float bounds_hi[3];
} pt_cache_stats;

/* (Re)initialises the cache (fresh weights, Adam state and counters) and resets
 * sample_index. Synchronises. */
int pt_cache_configure(pt_ctx* ctx, const pt_cache_config* config);
int pt_cache_get_stats(pt_ctx* ctx, pt_cache_stats* out);

/* Parameter vector: per layer W[out][in] (row-major) then b[out]; layers 96->64, 4 x 64->64,
 * 64->3. */
uint32_t pt_cache_param_count(void);
int pt_cache_get_params(pt_ctx* ctx, float* params, size_t count);
int pt_cache_set_params(pt_ctx* ctx, const float* params, size_t count);
/* The fp32 weights the renderer's queries read: the trained ones, or with ema_decay > 0 the
 * bias-corrected moving average (equal to the trained weights before the first step). */
int pt_cache_get_query_params(pt_ctx* ctx, float* params, size_t count);

/* Predictions of n explicit feature rows (n x PT_CACHE_FEATURES floats, host memory) into
 * out (n x 3). Inputs and weights are rounded to BF16 as in the renderer. */
int pt_cache_forward(pt_ctx* ctx, const float* features, uint32_t n, float* out);

/* Input encoding of n records (13 floats each: hit position, direction towards the previous
 * vertex, shading normal, roughness, base colour) into n x PT_CACHE_FEATURES floats: the
 * position normalised to the scene's world box, its sines and cosines at 12 octaves
 * (sin, cos of pi 2^b p_c for b = 0..11, c = x, y, z), then p, direction, normal, roughness,
 * base colour and zero padding, each rounded to BF16. Synchronises. */
int pt_cache_encode(pt_ctx* ctx, const float* records, uint32_t n, float* features);

/* Loss mean_{i,c} (y - t)^2 / (y^2 + 0.01) (the denominator held constant) of one batch and
 * its gradient with respect to the parameter vector, into grads / loss when non-null; with
 * apply != 0 one Adam step follows. Synchronises. */
int pt_cache_train_batch(pt_ctx* ctx, const float* features, const float* targets, uint32_t n, int apply,
                         float* grads, float* loss);

/* Device build of the light sampler's alias tables: TableLightSampler's light table and the
 * per-emissive-mesh triangle tables (src/scene/lightsample.cpp:5-178, src/aliastable.h:21-67),
 * recomputed on the device from the scene on the GPU (material emission - for an emission map
 * the Monte Carlo intensity of calculate_object_intensity, lightsample.cpp:41-87 - point and
 * directional light colours, triangle areas) and written over the words pt_create uploaded;
 * the result is bit-identical to the host packing. pt_create runs it; this call rebuilds.
 * Resets sample_index. PT_ESTATE when the words passed to pt_create do not follow that
 * packing. */
int pt_build_light_tables(pt_ctx* ctx);
/* Copies the light-sampler words (light_word_count of the scene) back to host memory. */
int pt_read_light_words(pt_ctx* ctx, uint32_t* words, size_t count);

/* Autotuner (no counterpart in the reference): times spp samples of the current view under
 * candidate launch configurations (coordinate descent over blocks per CU, LDS stack depth,
 * pool slots, BVH node width, BLAS leaf size and TLAS opening depth), keeps the fastest (device
 * time per traced ray) and clears the accumulation (sample_index 0).
 * Not with the radiance cache or instrumentation options on. Synchronises. */
typedef struct pt_tune_result {
    uint32_t blocks_per_cu;
    uint32_t pool_slots;
    uint32_t lds_stack;
    uint32_t leaf_size;
    uint32_t node_width;
    uint32_t trials;
    double ns_per_ray_start; /* the configuration the tuner started from */
    double ns_per_ray_best;
    uint32_t tlas_open;      /* PT_OPT_TLAS_OPEN */
    uint32_t reserved;
} pt_tune_result;
int pt_autotune(pt_ctx* ctx, uint32_t spp, pt_tune_result* out);

/* Waits for all work queued on the context's stream. */
int pt_synchronize(pt_ctx* ctx);

const char* pt_last_error(const pt_ctx* ctx);

void pt_destroy(pt_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif
