// Context of the device back end and the launchers of the wavefront stages.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "bvh.h"
#include "cache.h"
#include "common.h"
#include "shading.h"

namespace ptk {

constexpr uint32_t TRACE_BLOCK = 256;     // threads per traversal block (4 waves of 64)
constexpr uint32_t LDS_STACK = 16;        // stack entries per lane held in LDS
constexpr uint32_t SPILL_STACK = 112;     // further entries per lane spilled to global memory

// The path pool as two queues (ping-pong, struct-of-arrays): entry i of a queue is a ray
// together with the state of its path, so every stage reads and writes its entries in queue
// order (coalesced), and a path's state moves with it from queue to queue. Paths of a job
// (every pixel x every sample of a pt_render chunk) stream through: after each iteration the
// continuing paths are compacted to the front of the next queue and the job's next paths fill
// the rest, so every extend/shadow launch stays full until the job's last paths drain.
// The wavefront counters sit 4 KB apart: the atomics of the queue appends then go to
// different memory channels instead of queueing on one line.
constexpr uint32_t CTR_STRIDE = 1024;
constexpr uint32_t TOTALS = 16;  // words of Wavefront::totals
__host__ __device__ constexpr uint32_t CTR(uint32_t k) { return k * CTR_STRIDE; }

struct Wavefront {
    uint32_t capacity = 0;         // entries per queue
    float4* q_o[2] = {nullptr, nullptr};       // origin, tmin
    float4* q_d[2] = {nullptr, nullptr};       // direction, tmax
    // job path index, py << 16 | px of its pixel, sample index, view << 16 | segment
    uint4* q_ps[2] = {nullptr, nullptr};
    float4* q_thr[2] = {nullptr, nullptr};     // throughput rgb, brdf pdf (segments >= 1; implied for camera rays)
    float4* q_rad[2] = {nullptr, nullptr};     // radiance rgb, training record bits (~0: none; segments >= 1)
    uint32_t* q_scene[2] = {nullptr, nullptr}; // scene of each queued ray (scene batches)
    float4* hit = nullptr;         // t, u, v, instance bits (by queue position)
    uint32_t* hit_prim = nullptr;
    uint32_t* s_target = nullptr;  // shadow rays: radiance target (TARGET_ENDED | path, or next entry)
    float4* s_o = nullptr;         // origin, tmax
    float4* s_d = nullptr;         // direction
    float4* s_c = nullptr;         // unoccluded contribution
    uint32_t* s_scene = nullptr;   // scene of each shadow ray (scene batches)
    uint2* train_end = nullptr;    // (path, training record) of recorded paths that ended this iteration
    float4* rad_out = nullptr;     // radiance of every path of the job, by path id
    uint32_t rad_out_capacity = 0; // paths
    // at CTR(k): [0],[1] extension queues, [2] shadow queue, [3] error flags of the path
    // stages, [4] ended recorded paths, [5],[6] next path to start of job slot 0, 1, [7],[8]
    // ended paths of job slot 0, 1, [9] error flags of the ray queries (pt_trace_*), [10], [11],
    // [12] chunk counters of the extend, shadow and shade launches (one per XCD, 16 words apart),
    // [13] the last refill plan (6 words: continuing entries, older slot, first new path and new
    // paths per slot)
    uint32_t* counters = nullptr;
    // [0] extension rays, [1] shadow rays, [2]/[3] node/triangle fetches of extension
    // rays, [4]/[5] of shadow rays, [6]/[7] TLAS node fetches / instances entered by extension
    // rays, [8]/[9] by shadow rays (fetches counted only when fetch counting is on)
    unsigned long long* totals = nullptr;
    uint32_t* spill = nullptr;     // traversal stack spill, SPILL_STACK words per grid thread
    // material sort of the extension hits (allocated on first use)
    uint32_t* sort_keys[2] = {nullptr, nullptr};
    uint32_t* sort_vals[2] = {nullptr, nullptr};
    void* sort_temp = nullptr;
    size_t sort_temp_bytes = 0;
    uint32_t spill_threads = 0;
    uint32_t spill_depth = SPILL_STACK;  // spill entries per lane in use (PT_OPT_SPILL_DEPTH)
    uint32_t* snapshot = nullptr;     // per parity 8 words: next queue, ended[2], issued[2], error flags
    uint32_t* host_counts = nullptr;  // pinned copy of snapshot, read back with a one-iteration lag
    hipEvent_t count_ready[2] = {nullptr, nullptr};
};

// What the path stages need of the radiance cache.
struct CacheDev {
    float* records;           // training records (MLP_REC_FLOATS each)
    uint32_t* counters;       // [0] training records of the job, [1] queries of the iteration
    uint32_t record_capacity;
    float* queries;
    uint32_t* query_slot;     // radiance target of each query (TARGET_ENDED | path, or next entry)
    uint32_t train_stride;
    uint32_t collect;         // record training paths at the cache's vertex
    uint32_t segment;         // segment arriving at that vertex (query_vertex - 1)
    uint32_t query;           // end other paths there with the cache's prediction
    float cont_prob;          // > 0: queried paths go on with this probability (control variate)
    float max_pdf;            // > 0: only paths that arrived by a direction of pdf below it are queried
};

// One view of a render: its camera, image and accumulation buffer (device table).
struct ViewDev {
    CameraBlock cam;
    float4* accum;
    uint32_t width, height, npix;
    uint32_t scene;        // index into the scene table
    uint32_t max_bounces;
    uint32_t pad;
};

// A job is a run of the path pool over consecutive segments; a segment is a range of
// samples of one view, its paths numbered pixel-major from path_offset (a pixel's samples
// consecutive, pixels in the 8x8-pixel tile order of trace.hip's tile_pixel).
constexpr uint32_t MAX_SEGS = 1024;
struct SegDev {
    uint32_t view;
    uint32_t first_sample;  // sample index of the segment's first sample
    uint32_t samples;
    uint32_t path_offset;   // job path index of (pixel 0 of the tile order, first_sample)
    uint32_t pix_offset;    // sum of the pixel counts of the earlier segments
    uint32_t pad[3];
};
struct JobDev {
    uint32_t paths;            // paths of the job (0: the slot holds no job)
    uint32_t nseg;
    uint32_t signed_radiance;  // keep negative per-sample sums (control-variate estimates)
    uint32_t pix_total;        // sum of the segments' pixel counts
    uint32_t seq;              // order of the jobs in the two slots: the older one issues first
    uint32_t pad[3];
    SegDev seg[MAX_SEGS];
};

// Host side of a wavefront's path stream, kept between renders: a job is loaded into one of
// two slots (device tables d_job[0], d_job[1]); its paths are issued into the pool, the older
// slot's first, and once every path of a job has ended its radiance is accumulated and the
// slot is free again. The next job's paths thus fill the pool while the previous job's last
// paths drain, within a render and from one render call to the next.
struct StreamState {
    int q = 0;               // parity of the queue the next iteration traces
    uint64_t launched = 0;   // iterations queued so far
    uint32_t seq = 0;
    bool overflow = false;   // a traversal stack overflowed (seen in a snapshot)
    // the shadow queue holds rays of the last iteration queued, to be traced by the next one
    // (a pipelined iteration's; a job's paths end with their last shade stage, and its
    // accumulation is queued only after the iteration that traces those rays)
    bool shadow_pending = false;
    uint64_t flags_from = 0; // the first iteration queued after the error flags were last cleared
    struct Slot {
        bool busy = false;
        uint32_t paths = 0, seq = 0;
        uint64_t first = 0;   // the first iteration queued after the job was loaded
        uint32_t issued = 0;  // paths started, as last read back
        uint32_t ended = 0;   // paths ended, as last read back
    } slot[2];
    bool idle() const { return !slot[0].busy && !slot[1].busy; }
};

// How an iteration traces its shadow rays (launch_iteration).
struct IterMode {
    bool pipelined = false;  // with the next iteration's extension rays
    bool consume = false;    // this iteration traces the previous one's shadow rays with its own
                             // and starts the new paths it planned
    bool flush = false;      // a non-pipelined iteration after a pipelined one: those first
};

// Executable graphs of one wavefront iteration (queue parity 0 and 1), captured on first use
// and kept while the launch configuration they were captured with stays the same.
struct IterGraphs {
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    uint64_t sig[2] = {0, 0};
    void destroy()
    {
        for (int k = 0; k < 2; k++)
            if (exec[k])
                (void)hipGraphExecDestroy(exec[k]);
        exec[0] = exec[1] = nullptr;
        sig[0] = sig[1] = 0;
    }
};

enum KernelId { KERNEL_EXTEND = 0, KERNEL_SHADE = 1, KERNEL_SHADOW = 2, KERNEL_CACHE = 3, KERNEL_REFILL = 4, KERNEL_COUNT = 5 };

// Optional per-launch timing of the main stages with events on the context's stream.
struct KernelTimer {
    bool enabled = false;
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<int, size_t>> marks;  // (kernel, event index of the start mark)
    size_t used = 0;

    hipError_t mark(hipStream_t st, int kernel, bool start)
    {
        if (!enabled)
            return hipSuccess;
        if (used == pool.size()) {
            hipEvent_t ev;
            hipError_t e = hipEventCreate(&ev);
            if (e != hipSuccess)
                return e;
            pool.push_back(ev);
        }
        if (start)
            marks.push_back({kernel, used});
        return hipEventRecord(pool[used++], st);
    }
    void reset()
    {
        used = 0;
        marks.clear();
    }
    // Sums the durations per kernel (the stream must be idle).
    hipError_t collect(double* ms, uint32_t* launches)
    {
        for (int k = 0; k < KERNEL_COUNT; k++) {
            ms[k] = 0.0;
            launches[k] = 0;
        }
        for (const auto& m : marks) {
            float t = 0.0f;
            hipError_t e = hipEventElapsedTime(&t, pool[m.second], pool[m.second + 1]);
            if (e != hipSuccess)
                return e;
            ms[m.first] += t;
            launches[m.first]++;
        }
        return hipSuccess;
    }
    void destroy()
    {
        for (hipEvent_t ev : pool)
            (void)hipEventDestroy(ev);
        pool.clear();
        reset();
    }
};

// An emissive mesh's triangle alias table inside the light-sampler words.
struct EmitterTable {
    uint32_t mesh, word_offset, tris;
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    std::string error;

    // scene
    float4* d_vertices = nullptr;
    uint32_t* d_indices = nullptr;
    InstanceData* d_instances = nullptr;
    MaterialRec* d_materials = nullptr;
    uint32_t* d_light = nullptr;
    pt_mesh_desc* d_meshes = nullptr;
    uint32_t* d_instance_mesh = nullptr;
    float* d_inst_opacity = nullptr;
    pt_texture_desc* d_tex = nullptr;
    uint8_t* d_texels = nullptr;
    BvhBuffers bvh;
    DeviceScene scene{};
    double bvh_build_ms = 0.0;
    uint32_t instance_count = 0;

    // image and camera
    uint32_t width = 0, height = 0;
    float4* accum = nullptr;
    CameraBlock cam{};
    bool have_camera = false;
    uint32_t max_bounces = 2;
    uint32_t sample_index = 0;

    Wavefront wf;
    uint32_t grid_blocks = 0;  // persistent grid of the traversal kernels
    bool count_fetches = false;
    KernelTimer timer;
    bool stats_fresh = true;  // counters restart at the next pt_render
    CacheState cache;
    ViewDev* d_views = nullptr;  // the context's one view
    ViewDev* h_view = nullptr;   // pinned staging of it
    JobDev* d_job = nullptr;     // the two job slots
    JobDev* h_job = nullptr;     // pinned staging of the job slots
    StreamState stream_state;
    IterGraphs graphs;
    bool use_graphs = true;
    // launch configuration (PT_OPT_BLOCKS_PER_CU / POOL_SLOTS / LDS_STACK / LEAF_SIZE)
    uint32_t cus = 1;
    uint32_t blocks_per_cu = 6;
    uint32_t pool_slots = 0;
    uint32_t lds_stack = LDS_STACK;
    uint32_t spill_depth = SPILL_STACK;  // PT_OPT_SPILL_DEPTH
    bool transparent = false;  // some instance's material is transparent
    bool single_tri_leaves = false;  // every BLAS leaf of the built hierarchy holds one triangle
    uint32_t leaf_size = 1;  // single-triangle leaves: the traversal kernels' leaf-loop-free variants
    uint32_t node_width = 4;
    uint32_t bvh_builder = 0;  // BVH_BUILDER_*
    uint32_t tlas_open = 0;    // PT_OPT_TLAS_OPEN
    uint32_t tlas_boxes = 3;   // PT_OPT_TLAS_BOXES
    uint32_t tlas_builder = BVH_TLAS_HOST;  // PT_OPT_TLAS_BUILDER
    bool sort_hits = false;
    uint32_t job_paths = 0;  // capacity of the per-job radiance buffer
    std::vector<pt_mesh_desc> meshes;  // host copy (BVH rebuilds)
    uint32_t light_count = 0;
    uint64_t light_word_count = 0;
    std::vector<EmitterTable> emitter_tables;  // in word order
    bool light_layout_known = false;           // the words follow the packing of the host build

    // stats of the last render
    unsigned long long last_ext = 0, last_shadow = 0, last_paths = 0;
    double last_ms = 0.0;
};

hipError_t wavefront_alloc(Wavefront& wf, uint32_t capacity, uint32_t rad_out_paths, uint32_t spill_threads);
void wavefront_free(Wavefront& wf);

// What a render needs of its owner (a context or a batch).
struct RenderIO {
    hipStream_t st;
    Wavefront* wf;
    uint32_t grid_blocks;
    bool count_fetches;
    KernelTimer* timer;
    DeviceScene scene;              // the scene when there is one
    const DeviceScene* d_scenes;    // scene table when there are several, else nullptr
    const ViewDev* d_views;
    JobDev* d_job;                  // the two job slots (device)
    JobDev* h_job;                  // their pinned staging
    StreamState* stream;            // the path stream of this wavefront
    IterGraphs* graphs;             // nullptr: launch every kernel directly
    CacheState* cache;              // nullptr: no radiance cache
    uint32_t lds_stack;             // LDS stack entries per lane of the traversal kernels
    bool sort_hits;                 // shade the extension hits in (scene, material) order
    bool alpha_shadows = true;      // some instance is transparent: shadow rays weigh alpha
    bool leaf1 = false;             // every BLAS leaf holds one triangle (no leaf loop)
};

// Samples [sample_index, sample_index + spp) of one view.
struct ViewWork {
    uint32_t view, npix, sample_index, spp;
};

// Renders the given samples of every view, all views' paths streaming through one pool.
// Returns once every path of them has been started (the last ones may still be in flight,
// overlapping the next render), or, with drain, once every started path has ended and been
// accumulated.
hipError_t render_views(const RenderIO& io, const std::vector<ViewWork>& work, bool drain);

// Renders samples [ctx.sample_index, ctx.sample_index + spp) into ctx.accum (see
// render_views for drain).
hipError_t render_samples(Ctx& ctx, uint32_t spp, bool drain);

// Finishes every path in flight of the context's stream (a no-op when it is idle).
hipError_t drain_stream(Ctx& ctx);

// Primary hits of one sample index, into a device array of width*height pt_hit.
hipError_t primary_hits(Ctx& ctx, uint32_t sample_index, pt_hit* d_hits);

// Arbitrary rays (device arrays).
hipError_t trace_rays(Ctx& ctx, const pt_ray* d_rays, uint32_t count, pt_hit* d_hits, bool any_hit);

hipError_t upload_sobol_table();

// Rebuilds the light sampler's alias tables on the device (lights.hip).
hipError_t build_light_tables(Ctx& c);

// Radiance cache (cache.hip)
int mlp_param_count();
void cache_free(CacheState& c);
hipError_t cache_init(CacheState& c, uint32_t pool_slots, uint32_t record_capacity, hipStream_t st);
hipError_t cache_set_params(CacheState& c, const float* host, hipStream_t st);
hipError_t cache_get_params(CacheState& c, float* host, hipStream_t st);
hipError_t cache_get_query_params(CacheState& c, float* host, hipStream_t st);
hipError_t cache_forward(CacheState& c, const float* d_features, uint32_t n, float* d_out, hipStream_t st);
hipError_t cache_encode(CacheState& c, const float* d_rec13, uint32_t n, float* d_out, hipStream_t st);
// ACES tone mapping + 8-bit conversion of the PNG output path (post.hip): width x height
// RGBA8 words (R in the low byte) into d_out.
hipError_t tonemap_rgba8(const float4* accum, uint32_t width, uint32_t height, uint32_t* d_out, hipStream_t st);

hipError_t cache_query(CacheState& c, const uint32_t* d_count, float4* next_rad, float4* rad_out, unsigned blocks,
                       hipStream_t st);
hipError_t cache_backward(CacheState& c, const float* d_features, const float* d_targets, uint32_t n, uint32_t n_rec,
                          uint32_t start, uint32_t stride, hipStream_t st);
hipError_t cache_adam(CacheState& c, hipStream_t st);
hipError_t cache_train_records(CacheState& c, uint32_t n_records, hipStream_t st);
hipError_t cache_train_inflight(CacheState& c, hipStream_t st);

}  // namespace ptk
