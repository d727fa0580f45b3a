// Context of the device back end and the launchers of the wavefront stages.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "bvh.h"
#include "common.h"
#include "shading.h"

namespace ptk {

constexpr uint32_t TRACE_BLOCK = 256;     // threads per traversal block (4 waves of 64)
constexpr uint32_t LDS_STACK = 16;        // stack entries per lane held in LDS
constexpr uint32_t SPILL_STACK = 112;     // further entries per lane spilled to global memory

// Path state and ray queues of one batch of paths (struct-of-arrays, one slot per path).
struct Wavefront {
    uint32_t capacity = 0;         // paths
    float4* thr_pdf = nullptr;     // throughput rgb, brdf pdf
    float4* rad = nullptr;         // radiance rgb
    uint32_t* q_path[2] = {nullptr, nullptr};
    float4* q_o[2] = {nullptr, nullptr};  // origin, tmin
    float4* q_d[2] = {nullptr, nullptr};  // direction, tmax
    float4* hit = nullptr;         // t, u, v, instance bits
    uint32_t* hit_prim = nullptr;
    uint32_t* s_path = nullptr;    // shadow rays
    float4* s_o = nullptr;         // origin, tmax
    float4* s_d = nullptr;         // direction
    float4* s_c = nullptr;         // unoccluded contribution
    uint32_t* counters = nullptr;  // [0],[1] extension queues, [2] shadow queue, [3] error flags
    unsigned long long* totals = nullptr;  // [0] extension rays, [1] shadow rays
    uint32_t* spill = nullptr;     // traversal stack spill, SPILL_STACK words per grid thread
    uint32_t spill_threads = 0;
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

    // stats of the last render
    unsigned long long last_ext = 0, last_shadow = 0, last_paths = 0;
    double last_ms = 0.0;
};

hipError_t wavefront_alloc(Wavefront& wf, uint32_t capacity, uint32_t spill_threads);
void wavefront_free(Wavefront& wf);

// Renders samples [ctx.sample_index, ctx.sample_index + spp) into ctx.accum.
hipError_t render_samples(Ctx& ctx, uint32_t spp);

// Primary hits of one sample index, into a device array of width*height pt_hit.
hipError_t primary_hits(Ctx& ctx, uint32_t sample_index, pt_hit* d_hits);

// Arbitrary rays (device arrays).
hipError_t trace_rays(Ctx& ctx, const pt_ray* d_rays, uint32_t count, pt_hit* d_hits, bool any_hit);

hipError_t upload_sobol_table();

}  // namespace ptk
