// C ABI of the device back end (include/pthip.h).
#include <hip/hip_runtime.h>

#include <algorithm>

#include <cmath>
#include <cstring>
#include <exception>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "device.h"
#include "pthip.h"

using namespace ptk;

// Staging area of pt_trace_rays_noalloc: the rays of a call, then their hits, in device memory
// that exists from load time on, so that call allocates nothing. One copy per GPU (each device
// holds its own array until the process ends); calls that overlap on one GPU race for it.
static __device__ __align__(256) unsigned char scratch[size_t(PT_NOALLOC_MAX_RAYS) * (sizeof(pt_ray) + sizeof(pt_hit))];

namespace {

// The array's device address on each device: hipGetSymbolAddress, called once per device with
// that device current (by pt_create); the host-side name `scratch` is not a device address.
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[MAX_DEVICES] = {};

hipError_t scratch_setup(int device)
{
    if (device < 0 || device >= MAX_DEVICES)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device])
        return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess)
        g_scratch_ptr[device] = p;
    return e;
}

void* scratch_ptr(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    return (device >= 0 && device < MAX_DEVICES) ? g_scratch_ptr[device] : nullptr;
}

}  // namespace

struct pt_ctx {
    Ctx c;
};

namespace {

thread_local std::string g_create_error;

int fail(Ctx* c, int code, const std::string& msg)
{
    if (c)
        c->error = msg;
    else
        g_create_error = msg;
    return code;
}

int hip_fail(Ctx* c, hipError_t e, const char* what)
{
    const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? PT_ENOMEM : PT_EDEVICE;
    return fail(c, code, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(c, expr, what)                \
    do {                                      \
        hipError_t e_ = (expr);               \
        if (e_ != hipSuccess)                 \
            return hip_fail((c), e_, (what)); \
    } while (0)

template <typename T>
hipError_t upload(T*& dst, const void* src, size_t bytes, hipStream_t st)
{
    hipError_t e = hipMalloc((void**)&dst, std::max<size_t>(bytes, 16));
    if (e != hipSuccess)
        return e;
    if (bytes)
        e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
    return e;
}

void release(Ctx& c)
{
    void* ptrs[] = {c.d_vertices, c.d_indices, c.d_instances, c.d_materials, c.d_light,
                    c.d_meshes, c.d_instance_mesh, c.d_inst_opacity, c.accum, c.d_tex, c.d_texels};
    for (void* p : ptrs)
        if (p)
            (void)hipFree(p);
    free_bvh(c.bvh);
    wavefront_free(c.wf);
    cache_free(c.cache);
    c.graphs.destroy();
    if (c.d_views)
        (void)hipFree(c.d_views);
    if (c.d_job)
        (void)hipFree(c.d_job);
    if (c.h_view)
        (void)hipHostFree(c.h_view);
    if (c.h_job)
        (void)hipHostFree(c.h_job);
    c.timer.destroy();
    if (c.ev_start)
        (void)hipEventDestroy(c.ev_start);
    if (c.ev_stop)
        (void)hipEventDestroy(c.ev_stop);
    if (c.stream)
        (void)hipStreamDestroy(c.stream);
}

int validate_blob(const pt_scene_blob* b, std::string& why)
{
    if (!b)
        return (why = "scene is null", PT_EINVAL);
    if ((b->vertex_count && !b->vertices) || (b->index_count && !b->indices) || (b->instance_count && !b->instances) ||
        (b->material_count && !b->materials) || (b->light_word_count && !b->light_words) ||
        (b->mesh_count && !b->meshes) || (b->instance_count && !b->instance_mesh))
        return (why = "scene array pointer is null", PT_EINVAL);
    if (b->vertex_count > 0x7FFFFFFFull || b->index_count > 0xFFFFFFFFull)
        return (why = "scene exceeds 32-bit addressing", PT_EINVAL);
    for (uint32_t m = 0; m < b->mesh_count; m++) {
        const pt_mesh_desc& d = b->meshes[m];
        if (uint64_t(d.index_start) + uint64_t(d.tri_count) * 3 > b->index_count ||
            uint64_t(d.vertex_start) + d.vertex_count > b->vertex_count)
            return (why = "mesh " + std::to_string(m) + " exceeds the vertex/index arrays", PT_EINVAL);
        for (uint64_t k = 0; k < uint64_t(d.tri_count) * 3; k++)
            if (b->indices[d.index_start + k] >= d.vertex_count)
                return (why = "mesh " + std::to_string(m) + " has an index out of range", PT_EINVAL);
    }
    const auto* inst = static_cast<const InstanceData*>(b->instances);
    for (uint32_t i = 0; i < b->instance_count; i++) {
        if (b->instance_mesh[i] >= b->mesh_count)
            return (why = "instance " + std::to_string(i) + " names a missing mesh", PT_EINVAL);
        if (inst[i].material_index >= b->material_count)
            return (why = "instance " + std::to_string(i) + " names a missing material", PT_EINVAL);
        const pt_mesh_desc& d = b->meshes[b->instance_mesh[i]];
        if (inst[i].vertex_index != d.vertex_start || inst[i].index_index != d.index_start)
            return (why = "instance " + std::to_string(i) + " does not point at its mesh", PT_EINVAL);
    }
    if (b->light_count) {
        if (uint64_t(b->light_count) * PT_LIGHT_BIN_WORDS > b->light_word_count)
            return (why = "light table exceeds the light words", PT_EINVAL);
        for (uint32_t l = 0; l < b->light_count; l++) {
            const uint32_t* bin = b->light_words + size_t(l) * PT_LIGHT_BIN_WORDS;
            if (bin[2] >= b->light_count)
                return (why = "light alias out of range", PT_EINVAL);
            if (bin[3] == 3u) {
                const uint64_t off = bin[4], tris = bin[5], i0 = bin[6], ni = bin[7];
                if (tris == 0 || off + tris * PT_EMITTER_BIN_WORDS > b->light_word_count || ni == 0 ||
                    i0 + ni > b->instance_count)
                    return (why = "emitter light bin out of range", PT_EINVAL);
                uint32_t min_verts = 0xFFFFFFFFu;
                for (uint64_t i = i0; i < i0 + ni; i++)
                    min_verts = std::min(min_verts, b->meshes[b->instance_mesh[i]].vertex_count);
                for (uint64_t t = 0; t < tris; t++) {
                    const uint32_t* eb = b->light_words + off + t * PT_EMITTER_BIN_WORDS;
                    if (eb[2] >= tris || eb[3] >= min_verts || eb[4] >= min_verts || eb[5] >= min_verts)
                        return (why = "emitter triangle bin out of range", PT_EINVAL);
                }
            }
        }
    }
    for (uint32_t i = 0; i < b->instance_count; i++)
        if (inst[i].emitter_index != PT_NO_EMITTER && uint64_t(inst[i].emitter_index) >= b->light_count)
            return (why = "instance " + std::to_string(i) + " has an emitter index beyond the light table", PT_EINVAL);
    // textures: descriptors inside the texel bytes, 4-byte aligned RGBA rows
    if ((b->texture_count && !b->textures) || (b->texel_bytes && !b->texels))
        return (why = "texture array pointer is null", PT_EINVAL);
    for (uint32_t t = 0; t < b->texture_count; t++) {
        const pt_texture_desc& d = b->textures[t];
        const uint64_t bytes = uint64_t(d.width) * d.height * d.channels;
        if (d.width == 0 || d.height == 0 || (d.channels != 1 && d.channels != 2 && d.channels != 4) ||
            d.offset % 4 != 0 || d.offset + bytes > b->texel_bytes || (d.srgb && d.channels != 4))
            return (why = "texture " + std::to_string(t) + " is malformed", PT_EINVAL);
    }
    // every map a material names exists
    const auto* mats = static_cast<const MaterialRec*>(b->materials);
    for (uint32_t m = 0; m < b->material_count; m++) {
        const uint32_t bits = float_as_u32(mats[m].rmnb.w);
        const float* slots[] = {&mats[m].base_colour.x, &mats[m].emission.x, &mats[m].rmnb.x, &mats[m].rmnb.y,
                                &mats[m].rmnb.x, &mats[m].specular.x, &mats[m].sheen.x, &mats[m].rmnb.z};
        const uint32_t flags[] = {MAP_BASE_COLOUR, MAP_EMISSION, MAP_ROUGHNESS, MAP_METALNESS, MAP_ROUGHNESS_METALNESS,
                                  MAP_SPECULAR, MAP_SHEEN, MAP_NORMAL};
        for (int k = 0; k < 8; k++)
            if ((bits & flags[k]) && float_as_u32(*slots[k]) >= b->texture_count)
                return (why = "material " + std::to_string(m) + " names a missing texture", PT_EINVAL);
    }
    return PT_OK;
}

// Host check of the built hierarchy before any traversal kernel runs: every child index in
// range, every triangle record referenced exactly once, every instance reachable once.
int validate_bvh(Ctx& c, const std::vector<pt_mesh_desc>& meshes, std::string& why)
{
    const BvhBuffers& bv = c.bvh;
    // traversal addresses nodes by 32-bit byte offsets
    if (uint64_t(bv.node_count) * sizeof(Node4) > 0xFFFFFFFFull)
        return (why = "scene too large: " + std::to_string(bv.node_count) + " BVH nodes (at most 33554431)",
                PT_EINVAL);
    if (uint64_t(bv.tri_count) * sizeof(TriRec) > 0xFFFFFFFFull)
        return (why = "scene too large: " + std::to_string(bv.tri_count) + " triangles (at most 89478485)",
                PT_EINVAL);
    std::vector<Node4> nodes(bv.node_count);
    if (bv.node_count &&
        hipMemcpy(nodes.data(), bv.nodes, size_t(bv.node_count) * sizeof(Node4), hipMemcpyDeviceToHost) != hipSuccess)
        return (why = "BVH readback failed", PT_EDEVICE);
    std::vector<TriRec> tris(bv.tri_count);
    if (bv.tri_count &&
        hipMemcpy(tris.data(), bv.tris, size_t(bv.tri_count) * sizeof(TriRec), hipMemcpyDeviceToHost) != hipSuccess)
        return (why = "BVH readback failed", PT_EDEVICE);
    std::vector<uint8_t> tri_seen(bv.tri_count, 0), prim_seen(bv.tri_count, 0);
    std::vector<uint32_t> stack;
    uint32_t max_leaf = 0;
    uint32_t tri_base = 0;
    for (uint32_t m = 0; m < meshes.size(); m++) {
        const uint32_t nt = meshes[m].tri_count;
        if (nt) {
            stack.assign(1, m);
            while (!stack.empty()) {
                const uint32_t n = stack.back();
                stack.pop_back();
                if (n >= bv.node_count)
                    return (why = "BLAS child index out of range", PT_EDEVICE);
                for (int k = 0; k < 4; k++) {
                    const uint32_t ch = nodes[n].child[k];
                    if (ch == CHILD_EMPTY)
                        continue;
                    if (ch & CHILD_LEAF) {
                        const uint32_t f = ch & LEAF_FIRST_MASK, cnt = ((ch >> 27) & 0xFu) + 1u;
                        max_leaf = std::max(max_leaf, cnt);
                        if (f < tri_base || f + cnt > tri_base + nt)
                            return (why = "BLAS leaf outside its mesh", PT_EDEVICE);
                        for (uint32_t q = f; q < f + cnt; q++) {
                            if (tri_seen[q]++)
                                return (why = "triangle referenced twice", PT_EDEVICE);
                            const uint32_t prim = float_as_u32(tris[q].v0.w);
                            if (prim >= nt)
                                return (why = "triangle record with a bad primitive index", PT_EDEVICE);
                            if (prim_seen[tri_base + prim]++)
                                return (why = "primitive recorded twice", PT_EDEVICE);
                        }
                    } else {
                        if (ch >= bv.node_count || ch < meshes.size())
                            return (why = "BLAS child index out of range", PT_EDEVICE);
                        stack.push_back(ch);
                    }
                }
                if (stack.size() > bv.node_count)
                    return (why = "BLAS has a cycle", PT_EDEVICE);
            }
        }
        tri_base += nt;
    }
    for (uint32_t q = 0; q < bv.tri_count; q++)
        if (!tri_seen[q])
            return (why = "triangle record not reachable", PT_EDEVICE);
    c.single_tri_leaves = max_leaf <= 1;
    if (c.instance_count) {
        if (bv.tlas_root >= bv.node_count)
            return (why = "TLAS root out of range", PT_EDEVICE);
        std::vector<uint2> entries(bv.entry_count);
        if (bv.entry_count && hipMemcpy(entries.data(), bv.tlas_entries, size_t(bv.entry_count) * sizeof(uint2),
                                        hipMemcpyDeviceToHost) != hipSuccess)
            return (why = "BVH readback failed", PT_EDEVICE);
        std::vector<uint8_t> entry_seen(bv.entry_count, 0);
        stack.assign(1, bv.tlas_root);
        size_t visited = 0;
        while (!stack.empty()) {
            const uint32_t n = stack.back();
            stack.pop_back();
            if (++visited > bv.node_count)
                return (why = "TLAS has a cycle", PT_EDEVICE);
            for (int k = 0; k < 4; k++) {
                const uint32_t ch = nodes[n].child[k];
                if (ch == CHILD_EMPTY)
                    continue;
                if (ch & CHILD_LEAF) {
                    const uint32_t id = ch & ~CHILD_LEAF;
                    if (id >= bv.entry_count || entry_seen[id]++)
                        return (why = "TLAS leaf out of range or repeated", PT_EDEVICE);
                } else {
                    if (ch >= bv.node_count || ch < bv.tlas_root)
                        return (why = "TLAS child index out of range", PT_EDEVICE);
                    stack.push_back(ch);
                }
            }
        }
        // every entry reached; each instance's entries are an exact cut of its BLAS: walking down
        // from the root, every path meets one of them before any triangle leaf, and each is met once
        std::vector<std::vector<uint32_t>> cut(c.instance_count);
        for (uint32_t e = 0; e < bv.entry_count; e++) {
            if (!entry_seen[e])
                return (why = "TLAS entry not reachable", PT_EDEVICE);
            if (entries[e].x >= c.instance_count)
                return (why = "TLAS entry names a missing instance", PT_EDEVICE);
            cut[entries[e].x].push_back(entries[e].y);
        }
        std::vector<uint32_t> inst_mesh(c.instance_count);
        if (hipMemcpy(inst_mesh.data(), c.d_instance_mesh, size_t(c.instance_count) * 4, hipMemcpyDeviceToHost) !=
            hipSuccess)
            return (why = "BVH readback failed", PT_EDEVICE);
        for (uint32_t i = 0; i < c.instance_count; i++) {
            std::vector<uint32_t>& cu = cut[i];
            std::sort(cu.begin(), cu.end());
            if (cu.empty() || std::adjacent_find(cu.begin(), cu.end()) != cu.end())
                return (why = "instance " + std::to_string(i) + " has no TLAS entry or a repeated one", PT_EDEVICE);
            const uint32_t m = inst_mesh[i];
            if (meshes[m].tri_count == 0) {
                if (cu.size() != 1 || cu[0] != CHILD_EMPTY)
                    return (why = "instance of an empty mesh enters a node", PT_EDEVICE);
                continue;
            }
            size_t met = 0;
            stack.assign(1, m);
            while (!stack.empty()) {
                const uint32_t n = stack.back();
                stack.pop_back();
                if (std::binary_search(cu.begin(), cu.end(), n)) {
                    met++;
                    continue;
                }
                if (n >= bv.node_count)
                    return (why = "TLAS entry cut out of range", PT_EDEVICE);
                for (int k = 0; k < 4; k++) {
                    const uint32_t ch = nodes[n].child[k];
                    if (ch == CHILD_EMPTY)
                        continue;
                    if (ch & CHILD_LEAF)
                        return (why = "instance " + std::to_string(i) + ": triangles outside its TLAS entries",
                                PT_EDEVICE);
                    stack.push_back(ch);
                }
                if (stack.size() > 64)
                    return (why = "TLAS entries of instance " + std::to_string(i) + " are not a cut of its BLAS",
                            PT_EDEVICE);
            }
            if (met != cu.size())
                return (why = "TLAS entries of instance " + std::to_string(i) + " are not a cut of its BLAS",
                        PT_EDEVICE);
        }
        // world bounds (union of the root's children) normalise the radiance cache's positions
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        const Node4& root = nodes[bv.tlas_root];
        for (int k = 0; k < 4; k++) {
            if (root.child[k] == CHILD_EMPTY)
                continue;
            lo[0] = std::min(lo[0], root.lo_x[k]), hi[0] = std::max(hi[0], root.hi_x[k]);
            lo[1] = std::min(lo[1], root.lo_y[k]), hi[1] = std::max(hi[1], root.hi_y[k]);
            lo[2] = std::min(lo[2], root.lo_z[k]), hi[2] = std::max(hi[2], root.hi_z[k]);
        }
        for (int k = 0; k < 3; k++)
            if (std::isfinite(lo[k]) && std::isfinite(hi[k]) && hi[k] >= lo[k]) {
                c.cache.bounds_lo[k] = lo[k];
                c.cache.bounds_hi[k] = hi[k];
            }
    }
    return PT_OK;
}

// A ray query (pt_trace_*) whose traversal stack overflowed: its hits are not valid.
int query_overflow(Ctx& c)
{
    uint32_t flags = 0;
    HIP_TRY(&c, hipMemcpy(&flags, c.wf.counters + CTR(9), 4, hipMemcpyDeviceToHost), "error flags");
    if (flags)
        return fail(&c, PT_EDEVICE,
                    "traversal stack overflow: some rays missed BVH nodes (the hierarchy is deeper than the traversal "
                    "stack of " + std::to_string(c.lds_stack) + " + " + std::to_string(c.spill_depth) + " entries)");
    return PT_OK;
}

int need_cache(Ctx& c)
{
    if (!c.cache.configured)
        return fail(&c, PT_ESTATE, "radiance cache not configured");
    return PT_OK;
}

// Device copies of test inputs; freed by the destructor.
struct DevBufs {
    std::vector<void*> p;
    ~DevBufs()
    {
        for (void* q : p)
            (void)hipFree(q);
    }
    template <typename T>
    hipError_t alloc(T*& dst, size_t bytes)
    {
        hipError_t e = hipMalloc((void**)&dst, std::max<size_t>(bytes, 16));
        if (e == hipSuccess)
            p.push_back(dst);
        return e;
    }
};


// The emitter triangle tables follow the light table in the order their meshes first appear
// among the emitting lights (lightsample.cpp:89-178 as the host packs it). False when the words
// do not have that layout.
bool parse_emitter_tables(const pt_scene_blob& b, std::vector<EmitterTable>& out)
{
    out.clear();
    if (b.light_word_count < uint64_t(b.light_count) * PT_LIGHT_BIN_WORDS)
        return false;
    uint64_t ofs = uint64_t(b.light_count) * PT_LIGHT_BIN_WORDS;
    for (uint32_t i = 0; i < b.light_count; i++) {
        const uint32_t* w = b.light_words + size_t(i) * PT_LIGHT_BIN_WORDS;
        if (w[3] != 3u)  // not an emitting object
            continue;
        const uint32_t inst = w[6];
        if (inst >= b.instance_count)
            return false;
        const uint32_t mesh = b.instance_mesh[inst];
        if (mesh >= b.mesh_count || w[5] != b.meshes[mesh].tri_count)
            return false;
        bool seen = false;
        for (const EmitterTable& t : out)
            seen |= t.mesh == mesh;
        if (!seen) {
            out.push_back(EmitterTable{mesh, uint32_t(ofs), b.meshes[mesh].tri_count});
            ofs += uint64_t(b.meshes[mesh].tri_count) * PT_EMITTER_BIN_WORDS;
        }
    }
    return ofs == b.light_word_count;
}

// (Re)builds the BLAS+TLAS with the context's leaf size and checks the hierarchy.
int build_scene_bvh(Ctx& c)
{
    hipStream_t st = c.stream;
    free_bvh(c.bvh);
    HIP_TRY(&c, hipEventRecord(c.ev_start, st), "event");
    BvhBuildParams bp;
    bp.blas_leaf_size = c.leaf_size;
    bp.node_width = c.node_width;
    bp.builder = c.bvh_builder;
    bp.tlas_open = c.tlas_open;
    if (const char* env = getenv("PT_TLAS_OPEN_RATIO"))  // tuning of the opening criterion
        bp.tlas_open_ratio = float(std::max(0.0, std::min(16.0, atof(env))));
    bp.tlas_box_levels = c.tlas_boxes;
    bp.tlas_builder = c.tlas_builder;
    if (const char* env = getenv("PT_BVH_SAH_TRI_COST"))  // tuning of the SAH cost model
        bp.sah_tri_cost = float(std::max(0.01, std::min(100.0, atof(env))));
    HIP_TRY(&c,
            build_bvh(st, c.d_vertices, c.d_indices, c.d_meshes, c.meshes, c.d_instances, c.d_instance_mesh,
                      c.instance_count, c.d_materials, bp, c.bvh),
            "BVH build");
    HIP_TRY(&c, hipEventRecord(c.ev_stop, st), "event");
    HIP_TRY(&c, hipEventSynchronize(c.ev_stop), "event");
    float ms = 0.0f;
    HIP_TRY(&c, hipEventElapsedTime(&ms, c.ev_start, c.ev_stop), "event");
    c.bvh_build_ms = ms;
    std::string why;
    if (int v = validate_bvh(c, c.meshes, why))
        return fail(&c, v, "BVH validation: " + why);
    c.scene.nodes = c.bvh.nodes;
    c.scene.tris = c.bvh.tris;
    c.scene.inst_inv = c.bvh.inst_inv;
    c.scene.tlas_entries = c.bvh.tlas_entries;
    c.scene.tlas_root = c.bvh.tlas_root;
    c.graphs.destroy();
    return PT_OK;
}

// Grid and pool for the context's launch configuration (buffers reallocated).
hipError_t apply_launch(Ctx& c)
{
    c.grid_blocks = ((c.cus * c.blocks_per_cu + 7) / 8) * 8;
    c.graphs.destroy();
    hipError_t e = wavefront_alloc(c.wf, c.pool_slots, c.job_paths, c.grid_blocks * TRACE_BLOCK);
    c.wf.spill_depth = c.spill_depth;
    return e;
}

// The error flags of the path stages (traversal stack overflow): a render whose paths lost
// BVH nodes is reported until the accumulation restarts (a new camera, depth, cache or light
// tables), as the reference reports its failures by exception (src/pathtrace/ptdispatch.cpp:44-47).
bool overflowed(Ctx& c, std::string& why)
{
    uint32_t flags = 0;
    if (!c.stream_state.overflow &&
        hipMemcpy(&flags, c.wf.counters + CTR(3), 4, hipMemcpyDeviceToHost) == hipSuccess && flags)
        c.stream_state.overflow = true;
    if (c.stream_state.overflow)
        why = "traversal stack overflow: the render missed BVH nodes (the hierarchy is deeper than the "
              "traversal stack of " + std::to_string(c.lds_stack) + " + " + std::to_string(c.spill_depth) +
              " entries); the image is invalid until the accumulation restarts";
    return c.stream_state.overflow;
}

// Clears the error flags when the accumulation restarts.
hipError_t clear_overflow(Ctx& c)
{
    c.stream_state.overflow = false;
    c.stream_state.flags_from = c.stream_state.launched;
    return hipMemsetAsync(c.wf.counters + CTR(3), 0, 4, c.stream);
}

}  // namespace

struct pt_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    std::string error;
    std::vector<pt_ctx*> scenes;
    std::vector<pt_view> views;
    std::vector<uint32_t> sample_index;
    std::vector<size_t> accum_offset;  // in pixels
    float4* accum = nullptr;
    DeviceScene* d_scenes = nullptr;   // scene table (several scenes only)
    ViewDev* d_views = nullptr;
    ViewDev* h_views = nullptr;        // pinned staging
    JobDev* d_job = nullptr;   // the two job slots
    JobDev* h_job = nullptr;
    StreamState stream_state;
    Wavefront wf;
    uint32_t grid_blocks = 0;
    bool count_fetches = false;
    bool use_graphs = true;
    uint32_t lds_stack = LDS_STACK;
    bool sort_hits = false;
    KernelTimer timer;
    IterGraphs graphs;
    bool stats_fresh = true;
    unsigned long long last_paths = 0;
};

namespace {

thread_local std::string g_batch_error;

int bfail(pt_batch* b, int code, const std::string& msg)
{
    if (b)
        b->error = msg;
    else
        g_batch_error = msg;
    return code;
}

int bhip_fail(pt_batch* b, hipError_t e, const char* what)
{
    const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? PT_ENOMEM : PT_EDEVICE;
    return bfail(b, code, std::string(what) + ": " + hipGetErrorString(e));
}

#define BHIP_TRY(b, expr, what)                \
    do {                                       \
        hipError_t e_ = (expr);                \
        if (e_ != hipSuccess)                  \
            return bhip_fail((b), e_, (what)); \
    } while (0)

void batch_release(pt_batch& b)
{
    if (b.stream)
        (void)hipStreamSynchronize(b.stream);
    void* ptrs[] = {b.accum, b.d_scenes, b.d_views, b.d_job};
    for (void* p : ptrs)
        if (p)
            (void)hipFree(p);
    if (b.h_views)
        (void)hipHostFree(b.h_views);
    if (b.h_job)
        (void)hipHostFree(b.h_job);
    wavefront_free(b.wf);
    b.timer.destroy();
    b.graphs.destroy();
    if (b.ev_start)
        (void)hipEventDestroy(b.ev_start);
    if (b.ev_stop)
        (void)hipEventDestroy(b.ev_stop);
    if (b.stream)
        (void)hipStreamDestroy(b.stream);
}

int check_view(const pt_view& v, uint32_t scene_count, std::string& why)
{
    if (v.scene >= scene_count)
        return (why = "view scene index out of range", PT_EINVAL);
    if (v.width == 0 || v.height == 0 || v.width > 65535 || v.height > 65535 ||
        uint64_t(v.width) * v.height > (1ull << 27))
        return (why = "invalid view size", PT_EINVAL);
    if (v.spp == 0)
        return (why = "Sample count was zero.", PT_EINVAL);
    if (v.max_bounces > 1024)
        return (why = "max_bounces too large", PT_EINVAL);
    return PT_OK;
}

void view_dev(const pt_batch& b, uint32_t i, ViewDev& d)
{
    const pt_view& v = b.views[i];
    static_assert(sizeof(CameraBlock) == sizeof(pt_camera), "camera block layout");
    memcpy(&d.cam, &v.camera, sizeof(CameraBlock));
    d.accum = b.accum + b.accum_offset[i];
    d.width = v.width;
    d.height = v.height;
    d.npix = v.width * v.height;
    d.scene = b.scenes.size() > 1 ? v.scene : 0u;
    d.max_bounces = v.max_bounces;
    d.pad = 0;
}

}  // namespace

namespace {

// Renders return once all their paths have started; the last ones may still be in flight,
// overlapping the next render. Every call that reads the results or changes what those paths
// use finishes them first (and stamps the end of the render for the statistics).
int settle(Ctx& c)
{
    if (c.stream_state.idle())
        return PT_OK;
    HIP_TRY(&c, drain_stream(c), "render");
    HIP_TRY(&c, hipEventRecord(c.ev_stop, c.stream), "event");
    return PT_OK;
}

// settle() for the calls that hand out results: also reports a traversal stack overflow of
// the renders since the accumulation last restarted.
int settle_checked(Ctx& c)
{
    if (int v = settle(c))
        return v;
    HIP_TRY(&c, hipStreamSynchronize(c.stream), "sync");
    std::string why;
    if (overflowed(c, why))
        return fail(&c, PT_EDEVICE, why);
    return PT_OK;
}

}  // namespace

extern "C" {

int pt_create(const pt_scene_blob* scene, uint32_t width, uint32_t height, pt_ctx** out)
{
    if (!out)
        return fail(nullptr, PT_EINVAL, "out is null");
    *out = nullptr;
    if (width == 0 || height == 0 || width > 65535 || height > 65535 || uint64_t(width) * height > 0x7FFFFFFFull)
        return fail(nullptr, PT_EINVAL, "invalid image size (sides 1..65535)");
    std::string why;
    if (int v = validate_blob(scene, why))
        return fail(nullptr, v, why);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(nullptr, PT_ENODEV, "no HIP device");

    pt_ctx* h = new (std::nothrow) pt_ctx;
    if (!h)
        return fail(nullptr, PT_ENOMEM, "out of host memory");
    Ctx& c = h->c;
    auto bail = [&](int code) {
        g_create_error = c.error;
        release(c);
        delete h;
        return code;
    };
    int code;
    if ((code = [&]() -> int {
             HIP_TRY(&c, hipGetDevice(&c.device), "hipGetDevice");
             HIP_TRY(&c, scratch_setup(c.device), "staging area of pt_trace_rays_noalloc");
             HIP_TRY(&c, hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking), "hipStreamCreate");
             HIP_TRY(&c, hipEventCreate(&c.ev_start), "hipEventCreate");
             HIP_TRY(&c, hipEventCreate(&c.ev_stop), "hipEventCreate");
             HIP_TRY(&c, upload_sobol_table(), "sobol table");
             hipStream_t st = c.stream;
             HIP_TRY(&c, upload(c.d_vertices, scene->vertices, scene->vertex_count * 32, st), "vertices");
             HIP_TRY(&c, upload(c.d_indices, scene->indices, scene->index_count * 4, st), "indices");
             HIP_TRY(&c, upload(c.d_instances, scene->instances, size_t(scene->instance_count) * 128, st), "instances");
             HIP_TRY(&c, upload(c.d_materials, scene->materials, size_t(scene->material_count) * 96, st), "materials");
             HIP_TRY(&c, upload(c.d_light, scene->light_words, scene->light_word_count * 4, st), "light sampler");
             HIP_TRY(&c, upload(c.d_meshes, scene->meshes, size_t(scene->mesh_count) * sizeof(pt_mesh_desc), st),
                     "meshes");
             HIP_TRY(&c, upload(c.d_instance_mesh, scene->instance_mesh, size_t(scene->instance_count) * 4, st),
                     "instance meshes");
             // Opacity per instance: transparent materials (constant base-colour alpha) let
             // shadow rays through with transmittance 1 - alpha (shadowray.ah.hlsl:12-40).
             std::vector<float> opacity(std::max<uint32_t>(scene->instance_count, 1), 1.0f);
             const auto* inst = static_cast<const InstanceData*>(scene->instances);
             const auto* mats = static_cast<const MaterialRec*>(scene->materials);
             for (uint32_t i = 0; i < scene->instance_count; i++) {
                 const MaterialRec& m = mats[inst[i].material_index];
                 const uint32_t bits = float_as_u32(m.rmnb.w);
                 if (bits & (1u << 13))  // transparent: constant alpha, or -1 for the map's
                     opacity[i] = (bits & MAP_BASE_COLOUR) ? -1.0f : m.base_colour.w;
             }
             HIP_TRY(&c, upload(c.d_tex, scene->textures, size_t(scene->texture_count) * sizeof(pt_texture_desc), st),
                     "textures");
             HIP_TRY(&c, upload(c.d_texels, scene->texels, scene->texel_bytes, st), "textures");
             HIP_TRY(&c, upload(c.d_inst_opacity, opacity.data(), opacity.size() * 4, st), "opacity");
             c.transparent = std::any_of(opacity.begin(), opacity.end(), [](float o) { return o != 1.0f; });
             c.instance_count = scene->instance_count;

             c.meshes.assign(scene->meshes, scene->meshes + scene->mesh_count);
             c.light_count = scene->light_count;
             c.light_word_count = scene->light_word_count;
             c.light_layout_known = parse_emitter_tables(*scene, c.emitter_tables);
             if (const char* env = getenv("PT_BVH_LEAF"))
                 c.leaf_size = uint32_t(std::max(1, std::min(16, atoi(env))));
             if (int v = build_scene_bvh(c))
                 return v;

             DeviceScene& s = c.scene;
             s.vertices = c.d_vertices;
             s.indices = c.d_indices;
             s.instances = c.d_instances;
             s.materials = c.d_materials;
             s.light = c.d_light;
             s.light_count = scene->light_count;
             s.instance_count = scene->instance_count;
             s.nodes = c.bvh.nodes;
             s.tris = c.bvh.tris;
             s.inst_inv = c.bvh.inst_inv;
             s.tlas_entries = c.bvh.tlas_entries;
             s.inst_opacity = c.d_inst_opacity;
             s.tlas_root = c.bvh.tlas_root;
             s.envcol = make_float3(scene->envcol[0], scene->envcol[1], scene->envcol[2]);
             s.tex = c.d_tex;
             s.texels = c.d_texels;
             s.tex_count = scene->texture_count;

             c.width = width;
             c.height = height;
             HIP_TRY(&c, hipMalloc(&c.accum, size_t(width) * height * 16), "accumulation image");
             HIP_TRY(&c, hipMemsetAsync(c.accum, 0, size_t(width) * height * 16, st), "accumulation image");

             hipDeviceProp_t prop;
             HIP_TRY(&c, hipGetDeviceProperties(&prop, c.device), "device properties");
             c.cus = std::max(1, prop.multiProcessorCount);
             if (const char* env = getenv("PT_BLOCKS_PER_CU"))
                 c.blocks_per_cu = uint32_t(std::max(1, std::min(64, atoi(env))));
             const uint32_t npix = width * height;
             // Path pool: 16M slots (fewer for small images; the autotuner's usual pick at 1080p);
             // per-job radiance buffer for up to 2^27 paths (64 spp at 1920x1080).
             c.pool_slots = std::min<uint32_t>(16u << 20, std::max<uint32_t>(npix * 8u, 65536u));
             if (const char* env = getenv("PT_POOL_SLOTS"))
                 c.pool_slots = uint32_t(std::max(1024, atoi(env)));
             const uint32_t job_spp = std::max<uint32_t>(1, std::min<uint32_t>(1024, (1u << 27) / npix));
             c.job_paths = npix * job_spp;
             HIP_TRY(&c, apply_launch(c), "wavefront buffers");
             HIP_TRY(&c, hipMalloc(&c.d_views, sizeof(ViewDev)), "view table");
             HIP_TRY(&c, hipMalloc(&c.d_job, 2 * sizeof(JobDev)), "job table");
             HIP_TRY(&c, hipMemsetAsync(c.d_job, 0, 2 * sizeof(JobDev), st), "job table");
             HIP_TRY(&c, hipHostMalloc((void**)&c.h_view, sizeof(ViewDev), hipHostMallocDefault), "view table");
             HIP_TRY(&c, hipHostMalloc((void**)&c.h_job, 2 * sizeof(JobDev), hipHostMallocDefault), "job table");
             if (const char* env = getenv("PT_GRAPHS"))
                 c.use_graphs = atoi(env) != 0;
             // The light sampler's alias tables are built on the device from the uploaded scene
             // (TableLightSampler, src/scene/lightsample.cpp:5-178), over the words the caller
             // packed; words of another layout are used as given.
             if (c.light_layout_known && c.light_count)
                 HIP_TRY(&c, build_light_tables(c), "light tables");
             HIP_TRY(&c, hipStreamSynchronize(st), "upload");
             return PT_OK;
         }()) != PT_OK)
        return bail(code);
    *out = h;
    return PT_OK;
}

int pt_set_camera(pt_ctx* h, const pt_camera* cam)
{
    if (!h)
        return PT_EINVAL;
    if (int v = settle(h->c))
        return v;
    if (!cam)
        return fail(&h->c, PT_EINVAL, "camera is null");
    static_assert(sizeof(CameraBlock) == sizeof(pt_camera), "camera block layout");
    memcpy(&h->c.cam, cam, sizeof(CameraBlock));
    h->c.have_camera = true;
    h->c.sample_index = 0;
    h->c.stats_fresh = true;
    HIP_TRY(&h->c, clear_overflow(h->c), "error flags");
    return PT_OK;
}

int pt_set_max_bounces(pt_ctx* h, uint32_t max_bounces)
{
    if (!h)
        return PT_EINVAL;
    if (int v = settle(h->c))
        return v;
    if (max_bounces > 1024)
        return fail(&h->c, PT_EINVAL, "max_bounces too large");
    h->c.max_bounces = max_bounces;
    h->c.sample_index = 0;
    h->c.stats_fresh = true;
    HIP_TRY(&h->c, clear_overflow(h->c), "error flags");
    return PT_OK;
}

int pt_render(pt_ctx* h, uint32_t spp)
{
    if (!h)
        return PT_EINVAL;
    Ctx& c = h->c;
    if (spp == 0)
        return fail(&c, PT_EINVAL, "Sample count was zero.");
    if (!c.have_camera)
        return fail(&c, PT_ESTATE, "no camera set");
    if (uint64_t(c.sample_index) + spp > 0xFFFFFFFFull)
        return fail(&c, PT_EINVAL, "sample index overflow");
    if (c.stats_fresh) {
        c.timer.reset();
        HIP_TRY(&c, hipMemsetAsync(c.wf.totals, 0, TOTALS * 8, c.stream), "stats");
        HIP_TRY(&c, hipEventRecord(c.ev_start, c.stream), "event");
        c.last_paths = 0;
        c.stats_fresh = false;
    }
    HIP_TRY(&c, render_samples(c, spp, false), "render");
    HIP_TRY(&c, hipEventRecord(c.ev_stop, c.stream), "event");
    c.sample_index += spp;
    c.last_paths += uint64_t(spp) * c.width * c.height;
    if (c.stream_state.overflow) {  // seen in the counts the render steered by
        std::string why;
        overflowed(c, why);
        return fail(&c, PT_EDEVICE, why);
    }
    return PT_OK;
}

int pt_read_accumulation(pt_ctx* h, float* rgba, size_t row_pitch)
{
    if (!h)
        return PT_EINVAL;
    if (int v = settle_checked(h->c))
        return v;
    Ctx& c = h->c;
    if (!rgba || row_pitch < size_t(c.width) * 16)
        return fail(&c, PT_EINVAL, "bad output buffer");
    HIP_TRY(&c,
            hipMemcpy2DAsync(rgba, row_pitch, c.accum, size_t(c.width) * 16, size_t(c.width) * 16, c.height,
                             hipMemcpyDeviceToHost, c.stream),
            "readback");
    HIP_TRY(&c, hipStreamSynchronize(c.stream), "readback");
    return PT_OK;
}

namespace {

// ACES-tone-mapped RGBA8 of an accumulation image into host rows of row_pitch bytes.
hipError_t read_tonemapped(const float4* accum, uint32_t w, uint32_t h, uint8_t* out, size_t row_pitch,
                           hipStream_t st)
{
    uint32_t* d = nullptr;
    hipError_t e = hipMalloc(&d, std::max<size_t>(size_t(w) * h * 4, 4));
    if (e != hipSuccess)
        return e;
    e = tonemap_rgba8(accum, w, h, d, st);
    if (e == hipSuccess)
        e = hipMemcpy2DAsync(out, row_pitch, d, size_t(w) * 4, size_t(w) * 4, h, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    (void)hipFree(d);
    return e;
}

}  // namespace

int pt_read_tonemapped(pt_ctx* h, uint8_t* rgba8, size_t row_pitch)
{
    if (!h)
        return PT_EINVAL;
    if (int v = settle_checked(h->c))
        return v;
    Ctx& c = h->c;
    if (!rgba8 || row_pitch < size_t(c.width) * 4)
        return fail(&c, PT_EINVAL, "bad output buffer");
    HIP_TRY(&c, read_tonemapped(c.accum, c.width, c.height, rgba8, row_pitch, c.stream), "tone-mapped readback");
    return PT_OK;
}

int pt_trace_primary(pt_ctx* h, uint32_t sample_index, pt_hit* hits)
{
    if (!h)
        return PT_EINVAL;
    if (int v = settle(h->c))
        return v;
    Ctx& c = h->c;
    if (!hits)
        return fail(&c, PT_EINVAL, "hits is null");
    if (!c.have_camera)
        return fail(&c, PT_ESTATE, "no camera set");
    const size_t n = size_t(c.width) * c.height;
    pt_hit* d = nullptr;
    HIP_TRY(&c, hipMemsetAsync(c.wf.counters + CTR(9), 0, 4, c.stream), "error flags");
    HIP_TRY(&c, hipMalloc(&d, n * sizeof(pt_hit)), "hits");
    hipError_t e = primary_hits(c, sample_index, d);
    if (e == hipSuccess)
        e = hipMemcpyAsync(hits, d, n * sizeof(pt_hit), hipMemcpyDeviceToHost, c.stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(c.stream);
    (void)hipFree(d);
    if (e != hipSuccess)
        return hip_fail(&c, e, "primary hits");
    return query_overflow(c);
}

int pt_trace_rays(pt_ctx* h, const pt_ray* rays, uint32_t count, pt_hit* hits, int any_hit)
{
    if (!h)
        return PT_EINVAL;
    if (int v = settle(h->c))
        return v;
    Ctx& c = h->c;
    if (count == 0)
        return PT_OK;
    if (!rays || !hits)
        return fail(&c, PT_EINVAL, "null ray or hit array");
    pt_ray* dr = nullptr;
    pt_hit* dh = nullptr;
    HIP_TRY(&c, hipMemsetAsync(c.wf.counters + CTR(9), 0, 4, c.stream), "error flags");
    hipError_t e = hipMalloc(&dr, size_t(count) * sizeof(pt_ray));
    if (e == hipSuccess)
        e = hipMalloc(&dh, size_t(count) * sizeof(pt_hit));
    if (e == hipSuccess)
        e = hipMemcpyAsync(dr, rays, size_t(count) * sizeof(pt_ray), hipMemcpyHostToDevice, c.stream);
    if (e == hipSuccess)
        e = trace_rays(c, dr, count, dh, any_hit != 0);
    if (e == hipSuccess)
        e = hipMemcpyAsync(hits, dh, size_t(count) * sizeof(pt_hit), hipMemcpyDeviceToHost, c.stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(c.stream);
    if (dr)
        (void)hipFree(dr);
    if (dh)
        (void)hipFree(dh);
    if (e != hipSuccess)
        return hip_fail(&c, e, "trace rays");
    return query_overflow(c);
}

int pt_trace_rays_noalloc(pt_ctx* h, const pt_ray* rays, uint32_t count, pt_hit* hits, int any_hit)
{
    if (!h)
        return PT_EINVAL;
    if (int v = settle(h->c))
        return v;
    Ctx& c = h->c;
    if (count == 0)
        return PT_OK;
    if (!rays || !hits)
        return fail(&c, PT_EINVAL, "null ray or hit array");
    if (count > PT_NOALLOC_MAX_RAYS)
        return fail(&c, PT_EINVAL,
                    "at most " + std::to_string(PT_NOALLOC_MAX_RAYS) + " rays per pt_trace_rays_noalloc call");
    char* base = static_cast<char*>(scratch_ptr(c.device));
    if (!base)
        return fail(&c, PT_ESTATE, "staging area of pt_trace_rays_noalloc not set up on this device");
    pt_ray* dr = reinterpret_cast<pt_ray*>(base);
    pt_hit* dh = reinterpret_cast<pt_hit*>(base + size_t(PT_NOALLOC_MAX_RAYS) * sizeof(pt_ray));
    HIP_TRY(&c, hipMemsetAsync(c.wf.counters + CTR(9), 0, 4, c.stream), "error flags");
    HIP_TRY(&c, hipMemcpyAsync(dr, rays, size_t(count) * sizeof(pt_ray), hipMemcpyHostToDevice, c.stream), "upload");
    HIP_TRY(&c, trace_rays(c, dr, count, dh, any_hit != 0), "trace rays");
    HIP_TRY(&c, hipMemcpyAsync(hits, dh, size_t(count) * sizeof(pt_hit), hipMemcpyDeviceToHost, c.stream),
            "readback");
    HIP_TRY(&c, hipStreamSynchronize(c.stream), "trace rays");
    return query_overflow(c);
}

int pt_get_stats(pt_ctx* h, pt_stats* out)
{
    if (!h)
        return PT_EINVAL;
    if (int v = settle(h->c))
        return v;
    Ctx& c = h->c;
    if (!out)
        return fail(&c, PT_EINVAL, "out is null");
    HIP_TRY(&c, hipStreamSynchronize(c.stream), "sync");
    unsigned long long totals[10] = {};
    HIP_TRY(&c, hipMemcpy(totals, c.wf.totals, sizeof(totals), hipMemcpyDeviceToHost), "stats");
    float ms = 0.0f;
    if (c.last_paths)
        HIP_TRY(&c, hipEventElapsedTime(&ms, c.ev_start, c.ev_stop), "event");
    pt_stats s{};
    s.extension_rays = totals[0];
    s.shadow_rays = totals[1];
    s.paths = c.last_paths;
    s.render_ms = ms;
    s.bvh_build_ms = c.bvh_build_ms;
    s.bvh_bytes = uint64_t(c.bvh.node_count) * sizeof(Node4) + uint64_t(c.bvh.tri_count) * sizeof(TriRec);
    s.sample_index = c.sample_index;
    s.max_bounces = c.max_bounces;
    s.width = c.width;
    s.height = c.height;
    s.node_count = c.bvh.node_count;
    s.tri_record_count = c.bvh.tri_count;
    s.ext_node_fetches = totals[2];
    s.ext_tri_fetches = totals[3];
    s.shadow_node_fetches = totals[4];
    s.shadow_tri_fetches = totals[5];
    s.ext_tlas_node_fetches = totals[6];
    s.ext_instance_entries = totals[7];
    s.shadow_tlas_node_fetches = totals[8];
    s.shadow_instance_entries = totals[9];
    if (c.timer.enabled) {
        double kms[KERNEL_COUNT];
        uint32_t kl[KERNEL_COUNT];
        HIP_TRY(&c, c.timer.collect(kms, kl), "kernel timing");
        s.extend_ms = kms[KERNEL_EXTEND];
        s.shade_ms = kms[KERNEL_SHADE];
        s.shadow_ms = kms[KERNEL_SHADOW];
        s.extend_launches = kl[KERNEL_EXTEND];
        s.shade_launches = kl[KERNEL_SHADE];
        s.shadow_launches = kl[KERNEL_SHADOW];
        s.refill_ms = kms[KERNEL_REFILL];
        s.refill_launches = kl[KERNEL_REFILL];
    }
    *out = s;
    std::string why;
    if (overflowed(c, why))
        return fail(&c, PT_EDEVICE, why);
    return PT_OK;
}

int pt_reset_stats(pt_ctx* h)
{
    if (!h)
        return PT_EINVAL;
    if (int v = settle(h->c))
        return v;
    h->c.stats_fresh = true;
    return PT_OK;
}

int pt_set_option(pt_ctx* h, uint32_t option, int64_t value)
{
    if (!h)
        return PT_EINVAL;
    if (int v = settle(h->c))
        return v;
    Ctx& c = h->c;
    switch (option) {
    case PT_OPT_PROFILE_KERNELS:
        HIP_TRY(&c, hipStreamSynchronize(c.stream), "sync");
        c.timer.enabled = value != 0;
        c.stats_fresh = true;
        return PT_OK;
    case PT_OPT_COUNT_FETCHES:
        c.count_fetches = value != 0;
        return PT_OK;
    case PT_OPT_GRAPHS:
        HIP_TRY(&c, hipStreamSynchronize(c.stream), "sync");
        c.use_graphs = value != 0;
        return PT_OK;
    case PT_OPT_BLOCKS_PER_CU:
    case PT_OPT_POOL_SLOTS: {
        if (option == PT_OPT_BLOCKS_PER_CU && (value < 1 || value > 64))
            return fail(&c, PT_EINVAL, "blocks per CU must be 1..64");
        if (option == PT_OPT_POOL_SLOTS && (value < 1024 || value > (1 << 26)))
            return fail(&c, PT_EINVAL, "pool slots must be 1024..2^26");
        if (c.cache.configured)
            return fail(&c, PT_ESTATE, "disable the radiance cache before changing the path pool");
        HIP_TRY(&c, hipStreamSynchronize(c.stream), "sync");
        (option == PT_OPT_BLOCKS_PER_CU ? c.blocks_per_cu : c.pool_slots) = uint32_t(value);
        HIP_TRY(&c, apply_launch(c), "wavefront buffers");
        c.stats_fresh = true;
        return PT_OK;
    }
    case PT_OPT_SORT_HITS:
        HIP_TRY(&c, hipStreamSynchronize(c.stream), "sync");
        c.sort_hits = value != 0;
        return PT_OK;
    case PT_OPT_LDS_STACK:
        if (value != 8 && value != 16 && value != 32)
            return fail(&c, PT_EINVAL, "LDS stack depth must be 8, 16 or 32");
        HIP_TRY(&c, hipStreamSynchronize(c.stream), "sync");
        c.lds_stack = uint32_t(value);
        return PT_OK;
    case PT_OPT_SPILL_DEPTH:
        if (value < 0 || value > int64_t(SPILL_STACK))
            return fail(&c, PT_EINVAL, "spill depth must be 0.." + std::to_string(SPILL_STACK));
        HIP_TRY(&c, hipStreamSynchronize(c.stream), "sync");
        c.spill_depth = c.wf.spill_depth = uint32_t(value);
        c.graphs.destroy();  // the captured launches carry the Wavefront by value
        return PT_OK;
    case PT_OPT_LEAF_SIZE:
    case PT_OPT_NODE_WIDTH:
    case PT_OPT_TLAS_OPEN:
    case PT_OPT_TLAS_BOXES:
    case PT_OPT_TLAS_BUILDER:
    case PT_OPT_BVH_BUILDER: {
        if (option == PT_OPT_LEAF_SIZE && (value < 1 || value > 16))
            return fail(&c, PT_EINVAL, "leaf size must be 1..16");
        if (option == PT_OPT_NODE_WIDTH && (value < 2 || value > 4))
            return fail(&c, PT_EINVAL, "node width must be 2..4");
        if (option == PT_OPT_BVH_BUILDER && value != BVH_BUILDER_PLOC && value != BVH_BUILDER_LBVH)
            return fail(&c, PT_EINVAL, "BVH builder must be 0 (PLOC) or 1 (radix tree)");
        if (option == PT_OPT_TLAS_OPEN && (value < 0 || value > 2))
            return fail(&c, PT_EINVAL, "TLAS opening depth must be 0..2");
        if (option == PT_OPT_TLAS_BOXES && (value < 0 || value > 3))
            return fail(&c, PT_EINVAL, "TLAS entry boxes must be 0..3");
        if (option == PT_OPT_TLAS_BUILDER && value != BVH_TLAS_DEVICE && value != BVH_TLAS_HOST)
            return fail(&c, PT_EINVAL, "TLAS builder must be 0 (device) or 1 (host SAH sweep)");
        HIP_TRY(&c, hipStreamSynchronize(c.stream), "sync");
        uint32_t& field = option == PT_OPT_LEAF_SIZE    ? c.leaf_size
                          : option == PT_OPT_NODE_WIDTH ? c.node_width
                          : option == PT_OPT_TLAS_OPEN  ? c.tlas_open
                          : option == PT_OPT_TLAS_BOXES ? c.tlas_boxes
                          : option == PT_OPT_TLAS_BUILDER ? c.tlas_builder
                                                        : c.bvh_builder;
        field = uint32_t(value);
        if (int v = build_scene_bvh(c))
            return v;
        c.stats_fresh = true;
        return PT_OK;
    }
    default:
        return fail(&c, PT_EINVAL, "unknown option " + std::to_string(option));
    }
}

int pt_cache_configure(pt_ctx* h, const pt_cache_config* cfg)
{
    if (!h)
        return PT_EINVAL;
    if (int v = settle(h->c))
        return v;
    Ctx& c = h->c;
    if (!cfg)
        return fail(&c, PT_EINVAL, "config is null");
    HIP_TRY(&c, hipStreamSynchronize(c.stream), "sync");
    if (!cfg->enable) {
        cache_free(c.cache);
        c.sample_index = 0;
        c.stats_fresh = true;
        return PT_OK;
    }
    if (cfg->train_stride == 0 || cfg->batch == 0 || !(cfg->learning_rate > 0.0f) || cfg->batch > (1u << 24) ||
        !(cfg->continue_prob >= 0.0f && cfg->continue_prob <= 1.0f) || cfg->query_vertex == 1 ||
        cfg->query_vertex > 8 || !(cfg->ema_decay >= 0.0f && cfg->ema_decay < 1.0f) ||
        !(cfg->query_max_pdf >= 0.0f))
        return fail(&c, PT_EINVAL, "invalid cache configuration");
    CacheState& cs = c.cache;
    cs.query = cfg->query != 0;
    cs.train_stride = cfg->train_stride;
    cs.train_steps = cfg->train_steps;
    cs.batch = cfg->batch;
    cs.lr = cfg->learning_rate;
    cs.seed = cfg->seed;
    cs.frame_spp = cfg->frame_spp;
    cs.cont_prob = cfg->continue_prob;
    cs.query_vertex = cfg->query_vertex ? cfg->query_vertex : 2u;
    cs.ema_decay = cfg->ema_decay;
    cs.query_max_pdf = cfg->query_max_pdf;
    cs.records_total = 0;
    cs.last_records = 0;
    // one training record per training path of a job
    const uint64_t job_paths = uint64_t(c.wf.rad_out_capacity);
    const uint32_t rec_cap = uint32_t(std::min<uint64_t>(job_paths / cfg->train_stride + 1024, 1u << 24));
    HIP_TRY(&c, cache_init(cs, c.wf.capacity, rec_cap, c.stream), "radiance cache");
    HIP_TRY(&c, clear_overflow(c), "error flags");
    HIP_TRY(&c, hipStreamSynchronize(c.stream), "radiance cache");
    c.sample_index = 0;
    c.stats_fresh = true;
    return PT_OK;
}

int pt_cache_get_stats(pt_ctx* h, pt_cache_stats* out)
{
    if (!h)
        return PT_EINVAL;
    if (int v = settle(h->c))
        return v;
    Ctx& c = h->c;
    if (!out)
        return fail(&c, PT_EINVAL, "out is null");
    pt_cache_stats s{};
    const CacheState& cs = c.cache;
    s.enabled = cs.configured ? 1u : 0u;
    s.query = cs.query ? 1u : 0u;
    s.steps = cs.steps_done;
    s.last_records = cs.last_records;
    s.records_total = cs.records_total;
    s.param_count = uint32_t(mlp_param_count());
    if (cs.configured) {
        HIP_TRY(&c, hipStreamSynchronize(c.stream), "sync");
        HIP_TRY(&c, hipMemcpy(&s.last_loss, cs.loss_acc + 1, 4, hipMemcpyDeviceToHost), "cache loss");
        if (cs.frame_spp == 0) {  // in-flight training: records of the current job
            uint32_t n = 0;
            HIP_TRY(&c, hipMemcpy(&n, cs.counters, 4, hipMemcpyDeviceToHost), "cache records");
            s.last_records = std::min(n, cs.record_capacity);
        }
    }
    for (int k = 0; k < 3; k++) {
        s.bounds_lo[k] = cs.bounds_lo[k];
        s.bounds_hi[k] = cs.bounds_hi[k];
    }
    *out = s;
    return PT_OK;
}

uint32_t pt_cache_param_count(void) { return uint32_t(mlp_param_count()); }


int pt_cache_get_params(pt_ctx* h, float* params, size_t count)
{
    if (!h)
        return PT_EINVAL;
    if (int v = settle(h->c))
        return v;
    Ctx& c = h->c;
    if (int v = need_cache(c))
        return v;
    if (!params || count != size_t(mlp_param_count()))
        return fail(&c, PT_EINVAL, "parameter buffer must hold pt_cache_param_count() floats");
    HIP_TRY(&c, cache_get_params(c.cache, params, c.stream), "cache parameters");
    return PT_OK;
}

int pt_cache_get_query_params(pt_ctx* h, float* params, size_t count)
{
    if (!h)
        return PT_EINVAL;
    if (int v = settle(h->c))
        return v;
    Ctx& c = h->c;
    if (int v = need_cache(c))
        return v;
    if (!params || count != size_t(mlp_param_count()))
        return fail(&c, PT_EINVAL, "parameter buffer must hold pt_cache_param_count() floats");
    HIP_TRY(&c, cache_get_query_params(c.cache, params, c.stream), "cache parameters");
    return PT_OK;
}

int pt_cache_set_params(pt_ctx* h, const float* params, size_t count)
{
    if (!h)
        return PT_EINVAL;
    if (int v = settle(h->c))
        return v;
    Ctx& c = h->c;
    if (int v = need_cache(c))
        return v;
    if (!params || count != size_t(mlp_param_count()))
        return fail(&c, PT_EINVAL, "parameter buffer must hold pt_cache_param_count() floats");
    HIP_TRY(&c, cache_set_params(c.cache, params, c.stream), "cache parameters");
    HIP_TRY(&c, hipStreamSynchronize(c.stream), "cache parameters");
    return PT_OK;
}

int pt_cache_forward(pt_ctx* h, const float* features, uint32_t n, float* out)
{
    if (!h)
        return PT_EINVAL;
    if (int v = settle(h->c))
        return v;
    Ctx& c = h->c;
    if (int v = need_cache(c))
        return v;
    if (n == 0)
        return PT_OK;
    if (!features || !out)
        return fail(&c, PT_EINVAL, "null buffer");
    DevBufs d;
    float *df = nullptr, *dout = nullptr;
    HIP_TRY(&c, d.alloc(df, size_t(n) * PT_CACHE_FEATURES * 4), "cache buffers");
    HIP_TRY(&c, d.alloc(dout, size_t(n) * 3 * 4), "cache buffers");
    HIP_TRY(&c, hipMemcpyAsync(df, features, size_t(n) * PT_CACHE_FEATURES * 4, hipMemcpyHostToDevice, c.stream),
            "upload");
    HIP_TRY(&c, cache_forward(c.cache, df, n, dout, c.stream), "cache forward");
    HIP_TRY(&c, hipMemcpyAsync(out, dout, size_t(n) * 3 * 4, hipMemcpyDeviceToHost, c.stream), "readback");
    HIP_TRY(&c, hipStreamSynchronize(c.stream), "cache forward");
    return PT_OK;
}

int pt_cache_encode(pt_ctx* h, const float* records, uint32_t n, float* features)
{
    if (!h)
        return PT_EINVAL;
    if (int v = settle(h->c))
        return v;
    Ctx& c = h->c;
    if (int v = need_cache(c))
        return v;
    if (n == 0)
        return PT_OK;
    if (!records || !features)
        return fail(&c, PT_EINVAL, "null buffer");
    DevBufs d;
    float *dr = nullptr, *df = nullptr;
    HIP_TRY(&c, d.alloc(dr, size_t(n) * 13 * 4), "cache buffers");
    HIP_TRY(&c, d.alloc(df, size_t(n) * PT_CACHE_FEATURES * 4), "cache buffers");
    HIP_TRY(&c, hipMemcpyAsync(dr, records, size_t(n) * 13 * 4, hipMemcpyHostToDevice, c.stream), "upload");
    HIP_TRY(&c, cache_encode(c.cache, dr, n, df, c.stream), "cache encode");
    HIP_TRY(&c, hipMemcpyAsync(features, df, size_t(n) * PT_CACHE_FEATURES * 4, hipMemcpyDeviceToHost, c.stream),
            "readback");
    HIP_TRY(&c, hipStreamSynchronize(c.stream), "cache encode");
    return PT_OK;
}

int pt_cache_train_batch(pt_ctx* h, const float* features, const float* targets, uint32_t n, int apply,
                         float* grads, float* loss)
{
    if (!h)
        return PT_EINVAL;
    if (int v = settle(h->c))
        return v;
    Ctx& c = h->c;
    if (int v = need_cache(c))
        return v;
    if (n == 0 || !features || !targets)
        return fail(&c, PT_EINVAL, "empty or null batch");
    DevBufs d;
    float *df = nullptr, *dt = nullptr;
    HIP_TRY(&c, d.alloc(df, size_t(n) * PT_CACHE_FEATURES * 4), "cache buffers");
    HIP_TRY(&c, d.alloc(dt, size_t(n) * 3 * 4), "cache buffers");
    HIP_TRY(&c, hipMemcpyAsync(df, features, size_t(n) * PT_CACHE_FEATURES * 4, hipMemcpyHostToDevice, c.stream),
            "upload");
    HIP_TRY(&c, hipMemcpyAsync(dt, targets, size_t(n) * 3 * 4, hipMemcpyHostToDevice, c.stream), "upload");
    HIP_TRY(&c, hipMemsetAsync(c.cache.grads, 0, size_t(mlp_param_count()) * 4, c.stream), "cache gradients");
    HIP_TRY(&c, cache_backward(c.cache, df, dt, n, 0, 0, 1, c.stream), "cache backward");
    if (grads)
        HIP_TRY(&c,
                hipMemcpyAsync(grads, c.cache.grads, size_t(mlp_param_count()) * 4, hipMemcpyDeviceToHost, c.stream),
                "readback");
    if (loss)
        HIP_TRY(&c, hipMemcpyAsync(loss, c.cache.loss_acc, 4, hipMemcpyDeviceToHost, c.stream), "readback");
    if (apply) {
        HIP_TRY(&c, cache_adam(c.cache, c.stream), "cache update");
    } else {  // gradient and loss discarded: training steps start from zero
        HIP_TRY(&c, hipMemsetAsync(c.cache.grads, 0, size_t(mlp_param_count()) * 4, c.stream), "cache gradients");
        HIP_TRY(&c, hipMemsetAsync(c.cache.loss_acc, 0, 4, c.stream), "cache loss");
    }
    HIP_TRY(&c, hipStreamSynchronize(c.stream), "cache train");
    return PT_OK;
}

}  // extern "C"

namespace {

RenderIO batch_io(pt_batch& b)
{
    RenderIO io{b.stream,   &b.wf,   b.grid_blocks, b.count_fetches,
                &b.timer,   b.scenes[0]->c.scene,   b.d_scenes,     b.d_views,
                b.d_job,    b.h_job, &b.stream_state,
                b.use_graphs ? &b.graphs : nullptr, nullptr, b.lds_stack, b.sort_hits};
    io.alpha_shadows = std::any_of(b.scenes.begin(), b.scenes.end(), [](const pt_ctx* c) { return c->c.transparent; });
    io.leaf1 = std::all_of(b.scenes.begin(), b.scenes.end(), [](const pt_ctx* c) { return c->c.single_tri_leaves; });
    return io;
}

// As settle() for a context: finishes the paths a batch render left in flight.
int bsettle(pt_batch* b)
{
    if (b->stream_state.idle())
        return PT_OK;
    BHIP_TRY(b, render_views(batch_io(*b), {}, true), "render");
    BHIP_TRY(b, hipEventRecord(b->ev_stop, b->stream), "event");
    return PT_OK;
}

// As settle_checked() for a context: a traversal stack overflow of the batch's renders since a
// view last restarted is an error of the calls that hand out results.
int bsettle_checked(pt_batch* b)
{
    if (int v = bsettle(b))
        return v;
    BHIP_TRY(b, hipStreamSynchronize(b->stream), "sync");
    uint32_t flags = 0;
    BHIP_TRY(b, hipMemcpy(&flags, b->wf.counters + CTR(3), 4, hipMemcpyDeviceToHost), "error flags");
    if (flags || b->stream_state.overflow) {
        b->stream_state.overflow = true;
        return bfail(b, PT_EDEVICE,
                     "traversal stack overflow: the batch render missed BVH nodes; its images are invalid until "
                     "the views restart (pt_batch_set_view)");
    }
    return PT_OK;
}

}  // namespace

extern "C" {

int pt_batch_create(pt_ctx* const* scenes, uint32_t scene_count, const pt_view* views, uint32_t view_count,
                    pt_batch** out)
{
    if (!out)
        return bfail(nullptr, PT_EINVAL, "out is null");
    *out = nullptr;
    if (!scenes || scene_count == 0 || !views || view_count == 0)
        return bfail(nullptr, PT_EINVAL, "a batch needs at least one scene and one view");
    if (view_count > 65536)
        return bfail(nullptr, PT_EINVAL, "too many views");
    for (uint32_t i = 0; i < scene_count; i++)
        if (!scenes[i])
            return bfail(nullptr, PT_EINVAL, "scene context is null");
    std::string why;
    for (uint32_t i = 0; i < view_count; i++)
        if (int v = check_view(views[i], scene_count, why))
            return bfail(nullptr, v, "view " + std::to_string(i) + ": " + why);
    const int device = scenes[0]->c.device;
    for (uint32_t i = 1; i < scene_count; i++)
        if (scenes[i]->c.device != device)
            return bfail(nullptr, PT_EINVAL, "scenes live on different devices");
    pt_batch* b = new (std::nothrow) pt_batch;
    if (!b)
        return bfail(nullptr, PT_ENOMEM, "out of host memory");
    int code = [&]() -> int {
        b->device = device;
        BHIP_TRY(b, hipSetDevice(device), "hipSetDevice");
        BHIP_TRY(b, hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
        BHIP_TRY(b, hipEventCreate(&b->ev_start), "hipEventCreate");
        BHIP_TRY(b, hipEventCreate(&b->ev_stop), "hipEventCreate");
        b->scenes.assign(scenes, scenes + scene_count);
        b->views.assign(views, views + view_count);
        b->sample_index.assign(view_count, 0);
        size_t pix = 0;
        for (uint32_t i = 0; i < view_count; i++) {
            b->accum_offset.push_back(pix);
            pix += size_t(views[i].width) * views[i].height;
        }
        BHIP_TRY(b, hipMalloc(&b->accum, pix * 16), "accumulation images");
        BHIP_TRY(b, hipMemsetAsync(b->accum, 0, pix * 16, b->stream), "accumulation images");
        if (scene_count > 1) {
            std::vector<DeviceScene> tab(scene_count);
            for (uint32_t i = 0; i < scene_count; i++)
                tab[i] = scenes[i]->c.scene;
            BHIP_TRY(b, hipMalloc(&b->d_scenes, scene_count * sizeof(DeviceScene)), "scene table");
            BHIP_TRY(b,
                     hipMemcpyAsync(b->d_scenes, tab.data(), scene_count * sizeof(DeviceScene), hipMemcpyHostToDevice,
                                    b->stream),
                     "scene table");
        }
        BHIP_TRY(b, hipMalloc(&b->d_views, view_count * sizeof(ViewDev)), "view table");
        BHIP_TRY(b, hipHostMalloc((void**)&b->h_views, view_count * sizeof(ViewDev), hipHostMallocDefault),
                 "view table");
        BHIP_TRY(b, hipMalloc(&b->d_job, 2 * sizeof(JobDev)), "job table");
        BHIP_TRY(b, hipMemset(b->d_job, 0, 2 * sizeof(JobDev)), "job table");
        BHIP_TRY(b, hipHostMalloc((void**)&b->h_job, 2 * sizeof(JobDev), hipHostMallocDefault), "job table");
        for (uint32_t i = 0; i < view_count; i++)
            view_dev(*b, i, b->h_views[i]);
        BHIP_TRY(b,
                 hipMemcpyAsync(b->d_views, b->h_views, view_count * sizeof(ViewDev), hipMemcpyHostToDevice,
                                b->stream),
                 "view table");
        hipDeviceProp_t prop;
        BHIP_TRY(b, hipGetDeviceProperties(&prop, device), "device properties");
        const uint32_t cus = std::max(1, prop.multiProcessorCount);
        uint32_t blocks_per_cu = 6;
        if (const char* env = getenv("PT_BLOCKS_PER_CU"))
            blocks_per_cu = uint32_t(std::max(1, std::min(64, atoi(env))));
        b->grid_blocks = ((cus * blocks_per_cu + 7) / 8) * 8;
        // at most 16M paths in flight: measured against 8M, +3 % on 64 views of one scene and
        // +12 % on 16 scenes x 4 views
        uint32_t pool = std::min<uint64_t>(16u << 20, std::max<uint64_t>(pix * 4, 65536));
        if (const char* env = getenv("PT_POOL_SLOTS"))
            pool = uint32_t(std::max(1024, atoi(env)));
        BHIP_TRY(b, wavefront_alloc(b->wf, pool, 1u << 27, b->grid_blocks * TRACE_BLOCK), "wavefront buffers");
        if (const char* env = getenv("PT_GRAPHS"))
            b->use_graphs = atoi(env) != 0;
        BHIP_TRY(b, hipStreamSynchronize(b->stream), "upload");
        return PT_OK;
    }();
    if (code != PT_OK) {
        g_batch_error = b->error;
        batch_release(*b);
        delete b;
        return code;
    }
    *out = b;
    return PT_OK;
}

int pt_batch_set_view(pt_batch* b, uint32_t view, const pt_view* v)
{
    if (!b)
        return PT_EINVAL;
    if (int v = bsettle(b))
        return v;
    if (!v || view >= b->views.size())
        return bfail(b, PT_EINVAL, "no such view");
    const pt_view& old = b->views[view];
    if (v->width != old.width || v->height != old.height)
        return bfail(b, PT_EINVAL, "a view keeps its image size");
    std::string why;
    if (int c = check_view(*v, uint32_t(b->scenes.size()), why))
        return bfail(b, c, why);
    BHIP_TRY(b, hipStreamSynchronize(b->stream), "sync");
    b->views[view] = *v;
    b->sample_index[view] = 0;
    b->stream_state.overflow = false;
    b->stream_state.flags_from = b->stream_state.launched;
    BHIP_TRY(b, hipMemsetAsync(b->wf.counters + CTR(3), 0, 4, b->stream), "error flags");
    view_dev(*b, view, b->h_views[view]);
    BHIP_TRY(b,
             hipMemcpyAsync(b->d_views + view, b->h_views + view, sizeof(ViewDev), hipMemcpyHostToDevice, b->stream),
             "view table");
    b->stats_fresh = true;
    return PT_OK;
}

int pt_batch_render(pt_batch* b)
{
    if (!b)
        return PT_EINVAL;
    std::vector<ViewWork> work;
    unsigned long long paths = 0;
    for (uint32_t i = 0; i < b->views.size(); i++) {
        const pt_view& v = b->views[i];
        if (uint64_t(b->sample_index[i]) + v.spp > 0xFFFFFFFFull)
            return bfail(b, PT_EINVAL, "sample index overflow");
        work.push_back(ViewWork{i, v.width * v.height, b->sample_index[i], v.spp});
        paths += (unsigned long long)v.width * v.height * v.spp;
    }
    if (b->stats_fresh) {
        b->timer.reset();
        BHIP_TRY(b, hipMemsetAsync(b->wf.totals, 0, TOTALS * 8, b->stream), "stats");
        BHIP_TRY(b, hipEventRecord(b->ev_start, b->stream), "event");
        b->last_paths = 0;
        b->stats_fresh = false;
    }
    BHIP_TRY(b, render_views(batch_io(*b), work, false), "render");
    BHIP_TRY(b, hipEventRecord(b->ev_stop, b->stream), "event");
    for (uint32_t i = 0; i < b->views.size(); i++)
        b->sample_index[i] += b->views[i].spp;
    b->last_paths += paths;
    if (b->stream_state.overflow)
        return bfail(b, PT_EDEVICE, "traversal stack overflow: the batch render missed BVH nodes");
    return PT_OK;
}

int pt_batch_read(pt_batch* b, uint32_t view, float* rgba, size_t row_pitch)
{
    if (!b)
        return PT_EINVAL;
    if (int v = bsettle_checked(b))
        return v;
    if (view >= b->views.size())
        return bfail(b, PT_EINVAL, "no such view");
    const pt_view& v = b->views[view];
    if (!rgba || row_pitch < size_t(v.width) * 16)
        return bfail(b, PT_EINVAL, "bad output buffer");
    BHIP_TRY(b,
             hipMemcpy2DAsync(rgba, row_pitch, b->accum + b->accum_offset[view], size_t(v.width) * 16,
                              size_t(v.width) * 16, v.height, hipMemcpyDeviceToHost, b->stream),
             "readback");
    BHIP_TRY(b, hipStreamSynchronize(b->stream), "readback");
    return PT_OK;
}

int pt_batch_read_tonemapped(pt_batch* b, uint32_t view, uint8_t* rgba8, size_t row_pitch)
{
    if (!b)
        return PT_EINVAL;
    if (int v = bsettle_checked(b))
        return v;
    if (view >= b->views.size())
        return bfail(b, PT_EINVAL, "no such view");
    const pt_view& v = b->views[view];
    if (!rgba8 || row_pitch < size_t(v.width) * 4)
        return bfail(b, PT_EINVAL, "bad output buffer");
    BHIP_TRY(b, read_tonemapped(b->accum + b->accum_offset[view], v.width, v.height, rgba8, row_pitch, b->stream),
             "tone-mapped readback");
    return PT_OK;
}

int pt_batch_get_stats(pt_batch* b, pt_stats* out)
{
    if (!b)
        return PT_EINVAL;
    if (int v = bsettle(b))
        return v;
    if (!out)
        return bfail(b, PT_EINVAL, "out is null");
    BHIP_TRY(b, hipStreamSynchronize(b->stream), "sync");
    unsigned long long totals[10] = {};
    uint32_t error_flags = 0;
    BHIP_TRY(b, hipMemcpy(totals, b->wf.totals, sizeof(totals), hipMemcpyDeviceToHost), "stats");
    BHIP_TRY(b, hipMemcpy(&error_flags, b->wf.counters + CTR(3), 4, hipMemcpyDeviceToHost), "stats");
    error_flags |= b->stream_state.overflow ? 1u : 0u;
    float ms = 0.0f;
    if (b->last_paths)
        BHIP_TRY(b, hipEventElapsedTime(&ms, b->ev_start, b->ev_stop), "event");
    pt_stats s{};
    s.extension_rays = totals[0];
    s.shadow_rays = totals[1];
    s.paths = b->last_paths;
    s.render_ms = ms;
    for (pt_ctx* c : b->scenes) {
        s.bvh_build_ms += c->c.bvh_build_ms;
        s.bvh_bytes += uint64_t(c->c.bvh.node_count) * sizeof(Node4) + uint64_t(c->c.bvh.tri_count) * sizeof(TriRec);
        s.node_count += c->c.bvh.node_count;
        s.tri_record_count += c->c.bvh.tri_count;
    }
    s.ext_node_fetches = totals[2];
    s.ext_tri_fetches = totals[3];
    s.shadow_node_fetches = totals[4];
    s.shadow_tri_fetches = totals[5];
    s.ext_tlas_node_fetches = totals[6];
    s.ext_instance_entries = totals[7];
    s.shadow_tlas_node_fetches = totals[8];
    s.shadow_instance_entries = totals[9];
    if (b->timer.enabled) {
        double kms[KERNEL_COUNT];
        uint32_t kl[KERNEL_COUNT];
        BHIP_TRY(b, b->timer.collect(kms, kl), "kernel timing");
        s.extend_ms = kms[KERNEL_EXTEND];
        s.shade_ms = kms[KERNEL_SHADE];
        s.shadow_ms = kms[KERNEL_SHADOW];
        s.extend_launches = kl[KERNEL_EXTEND];
        s.shade_launches = kl[KERNEL_SHADE];
        s.shadow_launches = kl[KERNEL_SHADOW];
        s.refill_ms = kms[KERNEL_REFILL];
        s.refill_launches = kl[KERNEL_REFILL];
    }
    *out = s;
    if (error_flags)
        return bfail(b, PT_EDEVICE, "traversal stack overflow");
    return PT_OK;
}

int pt_batch_set_option(pt_batch* b, uint32_t option, int64_t value)
{
    if (!b)
        return PT_EINVAL;
    if (int v = bsettle(b))
        return v;
    switch (option) {
    case PT_OPT_PROFILE_KERNELS:
        BHIP_TRY(b, hipStreamSynchronize(b->stream), "sync");
        b->timer.enabled = value != 0;
        b->stats_fresh = true;
        return PT_OK;
    case PT_OPT_COUNT_FETCHES:
        b->count_fetches = value != 0;
        return PT_OK;
    case PT_OPT_GRAPHS:
        BHIP_TRY(b, hipStreamSynchronize(b->stream), "sync");
        b->use_graphs = value != 0;
        return PT_OK;
    case PT_OPT_SORT_HITS:
        BHIP_TRY(b, hipStreamSynchronize(b->stream), "sync");
        b->sort_hits = value != 0;
        return PT_OK;
    case PT_OPT_LDS_STACK:
        if (value != 8 && value != 16 && value != 32)
            return bfail(b, PT_EINVAL, "LDS stack depth must be 8, 16 or 32");
        BHIP_TRY(b, hipStreamSynchronize(b->stream), "sync");
        b->lds_stack = uint32_t(value);
        return PT_OK;
    default:
        return bfail(b, PT_EINVAL, "unknown option " + std::to_string(option));
    }
}

int pt_batch_reset_stats(pt_batch* b)
{
    if (!b)
        return PT_EINVAL;
    if (int v = bsettle(b))
        return v;
    b->stats_fresh = true;
    return PT_OK;
}

const char* pt_batch_last_error(const pt_batch* b) { return b ? b->error.c_str() : g_batch_error.c_str(); }

void pt_batch_destroy(pt_batch* b)
{
    if (!b)
        return;
    (void)bsettle(b);
    batch_release(*b);
    delete b;
}

int pt_autotune(pt_ctx* h, uint32_t spp, pt_tune_result* out)
{
    if (!h)
        return PT_EINVAL;
    if (int v = settle(h->c))
        return v;
    Ctx& c = h->c;
    if (spp == 0)
        return fail(&c, PT_EINVAL, "Sample count was zero.");
    if (!c.have_camera)
        return fail(&c, PT_ESTATE, "no camera set");
    if (c.cache.configured)
        return fail(&c, PT_ESTATE, "disable the radiance cache before tuning");
    if (c.timer.enabled || c.count_fetches)
        return fail(&c, PT_ESTATE, "instrumentation options are on");
    HIP_TRY(&c, hipStreamSynchronize(c.stream), "sync");
    struct Cfg {
        int64_t bpc, pool, lds, leaf, width, open;
        bool operator==(const Cfg& o) const
        {
            return bpc == o.bpc && pool == o.pool && lds == o.lds && leaf == o.leaf && width == o.width &&
                   open == o.open;
        }
    };
    const Cfg start{c.blocks_per_cu, c.pool_slots, c.lds_stack, c.leaf_size, c.node_width, c.tlas_open};
    uint32_t trials = 0;
    // device nanoseconds per traced ray of spp samples of the current view in the streaming
    // steady state the renders run in: calls that return once their paths have all started, the
    // next call's paths entering while the previous ones drain (best of two timed calls after
    // one that fills the pool, warms the caches and captures the graphs; the stream is drained
    // only when the configuration changes)
    auto measure = [&](const Cfg& g, double& ns) -> int {
        if (g.bpc != int64_t(c.blocks_per_cu))
            if (int v = pt_set_option(h, PT_OPT_BLOCKS_PER_CU, g.bpc))
                return v;
        if (g.pool != int64_t(c.pool_slots))
            if (int v = pt_set_option(h, PT_OPT_POOL_SLOTS, g.pool))
                return v;
        if (int v = pt_set_option(h, PT_OPT_LDS_STACK, g.lds))
            return v;
        if (g.leaf != int64_t(c.leaf_size))
            if (int v = pt_set_option(h, PT_OPT_LEAF_SIZE, g.leaf))
                return v;
        if (g.width != int64_t(c.node_width))
            if (int v = pt_set_option(h, PT_OPT_NODE_WIDTH, g.width))
                return v;
        if (g.open != int64_t(c.tlas_open))
            if (int v = pt_set_option(h, PT_OPT_TLAS_OPEN, g.open))
                return v;
        ns = 1e300;
        c.sample_index = 0;
        HIP_TRY(&c, render_samples(c, spp, false), "render");
        c.sample_index += spp;
        for (int rep = 0; rep < 2; rep++) {
            unsigned long long t0[2], t1[2];
            HIP_TRY(&c, hipStreamSynchronize(c.stream), "sync");
            HIP_TRY(&c, hipMemcpy(t0, c.wf.totals, sizeof(t0), hipMemcpyDeviceToHost), "stats");
            HIP_TRY(&c, hipEventRecord(c.ev_start, c.stream), "event");
            HIP_TRY(&c, render_samples(c, spp, false), "render");
            c.sample_index += spp;
            HIP_TRY(&c, hipEventRecord(c.ev_stop, c.stream), "event");
            HIP_TRY(&c, hipEventSynchronize(c.ev_stop), "event");
            float ms = 0.0f;
            HIP_TRY(&c, hipEventElapsedTime(&ms, c.ev_start, c.ev_stop), "event");
            HIP_TRY(&c, hipMemcpy(t1, c.wf.totals, sizeof(t1), hipMemcpyDeviceToHost), "stats");
            const double rays = double(t1[0] - t0[0]) + double(t1[1] - t0[1]);
            if (rays > 0)
                ns = std::min(ns, ms * 1e6 / rays);
        }
        HIP_TRY(&c, drain_stream(c), "render");
        trials++;
        return PT_OK;
    };
    double base_ns = 0.0;
    if (int v = measure(start, base_ns))
        return v;
    Cfg best = start;
    double best_ns = base_ns;
    const uint64_t npix = uint64_t(c.width) * c.height;
    const std::vector<int64_t> bpcs = {2, 3, 4, 5, 6, 8};
    const std::vector<int64_t> ldss = {8, 16, 32};
    std::vector<int64_t> pools;
    for (int64_t p : {1 << 20, 2 << 20, 4 << 20, 8 << 20, 16 << 20, 32 << 20})
        if (uint64_t(p) <= std::max<uint64_t>(npix * spp, 1u << 20))
            pools.push_back(p);
    const std::vector<int64_t> leaves = {1, 2, 4, 8};
    const std::vector<int64_t> widths = {4, 3, 2};
    const std::vector<int64_t> opens = {0, 1, 2};
    // coordinate descent over (blocks per CU, LDS stack depth, pool slots, node width, leaf size,
    // TLAS opening depth), repeated while a round changes the configuration (at most three
    // rounds: a parameter tried early is tried again once the others have moved); a trial
    // replaces the best only when it is 0.5 % faster, so timing noise does not move it
    for (int round = 0; round < 3; round++) {
        const Cfg round_start = best;
        for (int param = 0; param < 6; param++) {
            const std::vector<int64_t>& vals = param == 0   ? bpcs
                                               : param == 1 ? ldss
                                               : param == 2 ? pools
                                               : param == 3 ? widths
                                               : param == 4 ? leaves
                                                            : opens;
            for (int64_t v : vals) {
                Cfg g = best;
                (param == 0   ? g.bpc
                 : param == 1 ? g.lds
                 : param == 2 ? g.pool
                 : param == 3 ? g.width
                 : param == 4 ? g.leaf
                              : g.open) = v;
                if (g == best)
                    continue;
                double ns;
                if (int e = measure(g, ns))
                    return e;
                if (ns < best_ns * 0.995) {
                    best_ns = ns;
                    best = g;
                }
            }
        }
        if (best == round_start)
            break;
    }
    double final_ns;
    if (int v = measure(best, final_ns))
        return v;
    c.sample_index = 0;
    c.stats_fresh = true;
    HIP_TRY(&c, hipMemsetAsync(c.accum, 0, npix * 16, c.stream), "accumulation image");
    HIP_TRY(&c, clear_overflow(c), "error flags");
    HIP_TRY(&c, hipStreamSynchronize(c.stream), "sync");
    if (out) {
        pt_tune_result r{};
        r.blocks_per_cu = uint32_t(best.bpc);
        r.pool_slots = uint32_t(best.pool);
        r.lds_stack = uint32_t(best.lds);
        r.leaf_size = uint32_t(best.leaf);
        r.node_width = uint32_t(best.width);
        r.tlas_open = uint32_t(best.open);
        r.trials = trials;
        r.ns_per_ray_start = base_ns;
        r.ns_per_ray_best = best_ns;
        *out = r;
    }
    return PT_OK;
}

int pt_build_light_tables(pt_ctx* h)
{
    if (!h)
        return PT_EINVAL;
    if (int v = settle(h->c))
        return v;
    Ctx& c = h->c;
    if (!c.light_layout_known)
        return fail(&c, PT_ESTATE, "the light-sampler words do not have the packed layout");
    HIP_TRY(&c, hipStreamSynchronize(c.stream), "sync");
    HIP_TRY(&c, build_light_tables(c), "light tables");
    HIP_TRY(&c, clear_overflow(c), "error flags");
    c.sample_index = 0;
    c.stats_fresh = true;
    return PT_OK;
}

int pt_read_light_words(pt_ctx* h, uint32_t* words, size_t count)
{
    if (!h)
        return PT_EINVAL;
    if (int v = settle(h->c))
        return v;
    Ctx& c = h->c;
    if (!words || count != c.light_word_count)
        return fail(&c, PT_EINVAL, "the buffer must hold the scene's light_word_count words");
    if (count)
        HIP_TRY(&c, hipMemcpyAsync(words, c.d_light, count * 4, hipMemcpyDeviceToHost, c.stream), "readback");
    HIP_TRY(&c, hipStreamSynchronize(c.stream), "readback");
    return PT_OK;
}

int pt_synchronize(pt_ctx* h)
{
    if (!h)
        return PT_EINVAL;
    if (int v = settle_checked(h->c))
        return v;
    HIP_TRY(&h->c, hipStreamSynchronize(h->c.stream), "sync");
    return PT_OK;
}

const char* pt_last_error(const pt_ctx* h) { return h ? h->c.error.c_str() : g_create_error.c_str(); }

void pt_destroy(pt_ctx* h)
{
    if (!h)
        return;
    (void)settle(h->c);
    (void)hipStreamSynchronize(h->c.stream);
    release(h->c);
    delete h;
}

}  // extern "C"
