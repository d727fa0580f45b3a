// Device-built BVH: Morton codes -> radix sort -> binary tree (PLOC clustering, or Karras'
// (2012) radix tree) -> SAH-optimal collapse into 4-wide 128-byte nodes. One BLAS per mesh
// (all meshes in one batched build) and a TLAS over the instances (or over subtrees of their
// BLASes: tlas_open).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "common.h"

namespace ptk {

constexpr uint32_t BVH_BUILDER_PLOC = 0;  // parallel locally-ordered clustering (Meister & Bittner 2018)
constexpr uint32_t BVH_BUILDER_LBVH = 1;  // Karras (2012) binary radix tree over the Morton codes
constexpr uint32_t BVH_TLAS_DEVICE = 0;
constexpr uint32_t BVH_TLAS_HOST = 1;
constexpr uint32_t HOST_TLAS_MAX = 16384;

struct BvhBuildParams {
    uint32_t blas_leaf_size = 4;  // max triangles per BLAS leaf (1..16)
    uint32_t node_width = 4;      // children per node (2..4) the SAH collapse may use
    uint32_t builder = BVH_BUILDER_PLOC;
    // Levels of each instance's BLAS the TLAS may open (0..2): an instance is then entered
    // through the subtrees of its BLAS whose transformed boxes are tighter than its own (partial
    // re-braiding, Benthin et al. 2017); a node opens when all its children are internal nodes
    // and their world boxes' surface areas sum to less than tlas_open_ratio x its own.
    uint32_t tlas_open = 0;
    float tlas_open_ratio = 1.5f;
    // TLAS entry boxes: 0 = the 8 transformed corners of the entry node's box; 1 or 2 = the
    // union of the transformed boxes of its descendants that many levels down (tighter under
    // rotation); 3 = an instance entered at its BLAS root gets the box of its transformed
    // vertices, opened entries that of level 2
    uint32_t tlas_box_levels = 0;
    // TLAS builder: BVH_TLAS_DEVICE (the BLAS builder's) or BVH_TLAS_HOST (full-sweep SAH and
    // SAH-optimal collapse on the host, for at most HOST_TLAS_MAX entries; more fall back to
    // the device builder)
    uint32_t tlas_builder = BVH_TLAS_DEVICE;
    // SAH cost of entering a node, testing a triangle, entering an instance
    float sah_node_cost = 1.0f;
    float sah_tri_cost = 0.5f;
    float sah_instance_cost = 1.0f;
};

struct BvhBuffers {
    Node4* nodes = nullptr;         // [blas nodes | tlas nodes]
    TriRec* tris = nullptr;
    InstXform* inst_inv = nullptr;
    uint2* tlas_entries = nullptr;  // what a TLAS leaf enters: (instance, BLAS node or CHILD_EMPTY)
    uint32_t entry_count = 0;
    uint32_t node_count = 0;
    uint32_t node_capacity = 0;
    uint32_t tri_count = 0;
    uint32_t tlas_root = CHILD_EMPTY;
};

// Inputs are device pointers: vertices (2 float4 per vertex), indices (mesh-relative),
// per-mesh descriptors, per-instance records and mesh map. Host copies of the mesh table
// and instance data are also passed (small) for sizing. Returns hipSuccess or an error.
hipError_t build_bvh(hipStream_t stream, const float4* d_vertices, const uint32_t* d_indices,
                     const pt_mesh_desc* d_meshes, const std::vector<pt_mesh_desc>& meshes,
                     const InstanceData* d_instances, const uint32_t* d_instance_mesh, uint32_t instance_count,
                     const MaterialRec* d_materials, const BvhBuildParams& params, BvhBuffers& out);

void free_bvh(BvhBuffers& b);

}  // namespace ptk
