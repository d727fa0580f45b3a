// Device-built LBVH: Morton codes -> radix sort -> Karras (2012) binary radix tree ->
// bottom-up bounds -> collapse into 4-wide 128-byte nodes. One BLAS per mesh (all meshes in
// one batched build) and a TLAS over the instances.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "common.h"

namespace ptk {

struct BvhBuildParams {
    uint32_t blas_leaf_size = 4;  // max triangles per BLAS leaf (1..16)
};

struct BvhBuffers {
    Node4* nodes = nullptr;         // [blas nodes | tlas nodes]
    TriRec* tris = nullptr;
    InstXform* inst_inv = nullptr;
    uint32_t* inst_blas_root = nullptr;
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
