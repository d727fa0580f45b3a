// Shared device-side types and helpers of the MI355X back end.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pthip.h"

namespace ptk {

// InstanceData, 128 B (src/scene/scenebuffer.h:13-20): head words, M row-major, and the
// normal matrix rows (= columns of M^-1), each padded to 16 B.
struct InstanceData {
    uint32_t vertex_index, index_index, emitter_index, material_index;
    float m[4][4];
    float n[3][4];
};
static_assert(sizeof(InstanceData) == 128, "InstanceData layout");

// PbrMaterial, 96 B (src/scene/material.h:24-29).
struct MaterialRec {
    float4 base_colour, emission, rmnb, specular, sheen, cc;
};
static_assert(sizeof(MaterialRec) == 96, "PbrMaterial layout");

// A 4-wide BVH node: exactly one 128-byte cache line. Child bounds are stored as six
// 4-float rows (struct-of-arrays within the node) so one lane reads a node with eight
// 16-byte loads. child[k]: EMPTY, an internal node index, or a leaf (LEAF bit set):
//   BLAS leaf: bits 0..26 first triangle record, bits 27..30 count-1
//   TLAS leaf: bits 0..30 instance index
struct alignas(128) Node4 {
    float lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4];
    uint32_t child[4];
    uint32_t pad[4];
};
static_assert(sizeof(Node4) == 128, "Node4 layout");

constexpr uint32_t CHILD_EMPTY = 0xFFFFFFFFu;
constexpr uint32_t CHILD_LEAF = 0x80000000u;
constexpr uint32_t LEAF_FIRST_MASK = (1u << 27) - 1u;
constexpr uint32_t LEAF_MAX_TRIS = 16;

// A shadow ray's or cache query's radiance target: an entry of the next path queue (the
// path goes on) or, with this bit, the path's place in the job's radiance output (the path
// has ended).
constexpr uint32_t TARGET_ENDED = 0x80000000u;

// Triangle record, 48 B: object-space corners; v0.w holds the mesh-relative primitive
// index (as raw bits).
struct TriRec {
    float4 v0, v1, v2;
};
static_assert(sizeof(TriRec) == 48, "TriRec layout");

// World->object affine transform of an instance, three rows.
struct InstXform {
    float4 r0, r1, r2;
};

struct DeviceScene {
    const float4* vertices;  // two float4 per vertex (geometry.hlsli:27-36)
    const uint32_t* indices;
    const InstanceData* instances;
    const MaterialRec* materials;
    const uint32_t* light;
    uint32_t light_count;
    uint32_t instance_count;
    const Node4* nodes;
    const TriRec* tris;
    const InstXform* inst_inv;
    const uint2* tlas_entries;        // TLAS leaf k enters instance .x at BLAS node .y (CHILD_EMPTY: none)
    const float* inst_opacity;        // base-colour alpha of transparent instances, 1 = opaque,
                                      // -1 = alpha from the base-colour map at the hit
    uint32_t tlas_root;               // CHILD_EMPTY when the scene has no instances
    float3 envcol;
    const pt_texture_desc* tex;       // textures (pthip.h pt_texture_desc)
    const uint8_t* texels;
    uint32_t tex_count;
};

__host__ __device__ inline uint32_t float_as_u32(float f)
{
    union {
        float f;
        uint32_t u;
    } c;
    c.f = f;
    return c.u;
}

__host__ __device__ inline float u32_as_float(uint32_t u)
{
    union {
        float f;
        uint32_t u;
    } c;
    c.u = u;
    return c.f;
}

// float3 helpers
__device__ __forceinline__ float3 f3(float x, float y, float z) { return make_float3(x, y, z); }
__device__ __forceinline__ float3 operator+(float3 a, float3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 operator-(float3 a, float3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 operator-(float3 a) { return f3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float3 operator*(float3 a, float3 b) { return f3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ float3 operator*(float3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float3 operator*(float s, float3 a) { return f3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float3 operator/(float3 a, float s) { return f3(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 cross3(float3 a, float3 b)
{
    return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float3 normalize3(float3 v) { return v * (1.0f / sqrtf(dot3(v, v))); }
__device__ __forceinline__ float maxc3(float3 v) { return fmaxf(fmaxf(v.x, v.y), v.z); }
__device__ __forceinline__ float3 xyz(float4 v) { return f3(v.x, v.y, v.z); }

}  // namespace ptk
