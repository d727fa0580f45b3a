// Device LBVH builder (see bvh.h).
//
// Per build set (all meshes' triangles at once, or the instances): primitive bounds ->
// 30-bit Morton code of the centroid within its segment's centroid bounds -> one 64-bit
// radix sort of (segment << 32 | morton) -> Karras' binary radix tree per segment ->
// bottom-up bounds with atomic visit counters -> top-down collapse into 4-wide nodes, each
// step opening the child with the largest surface area, subtrees of <= leaf_size
// primitives becoming leaves (contiguous ranges of the sorted order).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "bvh.h"

namespace ptk {

namespace {

constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr uint32_t BREF_LEAF = 0x80000000u;  // binary-tree reference: leaf flag

__device__ __forceinline__ uint32_t ordered(float f)
{
    uint32_t u = float_as_u32(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unordered(uint32_t u)
{
    return u32_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

__device__ __forceinline__ uint32_t expand_bits(uint32_t v)
{
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

struct BuildSet {
    uint32_t n = 0;          // primitives
    uint32_t segs = 0;       // segments
    float4* lo = nullptr;    // prim bounds
    float4* hi = nullptr;
    uint32_t* seg = nullptr;        // segment of each (unsorted) primitive
    const uint32_t* seg_begin = nullptr;  // segs+1 prefix offsets (device)
    uint32_t* cmin = nullptr;  // 3 ordered uints per segment (centroid bounds)
    uint32_t* cmax = nullptr;
    uint64_t* keys = nullptr;
    uint64_t* keys_alt = nullptr;
    uint32_t* vals = nullptr;
    uint32_t* vals_alt = nullptr;
    // binary tree, indexed by global sorted position
    uint32_t* left = nullptr;
    uint32_t* right = nullptr;
    uint32_t* first = nullptr;
    uint32_t* last = nullptr;
    uint32_t* parent = nullptr;       // of internal nodes
    uint32_t* leaf_parent = nullptr;  // of leaves
    uint32_t* visit = nullptr;
    float* ibox = nullptr;  // 6 floats per internal node
    float* lbox = nullptr;  // 6 floats per leaf
};

__global__ void k_centroid_bounds(uint32_t n, const float4* lo, const float4* hi, const uint32_t* seg,
                                  uint32_t* cmin, uint32_t* cmax)
{
    // Whole waves iterate together so that a wave whose 64 primitives share one segment
    // reduces in registers and issues 6 atomics instead of 384.
    const uint32_t rounded = (n + 63u) & ~63u;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < rounded; i += gridDim.x * blockDim.x) {
        const bool valid = i < n;
        uint32_t s = 0xFFFFFFFFu;
        uint32_t mn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, mx[3] = {0u, 0u, 0u};
        if (valid) {
            const float4 a = lo[i], b = hi[i];
            s = seg[i];
            const float c[3] = {0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z)};
            for (int k = 0; k < 3; k++)
                mn[k] = mx[k] = ordered(c[k]);
        }
        const uint32_t s0 = __shfl(s, 0);
        const bool uniform = __all(!valid || s == s0) && s0 != 0xFFFFFFFFu;
        if (uniform) {
            for (int off = 32; off >= 1; off >>= 1)
                for (int k = 0; k < 3; k++) {
                    mn[k] = min(mn[k], (uint32_t)__shfl_xor(mn[k], off));
                    mx[k] = max(mx[k], (uint32_t)__shfl_xor(mx[k], off));
                }
            if (__lane_id() == 0)
                for (int k = 0; k < 3; k++) {
                    atomicMin(&cmin[s0 * 3 + k], mn[k]);
                    atomicMax(&cmax[s0 * 3 + k], mx[k]);
                }
        } else if (valid) {
            for (int k = 0; k < 3; k++) {
                atomicMin(&cmin[s * 3 + k], mn[k]);
                atomicMax(&cmax[s * 3 + k], mx[k]);
            }
        }
    }
}

__global__ void k_morton(uint32_t n, const float4* lo, const float4* hi, const uint32_t* seg, const uint32_t* cmin,
                         const uint32_t* cmax, uint64_t* keys, uint32_t* vals)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4 a = lo[i], b = hi[i];
        const uint32_t s = seg[i];
        const float c[3] = {0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z)};
        uint32_t q[3];
        for (int k = 0; k < 3; k++) {
            const float mn = unordered(cmin[s * 3 + k]), mx = unordered(cmax[s * 3 + k]);
            const float ext = mx - mn;
            float t = ext > 0.0f ? (c[k] - mn) / ext : 0.5f;
            t = fminf(fmaxf(t, 0.0f), 1.0f);
            q[k] = min(uint32_t(t * 1024.0f), 1023u);
        }
        const uint32_t code = (expand_bits(q[0]) << 2) | (expand_bits(q[1]) << 1) | expand_bits(q[2]);
        keys[i] = (uint64_t(s) << 32) | code;
        vals[i] = i;
    }
}

__device__ __forceinline__ int delta(const uint64_t* keys, uint32_t base, int n, int a, int b)
{
    if (b < 0 || b >= n)
        return -1;
    const uint32_t ka = uint32_t(keys[base + a]), kb = uint32_t(keys[base + b]);
    if (ka == kb)
        return 32 + __clz(uint32_t(a) ^ uint32_t(b));
    return __clz(ka ^ kb);
}

__global__ void k_karras(uint32_t n, const uint64_t* keys, const uint32_t* seg_begin, uint32_t* left, uint32_t* right,
                         uint32_t* first, uint32_t* last, uint32_t* parent, uint32_t* leaf_parent)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t s = uint32_t(keys[i] >> 32);
        const uint32_t base = seg_begin[s];
        const int nm = int(seg_begin[s + 1] - base);
        const int j = int(i - base);
        if (nm == 1) {
            leaf_parent[i] = NONE;
            continue;
        }
        if (j == 0)
            parent[i] = NONE;  // segment root
        if (j >= nm - 1)
            continue;
        const int d = (delta(keys, base, nm, j, j + 1) - delta(keys, base, nm, j, j - 1)) >= 0 ? 1 : -1;
        const int dmin = delta(keys, base, nm, j, j - d);
        int lmax = 2;
        while (delta(keys, base, nm, j, j + lmax * d) > dmin)
            lmax *= 2;
        int l = 0;
        for (int t = lmax / 2; t >= 1; t /= 2)
            if (delta(keys, base, nm, j, j + (l + t) * d) > dmin)
                l += t;
        const int jj = j + l * d;
        const int dnode = delta(keys, base, nm, j, jj);
        int sp = 0;
        int t = l;
        do {
            t = (t + 1) / 2;
            if (delta(keys, base, nm, j, j + (sp + t) * d) > dnode)
                sp += t;
        } while (t > 1);
        const int gamma = j + sp * d + min(d, 0);
        const int lo_i = min(j, jj), hi_i = max(j, jj);
        const uint32_t g = base + uint32_t(j);
        const uint32_t lc = base + uint32_t(gamma), rc = base + uint32_t(gamma + 1);
        if (lo_i == gamma) {
            left[g] = lc | BREF_LEAF;
            leaf_parent[lc] = g;
        } else {
            left[g] = lc;
            parent[lc] = g;
        }
        if (hi_i == gamma + 1) {
            right[g] = rc | BREF_LEAF;
            leaf_parent[rc] = g;
        } else {
            right[g] = rc;
            parent[rc] = g;
        }
        first[g] = base + uint32_t(lo_i);
        last[g] = base + uint32_t(hi_i);
    }
}

__device__ __forceinline__ void load_box(const float* b, float* out)
{
    for (int k = 0; k < 6; k++)
        out[k] = __hip_atomic_load(&b[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_box(float* b, const float* v)
{
    for (int k = 0; k < 6; k++)
        __hip_atomic_store(&b[k], v[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_bounds_up(uint32_t n, const uint32_t* vals, const float4* lo, const float4* hi,
                            const uint32_t* left, const uint32_t* right, const uint32_t* parent,
                            const uint32_t* leaf_parent, uint32_t* visit, float* ibox, float* lbox)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t p0 = vals[i];
        const float4 a = lo[p0], b = hi[p0];
        const float box[6] = {a.x, a.y, a.z, b.x, b.y, b.z};
        store_box(&lbox[size_t(i) * 6], box);
        __threadfence();
        uint32_t p = leaf_parent[i];
        while (p != NONE) {
            const uint32_t old = atomicAdd(&visit[p], 1u);
            if (old == 0)
                break;  // the sibling subtree finishes this node
            __threadfence();
            float bl[6], br[6];
            const uint32_t l = left[p], r = right[p];
            load_box((l & BREF_LEAF) ? &lbox[size_t(l & ~BREF_LEAF) * 6] : &ibox[size_t(l) * 6], bl);
            load_box((r & BREF_LEAF) ? &lbox[size_t(r & ~BREF_LEAF) * 6] : &ibox[size_t(r) * 6], br);
            float u[6];
            for (int k = 0; k < 3; k++) {
                u[k] = fminf(bl[k], br[k]);
                u[k + 3] = fmaxf(bl[k + 3], br[k + 3]);
            }
            store_box(&ibox[size_t(p) * 6], u);
            __threadfence();
            p = parent[p];
        }
    }
}

struct CollapseItem {
    uint32_t bref;  // binary reference (leaf flag) whose subtree becomes this node
    uint32_t slot;  // output node index
};

__device__ __forceinline__ float box_area(const float* b)
{
    const float dx = fmaxf(b[3] - b[0], 0.0f), dy = fmaxf(b[4] - b[1], 0.0f), dz = fmaxf(b[5] - b[2], 0.0f);
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

// tlas != 0: leaves hold one instance each (value = original primitive = instance index).
__global__ void k_collapse(const CollapseItem* in, const uint32_t* in_count, CollapseItem* out, uint32_t* out_count,
                           uint32_t* node_counter, uint32_t node_capacity, Node4* nodes, const uint32_t* left,
                           const uint32_t* right, const uint32_t* first, const uint32_t* last, const float* ibox,
                           const float* lbox, const uint32_t* vals, uint32_t leaf_size, int tlas)
{
    const uint32_t cnt = *in_count;
    for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < cnt; w += gridDim.x * blockDim.x) {
        const CollapseItem it = in[w];
        uint32_t c[4];
        int nc;
        if (it.bref & BREF_LEAF) {  // a single-primitive set
            c[0] = it.bref;
            nc = 1;
        } else {
            c[0] = left[it.bref];
            c[1] = right[it.bref];
            nc = 2;
        }
        auto size_of = [&](uint32_t r) -> uint32_t {
            return (r & BREF_LEAF) ? 1u : (last[r] - first[r] + 1u);
        };
        while (nc < 4) {
            int best = -1;
            float best_area = -1.0f;
            for (int k = 0; k < nc; k++) {
                if ((c[k] & BREF_LEAF) || size_of(c[k]) <= leaf_size)
                    continue;
                const float a = box_area(&ibox[size_t(c[k]) * 6]);
                if (a > best_area) {
                    best_area = a;
                    best = k;
                }
            }
            if (best < 0)
                break;
            const uint32_t r = c[best];
            c[best] = left[r];
            c[nc++] = right[r];
        }
        Node4 node;
        for (int k = 0; k < 4; k++) {
            if (k >= nc) {
                node.lo_x[k] = node.lo_y[k] = node.lo_z[k] = INFINITY;
                node.hi_x[k] = node.hi_y[k] = node.hi_z[k] = -INFINITY;
                node.child[k] = CHILD_EMPTY;
                continue;
            }
            const uint32_t r = c[k];
            const float* b = (r & BREF_LEAF) ? &lbox[size_t(r & ~BREF_LEAF) * 6] : &ibox[size_t(r) * 6];
            node.lo_x[k] = b[0];
            node.lo_y[k] = b[1];
            node.lo_z[k] = b[2];
            node.hi_x[k] = b[3];
            node.hi_y[k] = b[4];
            node.hi_z[k] = b[5];
            const uint32_t sz = size_of(r);
            if ((r & BREF_LEAF) || sz <= leaf_size) {
                const uint32_t f = (r & BREF_LEAF) ? (r & ~BREF_LEAF) : first[r];
                if (tlas)
                    node.child[k] = CHILD_LEAF | vals[f];
                else
                    node.child[k] = CHILD_LEAF | ((sz - 1u) << 27) | f;
            } else {
                const uint32_t slot = atomicAdd(node_counter, 1u);
                if (slot < node_capacity) {
                    const uint32_t q = atomicAdd(out_count, 1u);
                    out[q] = CollapseItem{r, slot};
                    node.child[k] = slot;
                } else {
                    node.child[k] = CHILD_EMPTY;  // capacity guard; the host reports the overflow
                }
            }
        }
        node.pad[0] = node.pad[1] = node.pad[2] = node.pad[3] = 0;
        nodes[it.slot] = node;
    }
}

__global__ void k_tri_bounds(uint32_t n, const float4* vertices, const uint32_t* indices, const pt_mesh_desc* meshes,
                             const uint32_t* tri_begin, uint32_t mesh_count, float4* lo, float4* hi, uint32_t* seg)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t a = 0, b = mesh_count;  // find m with tri_begin[m] <= i < tri_begin[m+1]
        while (b - a > 1) {
            const uint32_t mid = (a + b) / 2;
            if (tri_begin[mid] <= i)
                a = mid;
            else
                b = mid;
        }
        const pt_mesh_desc md = meshes[a];
        const uint32_t t = i - tri_begin[a];
        float3 mn = f3(INFINITY, INFINITY, INFINITY), mx = f3(-INFINITY, -INFINITY, -INFINITY);
        for (int k = 0; k < 3; k++) {
            const float4 p = vertices[size_t(md.vertex_start + indices[md.index_start + t * 3 + k]) * 2];
            mn = f3(fminf(mn.x, p.x), fminf(mn.y, p.y), fminf(mn.z, p.z));
            mx = f3(fmaxf(mx.x, p.x), fmaxf(mx.y, p.y), fmaxf(mx.z, p.z));
        }
        lo[i] = make_float4(mn.x, mn.y, mn.z, 0.0f);
        hi[i] = make_float4(mx.x, mx.y, mx.z, 0.0f);
        seg[i] = a;
    }
}

__global__ void k_tri_records(uint32_t n, const uint32_t* vals, const uint32_t* seg, const float4* vertices,
                              const uint32_t* indices, const pt_mesh_desc* meshes, const uint32_t* tri_begin,
                              TriRec* tris)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t p = vals[i];
        const uint32_t m = seg[p];
        const pt_mesh_desc md = meshes[m];
        const uint32_t t = p - tri_begin[m];
        float4 v[3];
        for (int k = 0; k < 3; k++)
            v[k] = vertices[size_t(md.vertex_start + indices[md.index_start + t * 3 + k]) * 2];
        TriRec r;
        r.v0 = make_float4(v[0].x, v[0].y, v[0].z, u32_as_float(t));
        r.v1 = make_float4(v[1].x, v[1].y, v[1].z, 0.0f);
        r.v2 = make_float4(v[2].x, v[2].y, v[2].z, 0.0f);
        tris[i] = r;
    }
}

// World bounds of each instance from its BLAS root box (8 transformed corners).
__global__ void k_instance_bounds(uint32_t n, const InstanceData* inst, const uint32_t* inst_mesh,
                                  const uint32_t* tri_begin, const float* ibox, const float* lbox,
                                  const uint32_t* mesh_tris, float4* lo, float4* hi, uint32_t* seg)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t m = inst_mesh[i];
        const uint32_t nt = mesh_tris[m];
        seg[i] = 0;
        if (nt == 0) {  // empty mesh: a degenerate box far from everything is never entered
            lo[i] = make_float4(INFINITY, INFINITY, INFINITY, 0.0f);
            hi[i] = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
            continue;
        }
        const uint32_t root = tri_begin[m];
        const float* b = nt == 1 ? &lbox[size_t(root) * 6] : &ibox[size_t(root) * 6];
        const InstanceData& d = inst[i];
        float3 mn = f3(INFINITY, INFINITY, INFINITY), mx = f3(-INFINITY, -INFINITY, -INFINITY);
        for (int c = 0; c < 8; c++) {
            const float px = (c & 1) ? b[3] : b[0], py = (c & 2) ? b[4] : b[1], pz = (c & 4) ? b[5] : b[2];
            const float wx = d.m[0][0] * px + d.m[0][1] * py + d.m[0][2] * pz + d.m[0][3];
            const float wy = d.m[1][0] * px + d.m[1][1] * py + d.m[1][2] * pz + d.m[1][3];
            const float wz = d.m[2][0] * px + d.m[2][1] * py + d.m[2][2] * pz + d.m[2][3];
            mn = f3(fminf(mn.x, wx), fminf(mn.y, wy), fminf(mn.z, wz));
            mx = f3(fmaxf(mx.x, wx), fmaxf(mx.y, wy), fmaxf(mx.z, wz));
        }
        // pad for the rounding of the transform
        const float3 ext = f3(fabsf(mn.x) + fabsf(mx.x), fabsf(mn.y) + fabsf(mx.y), fabsf(mn.z) + fabsf(mx.z));
        const float eps = 1e-6f;
        lo[i] = make_float4(mn.x - eps * ext.x, mn.y - eps * ext.y, mn.z - eps * ext.z, 0.0f);
        hi[i] = make_float4(mx.x + eps * ext.x, mx.y + eps * ext.y, mx.z + eps * ext.z, 0.0f);
    }
}

inline unsigned grid_for(uint32_t n, unsigned block = 256)
{
    const uint64_t g = (uint64_t(n) + block - 1) / block;
    return unsigned(std::max<uint64_t>(1, std::min<uint64_t>(g, 65535u * 4)));
}

#define CHECK(x)                        \
    do {                                \
        hipError_t e_ = (x);            \
        if (e_ != hipSuccess)           \
            return e_;                  \
    } while (0)

struct Scratch {
    std::vector<void*> ptrs;
    ~Scratch()
    {
        for (void* p : ptrs)
            (void)hipFree(p);
    }
    template <typename T>
    hipError_t alloc(T*& p, size_t count)
    {
        void* v = nullptr;
        hipError_t e = hipMalloc(&v, std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess)
            ptrs.push_back(v);
        p = (T*)v;
        return e;
    }
};

// Builds the radix tree and collapses it into nodes[slot0 + s] (root of segment s) and
// further slots from *node_counter (device). Prim bounds / seg must be filled.
hipError_t build_set(hipStream_t st, Scratch& sc, BuildSet& bs, const std::vector<uint32_t>& seg_begin_host,
                     uint32_t leaf_size, int tlas, Node4* nodes, uint32_t slot0, uint32_t* d_node_counter,
                     uint32_t node_capacity)
{
    const uint32_t n = bs.n;
    const unsigned B = 256;
    CHECK(sc.alloc(bs.cmin, size_t(bs.segs) * 3));
    CHECK(sc.alloc(bs.cmax, size_t(bs.segs) * 3));
    CHECK(hipMemsetAsync(bs.cmin, 0xFF, size_t(bs.segs) * 12, st));
    CHECK(hipMemsetAsync(bs.cmax, 0x00, size_t(bs.segs) * 12, st));
    hipLaunchKernelGGL(k_centroid_bounds, dim3(grid_for(n)), dim3(B), 0, st, n, bs.lo, bs.hi, bs.seg, bs.cmin, bs.cmax);
    CHECK(sc.alloc(bs.keys, n));
    CHECK(sc.alloc(bs.keys_alt, n));
    CHECK(sc.alloc(bs.vals, n));
    CHECK(sc.alloc(bs.vals_alt, n));
    hipLaunchKernelGGL(k_morton, dim3(grid_for(n)), dim3(B), 0, st, n, bs.lo, bs.hi, bs.seg, bs.cmin, bs.cmax,
                       bs.keys, bs.vals);
    int end_bit = 32;
    while ((1u << (end_bit - 32)) < bs.segs && end_bit < 64)
        end_bit++;
    hipcub::DoubleBuffer<uint64_t> dk(bs.keys, bs.keys_alt);
    hipcub::DoubleBuffer<uint32_t> dv(bs.vals, bs.vals_alt);
    size_t temp_bytes = 0;
    CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, dk, dv, int(n), 0, end_bit, st));
    void* temp = nullptr;
    CHECK(sc.alloc(*(uint8_t**)&temp, temp_bytes));
    CHECK(hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, dk, dv, int(n), 0, end_bit, st));
    uint64_t* keys = dk.Current();
    uint32_t* vals = dv.Current();

    CHECK(sc.alloc(bs.left, n));
    CHECK(sc.alloc(bs.right, n));
    CHECK(sc.alloc(bs.first, n));
    CHECK(sc.alloc(bs.last, n));
    CHECK(sc.alloc(bs.parent, n));
    CHECK(sc.alloc(bs.leaf_parent, n));
    CHECK(sc.alloc(bs.visit, n));
    CHECK(sc.alloc(bs.ibox, size_t(n) * 6));
    CHECK(sc.alloc(bs.lbox, size_t(n) * 6));
    CHECK(hipMemsetAsync(bs.visit, 0, size_t(n) * 4, st));
    hipLaunchKernelGGL(k_karras, dim3(grid_for(n)), dim3(B), 0, st, n, keys, bs.seg_begin, bs.left, bs.right, bs.first,
                       bs.last, bs.parent, bs.leaf_parent);
    hipLaunchKernelGGL(k_bounds_up, dim3(grid_for(n)), dim3(B), 0, st, n, vals, bs.lo, bs.hi, bs.left, bs.right,
                       bs.parent, bs.leaf_parent, bs.visit, bs.ibox, bs.lbox);
    bs.vals = vals;
    bs.keys = keys;

    // Top-down collapse, one launch per level.
    std::vector<CollapseItem> roots;
    for (uint32_t s = 0; s < bs.segs; s++) {
        const uint32_t b = seg_begin_host[s], e = seg_begin_host[s + 1];
        if (e == b)
            continue;
        roots.push_back(CollapseItem{(e - b == 1) ? (b | BREF_LEAF) : b, slot0 + s});
    }
    CollapseItem *qa, *qb;
    uint32_t* counts;  // [count_a, count_b]
    CHECK(sc.alloc(qa, std::max<size_t>(n, roots.size())));
    CHECK(sc.alloc(qb, std::max<size_t>(n, roots.size())));
    CHECK(sc.alloc(counts, 2));
    if (roots.empty())
        return hipSuccess;
    CHECK(hipMemcpyAsync(qa, roots.data(), roots.size() * sizeof(CollapseItem), hipMemcpyHostToDevice, st));
    uint32_t host_count = uint32_t(roots.size());
    CHECK(hipMemcpyAsync(&counts[0], &host_count, 4, hipMemcpyHostToDevice, st));
    int cur = 0;
    for (int level = 0; level < 128 && host_count > 0; level++) {
        CollapseItem* qin = cur ? qb : qa;
        CollapseItem* qout = cur ? qa : qb;
        CHECK(hipMemsetAsync(&counts[1 - cur], 0, 4, st));
        hipLaunchKernelGGL(k_collapse, dim3(grid_for(host_count)), dim3(B), 0, st, qin, &counts[cur], qout,
                           &counts[1 - cur], d_node_counter, node_capacity, nodes, bs.left, bs.right, bs.first,
                           bs.last, bs.ibox, bs.lbox, vals, leaf_size, tlas);
        CHECK(hipGetLastError());
        CHECK(hipMemcpyAsync(&host_count, &counts[1 - cur], 4, hipMemcpyDeviceToHost, st));
        CHECK(hipStreamSynchronize(st));
        cur = 1 - cur;
    }
    if (host_count > 0)
        return hipErrorUnknown;  // tree deeper than the level cap (cannot happen for 64-bit keys)
    return hipSuccess;
}

}  // namespace

void free_bvh(BvhBuffers& b)
{
    (void)hipFree(b.nodes);
    (void)hipFree(b.tris);
    (void)hipFree(b.inst_inv);
    (void)hipFree(b.inst_blas_root);
    b = BvhBuffers{};
}

hipError_t build_bvh(hipStream_t st, const float4* d_vertices, const uint32_t* d_indices, const pt_mesh_desc* d_meshes,
                     const std::vector<pt_mesh_desc>& meshes, const InstanceData* d_instances,
                     const uint32_t* d_instance_mesh, uint32_t instance_count, const MaterialRec* d_materials,
                     const BvhBuildParams& params, BvhBuffers& out)
{
    (void)d_materials;
    Scratch sc;
    const uint32_t M = uint32_t(meshes.size());
    std::vector<uint32_t> tri_begin(M + 1, 0), mesh_tris(std::max<uint32_t>(M, 1), 0);
    for (uint32_t m = 0; m < M; m++) {
        tri_begin[m + 1] = tri_begin[m] + meshes[m].tri_count;
        mesh_tris[m] = meshes[m].tri_count;
    }
    const uint32_t T = tri_begin[M];
    const uint32_t leaf = std::min<uint32_t>(std::max<uint32_t>(params.blas_leaf_size, 1), LEAF_MAX_TRIS);

    // Node budget: BLAS roots + at most one node per binary internal node; TLAS likewise.
    const uint32_t blas_cap = M + T + 1;
    const uint32_t tlas_cap = instance_count + 2;
    out.node_capacity = blas_cap + tlas_cap;
    CHECK(hipMalloc(&out.nodes, size_t(out.node_capacity) * sizeof(Node4)));
    CHECK(hipMalloc(&out.tris, std::max<size_t>(T, 1) * sizeof(TriRec)));
    out.tri_count = T;

    uint32_t *d_tri_begin, *d_mesh_tris, *d_counter;
    CHECK(sc.alloc(d_tri_begin, M + 1));
    CHECK(sc.alloc(d_mesh_tris, std::max<uint32_t>(M, 1)));
    CHECK(sc.alloc(d_counter, 1));
    CHECK(hipMemcpyAsync(d_tri_begin, tri_begin.data(), (M + 1) * 4, hipMemcpyHostToDevice, st));
    CHECK(hipMemcpyAsync(d_mesh_tris, mesh_tris.data(), mesh_tris.size() * 4, hipMemcpyHostToDevice, st));
    // Every mesh's root slot starts empty (meshes without triangles keep it so).
    {
        std::vector<Node4> empty(std::max<uint32_t>(M, 1));
        for (Node4& nd : empty) {
            for (int k = 0; k < 4; k++) {
                nd.lo_x[k] = nd.lo_y[k] = nd.lo_z[k] = INFINITY;
                nd.hi_x[k] = nd.hi_y[k] = nd.hi_z[k] = -INFINITY;
                nd.child[k] = CHILD_EMPTY;
                nd.pad[k] = 0;
            }
        }
        if (M)
            CHECK(hipMemcpyAsync(out.nodes, empty.data(), size_t(M) * sizeof(Node4), hipMemcpyHostToDevice, st));
        CHECK(hipStreamSynchronize(st));
    }
    uint32_t counter = M;
    CHECK(hipMemcpyAsync(d_counter, &counter, 4, hipMemcpyHostToDevice, st));

    BuildSet blas;
    float* blas_ibox = nullptr;
    float* blas_lbox = nullptr;
    if (T > 0) {
        blas.n = T;
        blas.segs = M;
        blas.seg_begin = d_tri_begin;
        CHECK(sc.alloc(blas.lo, T));
        CHECK(sc.alloc(blas.hi, T));
        CHECK(sc.alloc(blas.seg, T));
        hipLaunchKernelGGL(k_tri_bounds, dim3(grid_for(T)), dim3(256), 0, st, T, d_vertices, d_indices, d_meshes,
                           d_tri_begin, M, blas.lo, blas.hi, blas.seg);
        CHECK(build_set(st, sc, blas, tri_begin, leaf, 0, out.nodes, 0, d_counter, blas_cap));
        hipLaunchKernelGGL(k_tri_records, dim3(grid_for(T)), dim3(256), 0, st, T, blas.vals, blas.seg, d_vertices,
                           d_indices, d_meshes, d_tri_begin, out.tris);
        blas_ibox = blas.ibox;
        blas_lbox = blas.lbox;
    }
    CHECK(hipMemcpyAsync(&counter, d_counter, 4, hipMemcpyDeviceToHost, st));
    CHECK(hipStreamSynchronize(st));
    if (counter > blas_cap)
        return hipErrorOutOfMemory;
    const uint32_t blas_nodes = counter;

    // Per-instance data: world->object transforms (fp64 inverse on the host), BLAS roots,
    // opacity of transparent instances.
    std::vector<InstanceData> inst(instance_count);
    std::vector<uint32_t> inst_mesh(instance_count);
    if (instance_count) {
        CHECK(hipMemcpyAsync(inst.data(), d_instances, size_t(instance_count) * sizeof(InstanceData),
                             hipMemcpyDeviceToHost, st));
        CHECK(hipMemcpyAsync(inst_mesh.data(), d_instance_mesh, size_t(instance_count) * 4, hipMemcpyDeviceToHost,
                             st));
        CHECK(hipStreamSynchronize(st));
    }
    std::vector<InstXform> xf(std::max<uint32_t>(instance_count, 1));
    std::vector<uint32_t> roots(std::max<uint32_t>(instance_count, 1), CHILD_EMPTY);
    for (uint32_t i = 0; i < instance_count; i++) {
        const InstanceData& d = inst[i];
        double a[3][3], t[3];
        for (int r = 0; r < 3; r++) {
            for (int c = 0; c < 3; c++)
                a[r][c] = d.m[r][c];
            t[r] = d.m[r][3];
        }
        const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                           a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                           a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
        double iv[3][3];
        iv[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / det;
        iv[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det;
        iv[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det;
        iv[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) / det;
        iv[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det;
        iv[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
        iv[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / det;
        iv[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
        iv[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
        float4 rows[3];
        for (int r = 0; r < 3; r++) {
            const double tr = -(iv[r][0] * t[0] + iv[r][1] * t[1] + iv[r][2] * t[2]);
            rows[r] = make_float4(float(iv[r][0]), float(iv[r][1]), float(iv[r][2]), float(tr));
        }
        xf[i] = InstXform{rows[0], rows[1], rows[2]};
        const uint32_t m = inst_mesh[i];
        if (m >= M)
            return hipErrorInvalidValue;
        roots[i] = meshes[m].tri_count ? m : CHILD_EMPTY;
    }
    CHECK(hipMalloc(&out.inst_inv, xf.size() * sizeof(InstXform)));
    CHECK(hipMalloc(&out.inst_blas_root, roots.size() * 4));
    CHECK(hipMemcpyAsync(out.inst_inv, xf.data(), xf.size() * sizeof(InstXform), hipMemcpyHostToDevice, st));
    CHECK(hipMemcpyAsync(out.inst_blas_root, roots.data(), roots.size() * 4, hipMemcpyHostToDevice, st));

    // TLAS over instance world bounds.
    out.tlas_root = CHILD_EMPTY;
    if (instance_count > 0) {
        BuildSet tl;
        tl.n = instance_count;
        tl.segs = 1;
        std::vector<uint32_t> sb = {0, instance_count};
        uint32_t* d_sb;
        CHECK(sc.alloc(d_sb, 2));
        CHECK(hipMemcpyAsync(d_sb, sb.data(), 8, hipMemcpyHostToDevice, st));
        tl.seg_begin = d_sb;
        CHECK(sc.alloc(tl.lo, instance_count));
        CHECK(sc.alloc(tl.hi, instance_count));
        CHECK(sc.alloc(tl.seg, instance_count));
        float* dummy_box = nullptr;
        CHECK(sc.alloc(dummy_box, 6));
        hipLaunchKernelGGL(k_instance_bounds, dim3(grid_for(instance_count)), dim3(256), 0, st, instance_count,
                           d_instances, d_instance_mesh, d_tri_begin, blas_ibox ? blas_ibox : dummy_box,
                           blas_lbox ? blas_lbox : dummy_box, d_mesh_tris, tl.lo, tl.hi, tl.seg);
        const uint32_t tlas_slot = blas_nodes;
        counter = blas_nodes + 1;
        CHECK(hipMemcpyAsync(d_counter, &counter, 4, hipMemcpyHostToDevice, st));
        CHECK(build_set(st, sc, tl, sb, 1, 1, out.nodes, tlas_slot, d_counter, out.node_capacity));
        CHECK(hipMemcpyAsync(&counter, d_counter, 4, hipMemcpyDeviceToHost, st));
        CHECK(hipStreamSynchronize(st));
        if (counter > out.node_capacity)
            return hipErrorOutOfMemory;
        out.tlas_root = tlas_slot;
    }
    out.node_count = counter;
    CHECK(hipStreamSynchronize(st));
    return hipSuccess;
}

}  // namespace ptk
