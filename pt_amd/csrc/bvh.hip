// Device BVH builder (see bvh.h).
//
// Per build set (all meshes' triangles at once, or the instances): primitive bounds ->
// 30-bit Morton code of the centroid within its segment's centroid bounds -> one 64-bit
// radix sort of (segment << 32 | morton) -> a binary tree per segment, built either by
// PLOC (parallel locally-ordered clustering over the Morton order, the default) or as
// Karras' binary radix tree -> SAH cost of every binary node as a subtree of a W-wide tree
// (W = 2..4 children) -> top-down collapse into 4-wide 128-byte nodes following the
// SAH decisions; a triangle leaf is a contiguous range of the triangle records, which are
// written in leaf order.
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
    // Binary tree. A reference with BREF_LEAF is a sorted position (one primitive), any
    // other an internal node.
    uint32_t* left = nullptr;
    uint32_t* right = nullptr;
    uint32_t* cnt = nullptr;    // primitives under an internal node
    uint32_t* dec = nullptr;    // collapse decisions of an internal node (sah_node)
    float4* cost = nullptr;     // SAH cost of an internal node with 1..4 slots
    float* ibox = nullptr;      // 6 floats per internal node
    float* lbox = nullptr;      // 6 floats per leaf
    uint32_t* root = nullptr;   // per segment: reference of its root (NONE: no primitives)
    float* segbox = nullptr;    // per segment: 6 floats
    uint32_t* perm = nullptr;   // sorted position of each output position (leaf order)
};

// SAH cost model of the collapse into wide nodes: entering a node costs c_node, testing a
// primitive (a triangle, or entering an instance) c_leaf, each weighted by the surface area
// of the box through which a ray reaches it.
struct Sah {
    uint32_t width;     // children per wide node (2..4)
    uint32_t max_leaf;  // primitives per leaf (1..16)
    float c_node;
    float c_leaf;
};

__device__ __forceinline__ float half_area(const float* b)
{
    const float dx = fmaxf(b[3] - b[0], 0.0f), dy = fmaxf(b[4] - b[1], 0.0f), dz = fmaxf(b[5] - b[2], 0.0f);
    return dx * dy + dy * dz + dz * dx;
}

__device__ __forceinline__ void box_union(const float* a, const float* b, float* u)
{
    for (int k = 0; k < 3; k++) {
        u[k] = fminf(a[k], b[k]);
        u[k + 3] = fmaxf(a[k + 3], b[k + 3]);
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

// Cost of a child subtree when it may take 1..4 slots of its parent wide node.
__device__ __forceinline__ void child_cost(const Sah& sah, uint32_t ref, const float* box, const float4* cost,
                                           float* c)
{
    if (ref & BREF_LEAF) {
        const float v = sah.c_leaf * half_area(box);
        c[0] = c[1] = c[2] = c[3] = v;
        return;
    }
    const float* q = reinterpret_cast<const float*>(&cost[ref]);
    for (int k = 0; k < 4; k++)
        c[k] = __hip_atomic_load(&q[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The dynamic programme of Ylitie, Karras and Laine (HPG 2017) at one binary node with
// box `box`, `count` primitives and children costs cl, cr:
//   dist(i)  = min over kl of cl(kl) + cr(i - kl)          (i slots shared by the children)
//   C(n, 1)  = min(leaf: A c_leaf count, wide node: A c_node + dist(W))
//   C(n, i)  = min(C(n, i - 1), dist(i))                    (i = 2..W)
// dec bit 0: the node as one slot is a leaf; bits 4(i-1)..4(i-1)+3 (i = 2..4): how its
// subtree fills i slots (left slots | right slots << 2; 0: the node itself takes one slot);
// bits 16-17: left slots of the node's own wide node (the right child gets W minus them).
__device__ inline void sah_node(const Sah& sah, const float* box, uint32_t count, const float* cl, const float* cr,
                                float4& cost, uint32_t& dec)
{
    float dist[5];
    uint32_t split[5];
    for (uint32_t i = 2; i <= 4; i++) {
        float best = INFINITY;
        uint32_t bk = 1;
        for (uint32_t kl = 1; kl < i; kl++) {
            const float v = cl[kl - 1] + cr[i - kl - 1];
            if (v < best) {
                best = v;
                bk = kl;
            }
        }
        dist[i] = best;
        split[i] = bk;
    }
    const float a = half_area(box);
    const float c_int = a * sah.c_node + dist[sah.width];
    const float c_leaf = count <= sah.max_leaf ? a * sah.c_leaf * float(count) : INFINITY;
    float c[4];
    c[0] = fminf(c_leaf, c_int);
    uint32_t d = c_leaf <= c_int ? 1u : 0u;
    uint32_t code = 0;
    for (uint32_t i = 2; i <= 4; i++) {
        if (i <= sah.width && dist[i] < c[i - 2]) {
            c[i - 1] = dist[i];
            code = split[i] | ((i - split[i]) << 2);
        } else {
            c[i - 1] = c[i - 2];
        }
        d |= code << (4 * (i - 1));
    }
    dec = d | (split[sah.width] << 16);
    cost = make_float4(c[0], c[1], c[2], c[3]);
}

// Leaf boxes in sorted order, and the initial PLOC clusters (one per primitive).
struct Cluster {
    float4 lo;  // box min, segment (bits of w)
    float4 hi;  // box max, binary reference (bits of w)
};

__global__ void k_leaves(uint32_t n, const uint32_t* vals, const uint64_t* keys, const float4* lo, const float4* hi,
                         float* lbox, Cluster* cl)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t p = vals[i];
        const float4 a = lo[p], b = hi[p];
        float* o = &lbox[size_t(i) * 6];
        o[0] = a.x;
        o[1] = a.y;
        o[2] = a.z;
        o[3] = b.x;
        o[4] = b.y;
        o[5] = b.z;
        if (cl)
            cl[i] = Cluster{make_float4(a.x, a.y, a.z, u32_as_float(uint32_t(keys[i] >> 32))),
                            make_float4(b.x, b.y, b.z, u32_as_float(i | BREF_LEAF))};
    }
}

__global__ void k_empty_boxes(uint32_t n, float* box, uint32_t* root)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        for (int k = 0; k < 3; k++) {
            box[size_t(i) * 6 + k] = INFINITY;
            box[size_t(i) * 6 + k + 3] = -INFINITY;
        }
        root[i] = NONE;
    }
}

// ---- Karras (2012) binary radix tree

__device__ __forceinline__ int delta(const uint64_t* keys, uint32_t base, int n, int a, int b)
{
    if (b < 0 || b >= n)
        return -1;
    const uint32_t ka = uint32_t(keys[base + a]), kb = uint32_t(keys[base + b]);
    if (ka == kb)
        return 32 + __clz(uint32_t(a) ^ uint32_t(b));
    return __clz(ka ^ kb);
}

// Internal node g = base + j of a segment starting at base; its leaves are sorted positions.
__global__ void k_karras(uint32_t n, const uint64_t* keys, const uint32_t* seg_begin, uint32_t* left, uint32_t* right,
                         uint32_t* cnt, uint32_t* parent, uint32_t* leaf_parent)
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
        cnt[g] = uint32_t(hi_i - lo_i + 1);
    }
}

// Bottom-up bounds and SAH costs: the second child to finish completes its parent.
__global__ void k_bounds_up(uint32_t n, const uint32_t* left, const uint32_t* right, const uint32_t* cnt,
                            const uint32_t* parent, const uint32_t* leaf_parent, uint32_t* visit, float* ibox,
                            const float* lbox, float4* cost, uint32_t* dec, Sah sah)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
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
            box_union(bl, br, u);
            float cl[4], cr[4];
            child_cost(sah, l, bl, cost, cl);
            child_cost(sah, r, br, cost, cr);
            float4 cv;
            uint32_t dv;
            sah_node(sah, u, cnt[p], cl, cr, cv, dv);
            store_box(&ibox[size_t(p) * 6], u);
            float* q = reinterpret_cast<float*>(&cost[p]);
            const float cva[4] = {cv.x, cv.y, cv.z, cv.w};
            for (int k = 0; k < 4; k++)
                __hip_atomic_store(&q[k], cva[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            dec[p] = dv;
            __threadfence();
            p = parent[p];
        }
    }
}

__global__ void k_lbvh_roots(uint32_t segs, const uint32_t* seg_begin, const float* ibox, const float* lbox,
                             uint32_t* root, float* segbox)
{
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < segs; s += gridDim.x * blockDim.x) {
        const uint32_t b = seg_begin[s], e = seg_begin[s + 1];
        if (e == b)
            continue;
        const float* src = (e - b == 1) ? &lbox[size_t(b) * 6] : &ibox[size_t(b) * 6];
        root[s] = (e - b == 1) ? (b | BREF_LEAF) : b;
        for (int k = 0; k < 6; k++)
            segbox[size_t(s) * 6 + k] = src[k];
    }
}

// ---- PLOC (Meister and Bittner, IEEE TVCG 2018): each cluster finds its nearest neighbour
// among the PLOC_R clusters either side of it in Morton order (distance: surface area of
// the merged box), mutual nearest neighbours merge, the survivors are compacted in order.
// Pairs compare by (distance, index gap, parity of the lower index, lower index): a total
// order, so the closest pair of every segment is mutual and each pass merges at least once
// per unfinished segment; the parity term pairs (2k, 2k+1) among equal boxes.

constexpr int PLOC_R = 16;
constexpr int PLOC_B = 256;

__global__ __launch_bounds__(PLOC_B) void k_ploc_nn(uint32_t n, const Cluster* cl, uint32_t* nn)
{
    __shared__ float4 slo[PLOC_B + 2 * PLOC_R], shi[PLOC_B + 2 * PLOC_R];
    const int64_t b0 = int64_t(blockIdx.x) * PLOC_B;
    for (int t = threadIdx.x; t < PLOC_B + 2 * PLOC_R; t += PLOC_B) {
        const int64_t g = b0 - PLOC_R + t;
        if (g >= 0 && g < int64_t(n)) {
            slo[t] = cl[g].lo;
            shi[t] = cl[g].hi;
        } else {
            slo[t] = make_float4(0.0f, 0.0f, 0.0f, u32_as_float(NONE));  // matches no segment
            shi[t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    }
    __syncthreads();
    const int64_t i = b0 + threadIdx.x;
    if (i >= int64_t(n))
        return;
    const int li = int(threadIdx.x) + PLOC_R;
    const float4 a0 = slo[li], a1 = shi[li];
    const uint32_t sg = float_as_u32(a0.w);
    uint64_t best = ~0ull;
    uint32_t bj = NONE;
    for (int k = 1; k <= PLOC_R; k++) {
#pragma unroll
        for (int side = 0; side < 2; side++) {
            const int lj = side ? li + k : li - k;
            const float4 c0 = slo[lj];
            if (float_as_u32(c0.w) != sg)
                continue;
            const float4 c1 = shi[lj];
            const float dx = fmaxf(a1.x, c1.x) - fminf(a0.x, c0.x);
            const float dy = fmaxf(a1.y, c1.y) - fminf(a0.y, c0.y);
            const float dz = fmaxf(a1.z, c1.z) - fminf(a0.z, c0.z);
            const float d = fmaxf(dx * dy + dy * dz + dz * dx, 0.0f);
            const uint32_t lower = uint32_t(side ? i : i - k);
            const uint64_t key = (uint64_t(float_as_u32(d)) << 32) | (uint32_t(k) << 2) | ((lower & 1u) << 1) |
                                 uint32_t(side);
            if (key < best) {
                best = key;
                bj = uint32_t(side ? i + k : i - k);
            }
        }
    }
    nn[i] = bj;
}

// flags: bit 0 = the cluster survives this pass (unpaired, or the lower of a pair), bit 32 =
// it becomes a new internal node.
__global__ void k_ploc_flags(uint32_t n, const uint32_t* nn, uint64_t* flags)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t j = nn[i];
        const bool mutual = j != NONE && nn[j] == i;
        flags[i] = uint64_t(!mutual || i < j) | (uint64_t(mutual && i < j) << 32);
    }
}

__global__ void k_ploc_merge(uint32_t n, const Cluster* cin, const uint32_t* nn, const uint64_t* flags,
                             const uint64_t* scan, uint32_t node_base, Cluster* cout, uint32_t* left, uint32_t* right,
                             uint32_t* cnt, float* ibox, float4* cost, uint32_t* dec, uint32_t* totals, Sah sah)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint64_t f = flags[i], s = scan[i];
        if (i == n - 1) {
            totals[0] = uint32_t(s) + uint32_t(f);
            totals[1] = uint32_t(s >> 32) + uint32_t(f >> 32);
        }
        if (!(f & 1u))
            continue;
        const uint32_t p = uint32_t(s);
        if (!(f >> 32)) {
            cout[p] = cin[i];
            continue;
        }
        const Cluster a = cin[i], b = cin[nn[i]];
        const uint32_t id = node_base + uint32_t(s >> 32);
        const uint32_t ra = float_as_u32(a.hi.w), rb = float_as_u32(b.hi.w);
        const float ba[6] = {a.lo.x, a.lo.y, a.lo.z, a.hi.x, a.hi.y, a.hi.z};
        const float bb[6] = {b.lo.x, b.lo.y, b.lo.z, b.hi.x, b.hi.y, b.hi.z};
        float u[6];
        box_union(ba, bb, u);
        float cl[4], cr[4];
        child_cost(sah, ra, ba, cost, cl);
        child_cost(sah, rb, bb, cost, cr);
        const uint32_t count = ((ra & BREF_LEAF) ? 1u : cnt[ra]) + ((rb & BREF_LEAF) ? 1u : cnt[rb]);
        float4 cv;
        uint32_t dv;
        sah_node(sah, u, count, cl, cr, cv, dv);
        left[id] = ra;
        right[id] = rb;
        cnt[id] = count;
        for (int k = 0; k < 6; k++)
            ibox[size_t(id) * 6 + k] = u[k];
        cost[id] = cv;
        dec[id] = dv;
        cout[p] = Cluster{make_float4(u[0], u[1], u[2], a.lo.w), make_float4(u[3], u[4], u[5], u32_as_float(id))};
    }
}

// After the last pass: one cluster per non-empty segment.
__global__ void k_ploc_roots(uint32_t n, const Cluster* cl, uint32_t* root, float* segbox)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const Cluster c = cl[i];
        const uint32_t s = float_as_u32(c.lo.w);
        root[s] = float_as_u32(c.hi.w);
        float* o = &segbox[size_t(s) * 6];
        o[0] = c.lo.x;
        o[1] = c.lo.y;
        o[2] = c.lo.z;
        o[3] = c.hi.x;
        o[4] = c.hi.y;
        o[5] = c.hi.z;
    }
}

// ---- collapse into 4-wide nodes

struct CollapseItem {
    uint32_t bref;    // binary reference whose subtree becomes this node
    uint32_t slot;    // output node index
    uint32_t offset;  // first output position of the subtree's primitives
    uint32_t pad;
};

// One wide node per item: its W child slots are filled as the decisions of sah_node say;
// a child that is a leaf gets the next output positions of its primitives (depth first,
// left before right), a child that is an internal node becomes an item of the next level.
// tlas != 0: every leaf holds one instance (its original primitive index).
__global__ void k_collapse(const CollapseItem* in, const uint32_t* in_count, CollapseItem* out, uint32_t* out_count,
                           uint32_t* node_counter, uint32_t node_capacity, Node4* nodes, const uint32_t* left,
                           const uint32_t* right, const uint32_t* cnt, const uint32_t* dec, const float* ibox,
                           const float* lbox, const uint32_t* vals, uint32_t* perm, uint32_t n, uint32_t width,
                           int tlas)
{
    const uint32_t items = *in_count;
    for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < items; w += gridDim.x * blockDim.x) {
        const CollapseItem it = in[w];
        uint32_t c[4], off[4];
        int nc = 0;
        if (it.bref & BREF_LEAF) {  // a single-primitive set
            c[0] = it.bref;
            off[0] = it.offset;
            nc = 1;
        } else {
            struct Open {
                uint32_t ref, slots, offset;
            };
            Open st[4];
            int sp = 0;
            const uint32_t kl = (dec[it.bref] >> 16) & 3u;
            const uint32_t l = left[it.bref], r = right[it.bref];
            st[sp++] = Open{r, width - kl, it.offset + ((l & BREF_LEAF) ? 1u : cnt[l])};
            st[sp++] = Open{l, kl, it.offset};
            // slots sum to at most width <= 4, so at most 4 entries are ever pending
            while (sp > 0 && nc < 4) {
                const Open e = st[--sp];
                uint32_t code = 0;
                if (!(e.ref & BREF_LEAF) && e.slots >= 2)
                    code = (dec[e.ref] >> (4 * (e.slots - 1))) & 15u;
                if (code == 0 || sp > 2) {
                    c[nc] = e.ref;
                    off[nc] = e.offset;
                    nc++;
                    continue;
                }
                const uint32_t el = left[e.ref], er = right[e.ref];
                st[sp++] = Open{er, code >> 2, e.offset + ((el & BREF_LEAF) ? 1u : cnt[el])};
                st[sp++] = Open{el, code & 3u, e.offset};
            }
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
            if (r & BREF_LEAF) {
                const uint32_t pos = r & ~BREF_LEAF;
                if (tlas) {
                    node.child[k] = CHILD_LEAF | vals[pos];
                } else {
                    if (off[k] < n)
                        perm[off[k]] = pos;
                    node.child[k] = CHILD_LEAF | off[k];
                }
            } else if (!tlas && (dec[r] & 1u)) {
                // a leaf of cnt[r] <= 16 primitives: its subtree's leaves in order
                uint32_t st2[16];
                int sp2 = 0;
                uint32_t o = off[k];
                st2[sp2++] = r;
                while (sp2 > 0) {
                    const uint32_t x = st2[--sp2];
                    if (x & BREF_LEAF) {
                        if (o < n)
                            perm[o] = x & ~BREF_LEAF;
                        o++;
                    } else if (sp2 <= 14) {
                        st2[sp2++] = right[x];
                        st2[sp2++] = left[x];
                    }
                }
                node.child[k] = CHILD_LEAF | ((cnt[r] - 1u) << 27) | off[k];
            } else {
                const uint32_t slot = atomicAdd(node_counter, 1u);
                if (slot < node_capacity) {
                    const uint32_t q = atomicAdd(out_count, 1u);
                    out[q] = CollapseItem{r, slot, off[k], 0u};
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

// Record i holds the primitive at sorted position perm[i] (leaf order).
__global__ void k_tri_records(uint32_t n, const uint32_t* perm, const uint32_t* vals, const uint32_t* seg,
                              const float4* vertices, const uint32_t* indices, const pt_mesh_desc* meshes,
                              const uint32_t* tri_begin, TriRec* tris)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t p = vals[perm[i]];
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

constexpr uint32_t MAX_ENTRIES = 16;  // TLAS entries of one instance (two opened levels of 4)

// World box of an object-space box under an instance transform: the 8 corners transformed,
// padded for the rounding of the transform.
__device__ void world_box(const InstanceData& d, float x0, float y0, float z0, float x1, float y1, float z1,
                          float4& lo, float4& hi)
{
    float3 mn = f3(INFINITY, INFINITY, INFINITY), mx = f3(-INFINITY, -INFINITY, -INFINITY);
    for (int c = 0; c < 8; c++) {
        const float px = (c & 1) ? x1 : x0, py = (c & 2) ? y1 : y0, pz = (c & 4) ? z1 : z0;
        const float wx = d.m[0][0] * px + d.m[0][1] * py + d.m[0][2] * pz + d.m[0][3];
        const float wy = d.m[1][0] * px + d.m[1][1] * py + d.m[1][2] * pz + d.m[1][3];
        const float wz = d.m[2][0] * px + d.m[2][1] * py + d.m[2][2] * pz + d.m[2][3];
        mn = f3(fminf(mn.x, wx), fminf(mn.y, wy), fminf(mn.z, wz));
        mx = f3(fmaxf(mx.x, wx), fmaxf(mx.y, wy), fmaxf(mx.z, wz));
    }
    const float3 ext = f3(fabsf(mn.x) + fabsf(mx.x), fabsf(mn.y) + fabsf(mx.y), fabsf(mn.z) + fabsf(mx.z));
    const float eps = 1e-6f;
    lo = make_float4(mn.x - eps * ext.x, mn.y - eps * ext.y, mn.z - eps * ext.z, 0.0f);
    hi = make_float4(mx.x + eps * ext.x, mx.y + eps * ext.y, mx.z + eps * ext.z, 0.0f);
}

// World box of BLAS node `node` under an instance transform, tightened: the union of the world
// boxes of its descendants `levels` (1 or 2) levels down, a triangle leaf ending its branch with
// its own box. Each descendant box lies inside the node's box and together they hold all of its
// triangles, so the union contains the node's geometry and is no larger than the node box's own.
__device__ void tight_world_box(const InstanceData& d, const Node4* nodes, uint32_t node, uint32_t levels, float4& lo,
                                float4& hi)
{
    lo = make_float4(INFINITY, INFINITY, INFINITY, 0.0f);
    hi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
    auto add = [&](const Node4& nd, int k) {
        float4 bl, bh;
        world_box(d, nd.lo_x[k], nd.lo_y[k], nd.lo_z[k], nd.hi_x[k], nd.hi_y[k], nd.hi_z[k], bl, bh);
        lo = make_float4(fminf(lo.x, bl.x), fminf(lo.y, bl.y), fminf(lo.z, bl.z), 0.0f);
        hi = make_float4(fmaxf(hi.x, bh.x), fmaxf(hi.y, bh.y), fmaxf(hi.z, bh.z), 0.0f);
    };
    const Node4& nd = nodes[node];
    for (int k = 0; k < 4; k++) {
        const uint32_t ch = nd.child[k];
        if (ch == CHILD_EMPTY)
            continue;
        if (levels < 2 || (ch & CHILD_LEAF)) {
            add(nd, k);
            continue;
        }
        const Node4& g = nodes[ch];
        for (int j = 0; j < 4; j++)
            if (g.child[j] != CHILD_EMPTY)
                add(g, j);
    }
}

// World box of every instance's transformed vertices, padded for the rounding of the transform
// as world_box pads: it holds all of the instance's triangles and, for a rotated instance, is
// far tighter than the transformed corners of the mesh box. Each instance's vertex range is
// split into VBOX_PARTS slices; k_vertex_box_parts reduces each slice in one block, and
// k_vertex_boxes combines the slices of each instance.
constexpr uint32_t VBOX_PARTS = 16;

__global__ __launch_bounds__(256) void k_vertex_box_parts(uint32_t n, const InstanceData* inst,
                                                          const uint32_t* inst_mesh, const pt_mesh_desc* meshes,
                                                          const float4* vertices, float4* part_lo, float4* part_hi)
{
    __shared__ float red[6][256];
    const uint32_t t = threadIdx.x;
    for (uint32_t b = blockIdx.x; b < n * VBOX_PARTS; b += gridDim.x) {
        const uint32_t i = b / VBOX_PARTS, part = b % VBOX_PARTS;
        const InstanceData& d = inst[i];
        const pt_mesh_desc md = meshes[inst_mesh[i]];
        const uint32_t per = (md.vertex_count + VBOX_PARTS - 1) / VBOX_PARTS;
        const uint32_t v0 = min(md.vertex_count, part * per), v1 = min(md.vertex_count, v0 + per);
        float3 mn = f3(INFINITY, INFINITY, INFINITY), mx = f3(-INFINITY, -INFINITY, -INFINITY);
        for (uint32_t v = v0 + t; v < v1; v += 256) {
            const float4 p = vertices[size_t(md.vertex_start + v) * 2];
            const float wx = d.m[0][0] * p.x + d.m[0][1] * p.y + d.m[0][2] * p.z + d.m[0][3];
            const float wy = d.m[1][0] * p.x + d.m[1][1] * p.y + d.m[1][2] * p.z + d.m[1][3];
            const float wz = d.m[2][0] * p.x + d.m[2][1] * p.y + d.m[2][2] * p.z + d.m[2][3];
            mn = f3(fminf(mn.x, wx), fminf(mn.y, wy), fminf(mn.z, wz));
            mx = f3(fmaxf(mx.x, wx), fmaxf(mx.y, wy), fmaxf(mx.z, wz));
        }
        red[0][t] = mn.x, red[1][t] = mn.y, red[2][t] = mn.z;
        red[3][t] = mx.x, red[4][t] = mx.y, red[5][t] = mx.z;
        __syncthreads();
        for (uint32_t h = 128; h > 0; h >>= 1) {
            if (t < h) {
                for (int k = 0; k < 3; k++) {
                    red[k][t] = fminf(red[k][t], red[k][t + h]);
                    red[3 + k][t] = fmaxf(red[3 + k][t], red[3 + k][t + h]);
                }
            }
            __syncthreads();
        }
        if (t == 0) {
            part_lo[b] = make_float4(red[0][0], red[1][0], red[2][0], 0.0f);
            part_hi[b] = make_float4(red[3][0], red[4][0], red[5][0], 0.0f);
        }
        __syncthreads();
    }
}

__global__ void k_vertex_boxes(uint32_t n, const float4* part_lo, const float4* part_hi, float4* lo, float4* hi)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float3 a = f3(INFINITY, INFINITY, INFINITY), b = f3(-INFINITY, -INFINITY, -INFINITY);
        for (uint32_t k = 0; k < VBOX_PARTS; k++) {
            const float4 l = part_lo[i * VBOX_PARTS + k], h = part_hi[i * VBOX_PARTS + k];
            a = f3(fminf(a.x, l.x), fminf(a.y, l.y), fminf(a.z, l.z));
            b = f3(fmaxf(b.x, h.x), fmaxf(b.y, h.y), fmaxf(b.z, h.z));
        }
        const float3 ext = f3(fabsf(a.x) + fabsf(b.x), fabsf(a.y) + fabsf(b.y), fabsf(a.z) + fabsf(b.z));
        const float eps = 1e-6f;
        lo[i] = make_float4(a.x - eps * ext.x, a.y - eps * ext.y, a.z - eps * ext.z, 0.0f);
        hi[i] = make_float4(b.x + eps * ext.x, b.y + eps * ext.y, b.z + eps * ext.z, 0.0f);
    }
}

__device__ __forceinline__ float half_area(float4 lo, float4 hi)
{
    const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
    return dx * dy + dy * dz + dz * dx;
}

// The TLAS entries of each instance (at most MAX_ENTRIES, from slot i * MAX_ENTRIES): its BLAS
// root, or - for up to `depth` levels - the children of an entry that opens: every child an
// internal node, and the surface areas of their world boxes summing to less than ratio x that
// of the entry's own world box. An instance of an empty mesh gets one entry that enters nothing.
// With box_levels > 0 each entry's box is then tightened (tight_world_box, at most two levels);
// with vertex boxes (vlo, vhi: k_vertex_boxes) an entry at the BLAS root takes its instance's.
__global__ void k_open_instances(uint32_t n, const InstanceData* inst, const uint32_t* inst_mesh,
                                 const uint32_t* mesh_tris, const float* mesh_box, const Node4* nodes, uint32_t depth,
                                 float ratio, uint32_t box_levels, const float4* vlo, const float4* vhi,
                                 uint32_t* count, uint2* ent, float4* lo, float4* hi)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t m = inst_mesh[i], base = i * MAX_ENTRIES;
        if (mesh_tris[m] == 0) {
            ent[base] = make_uint2(i, CHILD_EMPTY);
            lo[base] = make_float4(INFINITY, INFINITY, INFINITY, 0.0f);
            hi[base] = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
            count[i] = 1;
            continue;
        }
        const InstanceData& d = inst[i];
        uint32_t nodes_of[MAX_ENTRIES];
        float4 blo[MAX_ENTRIES], bhi[MAX_ENTRIES];
        uint32_t cnt = 1;
        nodes_of[0] = m;  // the BLAS root of mesh m is node m
        const float* b = &mesh_box[size_t(m) * 6];
        world_box(d, b[0], b[1], b[2], b[3], b[4], b[5], blo[0], bhi[0]);
        for (uint32_t level = 0; level < depth; level++) {
            uint32_t next_nodes[MAX_ENTRIES];
            float4 nlo[MAX_ENTRIES], nhi[MAX_ENTRIES];
            uint32_t nc = 0;
            for (uint32_t e = 0; e < cnt; e++) {
                const Node4& nd = nodes[nodes_of[e]];
                float4 clo[4], chi[4];
                uint32_t cn[4], k_used = 0;
                bool open = true;
                float sum = 0.0f;
                for (int k = 0; k < 4; k++) {
                    const uint32_t ch = nd.child[k];
                    if (ch == CHILD_EMPTY)
                        continue;
                    if (ch & CHILD_LEAF) {
                        open = false;  // a triangle leaf stays reachable only through this node
                        break;
                    }
                    world_box(d, nd.lo_x[k], nd.lo_y[k], nd.lo_z[k], nd.hi_x[k], nd.hi_y[k], nd.hi_z[k], clo[k_used],
                              chi[k_used]);
                    sum += half_area(clo[k_used], chi[k_used]);
                    cn[k_used++] = ch;
                }
                open = open && k_used > 0 && sum < ratio * half_area(blo[e], bhi[e]) &&
                       nc + k_used + (cnt - e - 1) <= MAX_ENTRIES;
                if (open) {
                    for (uint32_t k = 0; k < k_used; k++) {
                        next_nodes[nc] = cn[k];
                        nlo[nc] = clo[k];
                        nhi[nc] = chi[k];
                        nc++;
                    }
                } else {
                    next_nodes[nc] = nodes_of[e];
                    nlo[nc] = blo[e];
                    nhi[nc] = bhi[e];
                    nc++;
                }
            }
            for (uint32_t e = 0; e < nc; e++) {
                nodes_of[e] = next_nodes[e];
                blo[e] = nlo[e];
                bhi[e] = nhi[e];
            }
            cnt = nc;
        }
        for (uint32_t e = 0; e < cnt; e++) {
            if (vlo && nodes_of[e] == m) {
                blo[e] = vlo[i];
                bhi[e] = vhi[i];
            } else if (box_levels > 0) {
                tight_world_box(d, nodes, nodes_of[e], std::min(box_levels, 2u), blo[e], bhi[e]);
            }
            ent[base + e] = make_uint2(i, nodes_of[e]);
            lo[base + e] = blo[e];
            hi[base + e] = bhi[e];
        }
        count[i] = cnt;
    }
}

// Packs the entries of every instance at its offset (exclusive prefix sum of the counts).
__global__ void k_pack_entries(uint32_t n, const uint32_t* count, const uint32_t* offset, const uint2* ent,
                               const float4* lo, const float4* hi, uint2* out_ent, float4* out_lo, float4* out_hi,
                               uint32_t* seg)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        for (uint32_t e = 0; e < count[i]; e++) {
            const uint32_t src = i * MAX_ENTRIES + e, dst = offset[i] + e;
            out_ent[dst] = ent[src];
            out_lo[dst] = lo[src];
            out_hi[dst] = hi[src];
            seg[dst] = 0;
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

// Builds the binary tree of every segment (PLOC or Karras), its SAH collapse decisions, and
// collapses it into nodes[slot0 + s] (root of segment s) and further slots from
// *node_counter (device). Prim bounds / seg must be filled. BLAS sets also get bs.perm.
hipError_t build_set(hipStream_t st, Scratch& sc, BuildSet& bs, const std::vector<uint32_t>& seg_begin_host,
                     const Sah& sah, uint32_t builder, int tlas, Node4* nodes, uint32_t slot0,
                     uint32_t* d_node_counter, uint32_t node_capacity)
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
    bs.vals = vals;
    bs.keys = keys;

    CHECK(sc.alloc(bs.left, n));
    CHECK(sc.alloc(bs.right, n));
    CHECK(sc.alloc(bs.cnt, n));
    CHECK(sc.alloc(bs.dec, n));
    CHECK(sc.alloc(bs.cost, n));
    CHECK(sc.alloc(bs.ibox, size_t(n) * 6));
    CHECK(sc.alloc(bs.lbox, size_t(n) * 6));
    CHECK(sc.alloc(bs.root, bs.segs));
    CHECK(sc.alloc(bs.segbox, size_t(bs.segs) * 6));
    hipLaunchKernelGGL(k_empty_boxes, dim3(grid_for(bs.segs)), dim3(B), 0, st, bs.segs, bs.segbox, bs.root);

    if (builder == BVH_BUILDER_LBVH) {
        uint32_t *parent, *leaf_parent, *visit;
        CHECK(sc.alloc(parent, n));
        CHECK(sc.alloc(leaf_parent, n));
        CHECK(sc.alloc(visit, n));
        CHECK(hipMemsetAsync(visit, 0, size_t(n) * 4, st));
        hipLaunchKernelGGL(k_leaves, dim3(grid_for(n)), dim3(B), 0, st, n, vals, keys, bs.lo, bs.hi, bs.lbox,
                           (Cluster*)nullptr);
        hipLaunchKernelGGL(k_karras, dim3(grid_for(n)), dim3(B), 0, st, n, keys, bs.seg_begin, bs.left, bs.right,
                           bs.cnt, parent, leaf_parent);
        hipLaunchKernelGGL(k_bounds_up, dim3(grid_for(n)), dim3(B), 0, st, n, bs.left, bs.right, bs.cnt, parent,
                           leaf_parent, visit, bs.ibox, bs.lbox, bs.cost, bs.dec, sah);
        hipLaunchKernelGGL(k_lbvh_roots, dim3(grid_for(bs.segs)), dim3(B), 0, st, bs.segs, bs.seg_begin, bs.ibox,
                           bs.lbox, bs.root, bs.segbox);
        CHECK(hipGetLastError());
    } else {
        Cluster *ca, *cb;
        uint32_t *nn, *totals;
        uint64_t *flags, *scan;
        CHECK(sc.alloc(ca, n));
        CHECK(sc.alloc(cb, n));
        CHECK(sc.alloc(nn, n));
        CHECK(sc.alloc(flags, n));
        CHECK(sc.alloc(scan, n));
        CHECK(sc.alloc(totals, 2));
        size_t scan_bytes = 0;
        CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, flags, scan, int(n), st));
        uint8_t* scan_temp;
        CHECK(sc.alloc(scan_temp, scan_bytes));
        hipLaunchKernelGGL(k_leaves, dim3(grid_for(n)), dim3(B), 0, st, n, vals, keys, bs.lo, bs.hi, bs.lbox, ca);
        uint32_t m = n, node_base = 0;
        // every pass merges at least one pair of each segment that still has two clusters
        for (uint32_t pass = 0; pass < n; pass++) {
            hipLaunchKernelGGL(k_ploc_nn, dim3((m + PLOC_B - 1) / PLOC_B), dim3(PLOC_B), 0, st, m, ca, nn);
            hipLaunchKernelGGL(k_ploc_flags, dim3(grid_for(m)), dim3(B), 0, st, m, nn, flags);
            CHECK(hipcub::DeviceScan::ExclusiveSum(scan_temp, scan_bytes, flags, scan, int(m), st));
            hipLaunchKernelGGL(k_ploc_merge, dim3(grid_for(m)), dim3(B), 0, st, m, ca, nn, flags, scan, node_base, cb,
                               bs.left, bs.right, bs.cnt, bs.ibox, bs.cost, bs.dec, totals, sah);
            CHECK(hipGetLastError());
            uint32_t t[2];
            CHECK(hipMemcpyAsync(t, totals, 8, hipMemcpyDeviceToHost, st));
            CHECK(hipStreamSynchronize(st));
            std::swap(ca, cb);
            m = t[0];
            node_base += t[1];
            if (t[1] == 0)
                break;
        }
        hipLaunchKernelGGL(k_ploc_roots, dim3(grid_for(m)), dim3(B), 0, st, m, ca, bs.root, bs.segbox);
        CHECK(hipGetLastError());
    }

    // Top-down collapse, one launch per level.
    std::vector<uint32_t> root(bs.segs);
    CHECK(hipMemcpyAsync(root.data(), bs.root, size_t(bs.segs) * 4, hipMemcpyDeviceToHost, st));
    CHECK(hipStreamSynchronize(st));
    std::vector<CollapseItem> roots;
    for (uint32_t s = 0; s < bs.segs; s++)
        if (root[s] != NONE)
            roots.push_back(CollapseItem{root[s], slot0 + s, seg_begin_host[s], 0u});
    if (!tlas) {
        CHECK(sc.alloc(bs.perm, n));
        CHECK(hipMemsetAsync(bs.perm, 0, size_t(n) * 4, st));  // a hole would repeat primitive 0 (validation)
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
    for (int level = 0; level < 4096 && host_count > 0; level++) {
        CollapseItem* qin = cur ? qb : qa;
        CollapseItem* qout = cur ? qa : qb;
        CHECK(hipMemsetAsync(&counts[1 - cur], 0, 4, st));
        hipLaunchKernelGGL(k_collapse, dim3(grid_for(host_count)), dim3(B), 0, st, qin, &counts[cur], qout,
                           &counts[1 - cur], d_node_counter, node_capacity, nodes, bs.left, bs.right, bs.cnt, bs.dec,
                           bs.ibox, bs.lbox, vals, bs.perm, n, sah.width, tlas);
        CHECK(hipGetLastError());
        CHECK(hipMemcpyAsync(&host_count, &counts[1 - cur], 4, hipMemcpyDeviceToHost, st));
        CHECK(hipStreamSynchronize(st));
        cur = 1 - cur;
    }
    if (host_count > 0)
        return hipErrorUnknown;  // deeper than the level cap
    return hipSuccess;
}

}  // namespace

void free_bvh(BvhBuffers& b)
{
    (void)hipFree(b.nodes);
    (void)hipFree(b.tris);
    (void)hipFree(b.inst_inv);
    (void)hipFree(b.tlas_entries);
    b = BvhBuffers{};
}

// The TLAS built on the host, for a modest number of entries: a binary tree split by the full
// SAH sweep (every position of the centroid order along every axis; a split costs
// area(left) x count(left) + area(right) x count(right)), then collapsed SAH-optimally into nodes
// of at most sah.width children. The collapse is a dynamic program over the binary tree: f(b, k)
// is the cheapest cover of b's subtree by k child roots, where a root costs c_leaf x area for an
// entry and c_node x area plus its own best cover of 2..width roots for an internal node. Nodes
// are numbered from slot0 (the root) in the order they are made.
static void host_tlas(const std::vector<float4>& lo, const std::vector<float4>& hi, const Sah& sah, uint32_t slot0,
                      std::vector<Node4>& out)
{
    const uint32_t E = uint32_t(lo.size());
    const uint32_t W = std::max(2u, std::min(sah.width, 4u));
    struct B {
        float lo[3], hi[3];
        int32_t l, r;  // children (binary node ids), or -1 with entry id in l for a leaf
    };
    std::vector<B> bn;
    bn.reserve(2 * E);
    auto area = [](const B& b) {
        const double dx = std::max(0.0, double(b.hi[0] - b.lo[0])), dy = std::max(0.0, double(b.hi[1] - b.lo[1])),
                     dz = std::max(0.0, double(b.hi[2] - b.lo[2]));
        return dx * dy + dy * dz + dz * dx;
    };
    auto box_of = [&](const uint32_t* ids, size_t n, B& b) {
        for (int k = 0; k < 3; k++) {
            b.lo[k] = INFINITY;
            b.hi[k] = -INFINITY;
        }
        for (size_t i = 0; i < n; i++) {
            const float4 a = lo[ids[i]], c = hi[ids[i]];
            b.lo[0] = std::min(b.lo[0], a.x), b.lo[1] = std::min(b.lo[1], a.y), b.lo[2] = std::min(b.lo[2], a.z);
            b.hi[0] = std::max(b.hi[0], c.x), b.hi[1] = std::max(b.hi[1], c.y), b.hi[2] = std::max(b.hi[2], c.z);
        }
    };
    std::vector<uint32_t> ids(E), tmp(E), best_order;
    for (uint32_t i = 0; i < E; i++)
        ids[i] = i;
    std::vector<double> suffix(E + 1);
    // iterative top-down build over [begin, end) ranges of ids
    struct Task {
        uint32_t begin, end;
        int32_t node;
    };
    bn.push_back(B{});
    std::vector<Task> work{{0, E, 0}};
    while (!work.empty()) {
        const Task tk = work.back();
        work.pop_back();
        B& nb = bn[tk.node];
        box_of(&ids[tk.begin], tk.end - tk.begin, nb);
        const uint32_t n = tk.end - tk.begin;
        if (n == 1) {
            nb.l = int32_t(ids[tk.begin]);
            nb.r = -1;
            continue;
        }
        double best = INFINITY;
        uint32_t best_split = n / 2;
        int best_axis = -1;
        for (int ax = 0; ax < 3; ax++) {
            std::copy(ids.begin() + tk.begin, ids.begin() + tk.end, tmp.begin());
            auto cen = [&](uint32_t e) {
                const float c = ax == 0 ? lo[e].x + hi[e].x : (ax == 1 ? lo[e].y + hi[e].y : lo[e].z + hi[e].z);
                return c == c ? c : 0.0f;  // an empty box (an instance of an empty mesh) sorts at 0
            };
            std::stable_sort(tmp.begin(), tmp.begin() + n, [&](uint32_t a, uint32_t b) { return cen(a) < cen(b); });
            B acc;
            box_of(&tmp[n - 1], 1, acc);
            suffix[n - 1] = area(acc);
            for (uint32_t i = n - 1; i-- > 0;) {
                B one;
                box_of(&tmp[i], 1, one);
                for (int k = 0; k < 3; k++)
                    acc.lo[k] = std::min(acc.lo[k], one.lo[k]), acc.hi[k] = std::max(acc.hi[k], one.hi[k]);
                suffix[i] = area(acc);
            }
            box_of(&tmp[0], 1, acc);
            for (uint32_t i = 1; i < n; i++) {  // left = [0, i), right = [i, n)
                const double c = area(acc) * double(i) + suffix[i] * double(n - i);
                if (c < best) {
                    best = c;
                    best_split = i;
                    best_axis = ax;
                }
                B one;
                box_of(&tmp[i], 1, one);
                for (int k = 0; k < 3; k++)
                    acc.lo[k] = std::min(acc.lo[k], one.lo[k]), acc.hi[k] = std::max(acc.hi[k], one.hi[k]);
            }
            if (best_axis == ax)
                best_order.assign(tmp.begin(), tmp.begin() + n);
        }
        if (best_axis >= 0)
            std::copy(best_order.begin(), best_order.end(), ids.begin() + tk.begin);
        const int32_t l = int32_t(bn.size()), r = l + 1;
        bn.push_back(B{});
        bn.push_back(B{});
        bn[tk.node].l = l;
        bn[tk.node].r = r;
        work.push_back({tk.begin, tk.begin + best_split, l});
        work.push_back({tk.begin + best_split, tk.end, r});
    }
    // collapse: children are made after parents, so a reverse sweep sees children first
    const size_t NB = bn.size();
    std::vector<double> f(NB * 5, INFINITY);
    std::vector<uint8_t> split(NB * 5, 0), own(NB, 0);  // k1 of f(b, k); best root count of b as a node
    for (size_t b = NB; b-- > 0;) {
        const B& x = bn[b];
        if (x.r < 0) {
            f[b * 5 + 1] = double(sah.c_leaf) * area(x);
            continue;
        }
        for (uint32_t k = 2; k <= W; k++)
            for (uint32_t k1 = 1; k1 < k; k1++) {
                const double c = f[size_t(x.l) * 5 + k1] + f[size_t(x.r) * 5 + (k - k1)];
                if (c < f[b * 5 + k]) {
                    f[b * 5 + k] = c;
                    split[b * 5 + k] = uint8_t(k1);
                }
            }
        double g = INFINITY;
        for (uint32_t k = 2; k <= W; k++)
            if (f[b * 5 + k] < g) {
                g = f[b * 5 + k];
                own[b] = uint8_t(k);
            }
        f[b * 5 + 1] = double(sah.c_node) * area(x) + g;
    }
    // emit
    out.clear();
    auto empty_node = []() {
        Node4 nd;
        for (int k = 0; k < 4; k++) {
            nd.lo_x[k] = nd.lo_y[k] = nd.lo_z[k] = INFINITY;
            nd.hi_x[k] = nd.hi_y[k] = nd.hi_z[k] = -INFINITY;
            nd.child[k] = CHILD_EMPTY;
            nd.pad[k] = 0;
        }
        return nd;
    };
    if (bn[0].r < 0) {  // one entry: a root with one leaf child
        Node4 nd = empty_node();
        const B& x = bn[0];
        nd.lo_x[0] = x.lo[0], nd.lo_y[0] = x.lo[1], nd.lo_z[0] = x.lo[2];
        nd.hi_x[0] = x.hi[0], nd.hi_y[0] = x.hi[1], nd.hi_z[0] = x.hi[2];
        nd.child[0] = CHILD_LEAF | uint32_t(x.l);
        out.push_back(nd);
        return;
    }
    std::vector<std::pair<int32_t, uint32_t>> q{{0, 0}};  // (binary node, output node)
    out.push_back(empty_node());
    for (size_t qi = 0; qi < q.size(); qi++) {
        const int32_t b = q[qi].first;
        const uint32_t o = q[qi].second;
        std::vector<int32_t> roots;
        std::vector<std::pair<int32_t, uint32_t>> cover{{b, own[b]}};
        while (!cover.empty()) {
            const auto c = cover.back();
            cover.pop_back();
            if (c.second == 1) {
                roots.push_back(c.first);
                continue;
            }
            const uint32_t k1 = split[size_t(c.first) * 5 + c.second];
            cover.push_back({bn[c.first].r, c.second - k1});
            cover.push_back({bn[c.first].l, k1});
        }
        for (size_t k = 0; k < roots.size(); k++) {
            const B& x = bn[roots[k]];
            Node4& nd = out[o];
            nd.lo_x[k] = x.lo[0], nd.lo_y[k] = x.lo[1], nd.lo_z[k] = x.lo[2];
            nd.hi_x[k] = x.hi[0], nd.hi_y[k] = x.hi[1], nd.hi_z[k] = x.hi[2];
            if (x.r < 0) {
                nd.child[k] = CHILD_LEAF | uint32_t(x.l);
            } else {
                const uint32_t id = uint32_t(out.size());
                out[o].child[k] = slot0 + id;
                out.push_back(empty_node());
                q.push_back({roots[k], id});
            }
        }
    }
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
    const uint32_t width = std::min<uint32_t>(std::max<uint32_t>(params.node_width, 2), 4);
    const Sah blas_sah{width, leaf, params.sah_node_cost, params.sah_tri_cost};
    const Sah tlas_sah{width, 1, params.sah_node_cost, params.sah_instance_cost};

    // Node budget: BLAS roots + at most one node per binary internal node; TLAS likewise, over
    // at most MAX_ENTRIES entries per instance when instances may open.
    const uint32_t blas_cap = M + T + 1;
    const uint32_t open_depth = std::min<uint32_t>(params.tlas_open, 2);
    const uint32_t per_inst = open_depth == 0 ? 1u : (open_depth == 1 ? 4u : MAX_ENTRIES);
    const uint32_t tlas_cap = instance_count * per_inst + 2;
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
    float* mesh_box = nullptr;  // root box of each mesh's BLAS
    if (T > 0) {
        blas.n = T;
        blas.segs = M;
        blas.seg_begin = d_tri_begin;
        CHECK(sc.alloc(blas.lo, T));
        CHECK(sc.alloc(blas.hi, T));
        CHECK(sc.alloc(blas.seg, T));
        hipLaunchKernelGGL(k_tri_bounds, dim3(grid_for(T)), dim3(256), 0, st, T, d_vertices, d_indices, d_meshes,
                           d_tri_begin, M, blas.lo, blas.hi, blas.seg);
        CHECK(build_set(st, sc, blas, tri_begin, blas_sah, params.builder, 0, out.nodes, 0, d_counter, blas_cap));
        hipLaunchKernelGGL(k_tri_records, dim3(grid_for(T)), dim3(256), 0, st, T, blas.perm, blas.vals, blas.seg,
                           d_vertices, d_indices, d_meshes, d_tri_begin, out.tris);
        mesh_box = blas.segbox;
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
        if (inst_mesh[i] >= M)
            return hipErrorInvalidValue;
    }
    CHECK(hipMalloc(&out.inst_inv, xf.size() * sizeof(InstXform)));
    CHECK(hipMemcpyAsync(out.inst_inv, xf.data(), xf.size() * sizeof(InstXform), hipMemcpyHostToDevice, st));

    // TLAS over the instances' entries (world bounds of their BLAS roots or opened subtrees).
    out.tlas_root = CHILD_EMPTY;
    out.entry_count = 0;
    if (instance_count > 0) {
        float* dummy_box = nullptr;
        CHECK(sc.alloc(dummy_box, 6));
        uint32_t *d_cnt, *d_off;
        uint2* d_ent;
        float4 *d_lo, *d_hi;
        const size_t slots = size_t(instance_count) * MAX_ENTRIES;
        CHECK(sc.alloc(d_cnt, instance_count));
        CHECK(sc.alloc(d_off, instance_count));
        CHECK(sc.alloc(d_ent, slots));
        CHECK(sc.alloc(d_lo, slots));
        CHECK(sc.alloc(d_hi, slots));
        float4 *d_vlo = nullptr, *d_vhi = nullptr;
        if (T > 0 && params.tlas_box_levels >= 3) {
            float4 *d_plo, *d_phi;
            CHECK(sc.alloc(d_vlo, instance_count));
            CHECK(sc.alloc(d_vhi, instance_count));
            CHECK(sc.alloc(d_plo, size_t(instance_count) * VBOX_PARTS));
            CHECK(sc.alloc(d_phi, size_t(instance_count) * VBOX_PARTS));
            hipLaunchKernelGGL(k_vertex_box_parts, dim3(std::min<uint32_t>(instance_count * VBOX_PARTS, 65536)),
                               dim3(256), 0, st, instance_count, d_instances, d_instance_mesh, d_meshes, d_vertices,
                               d_plo, d_phi);
            hipLaunchKernelGGL(k_vertex_boxes, dim3(grid_for(instance_count)), dim3(256), 0, st, instance_count,
                               d_plo, d_phi, d_vlo, d_vhi);
        }
        hipLaunchKernelGGL(k_open_instances, dim3(grid_for(instance_count, 64)), dim3(64), 0, st, instance_count,
                           d_instances, d_instance_mesh, d_mesh_tris, mesh_box ? mesh_box : dummy_box, out.nodes,
                           T > 0 ? open_depth : 0u, params.tlas_open_ratio, T > 0 ? params.tlas_box_levels : 0u,
                           d_vlo, d_vhi, d_cnt, d_ent, d_lo, d_hi);
        CHECK(hipGetLastError());
        std::vector<uint32_t> cnt(instance_count), off(instance_count);
        CHECK(hipMemcpyAsync(cnt.data(), d_cnt, size_t(instance_count) * 4, hipMemcpyDeviceToHost, st));
        CHECK(hipStreamSynchronize(st));
        uint32_t E = 0;
        for (uint32_t i = 0; i < instance_count; i++) {
            if (cnt[i] == 0 || cnt[i] > MAX_ENTRIES)
                return hipErrorUnknown;
            off[i] = E;
            E += cnt[i];
        }
        CHECK(hipMemcpyAsync(d_off, off.data(), size_t(instance_count) * 4, hipMemcpyHostToDevice, st));
        CHECK(hipMalloc(&out.tlas_entries, size_t(E) * sizeof(uint2)));
        out.entry_count = E;
        BuildSet tl;
        tl.n = E;
        tl.segs = 1;
        std::vector<uint32_t> sb = {0, E};
        uint32_t* d_sb;
        CHECK(sc.alloc(d_sb, 2));
        CHECK(hipMemcpyAsync(d_sb, sb.data(), 8, hipMemcpyHostToDevice, st));
        tl.seg_begin = d_sb;
        CHECK(sc.alloc(tl.lo, E));
        CHECK(sc.alloc(tl.hi, E));
        CHECK(sc.alloc(tl.seg, E));
        hipLaunchKernelGGL(k_pack_entries, dim3(grid_for(instance_count)), dim3(256), 0, st, instance_count, d_cnt,
                           d_off, d_ent, d_lo, d_hi, out.tlas_entries, tl.lo, tl.hi, tl.seg);
        const uint32_t tlas_slot = blas_nodes;
        if (params.tlas_builder == BVH_TLAS_HOST && E <= HOST_TLAS_MAX) {
            std::vector<float4> hlo(E), hhi(E);
            CHECK(hipMemcpyAsync(hlo.data(), tl.lo, size_t(E) * sizeof(float4), hipMemcpyDeviceToHost, st));
            CHECK(hipMemcpyAsync(hhi.data(), tl.hi, size_t(E) * sizeof(float4), hipMemcpyDeviceToHost, st));
            CHECK(hipStreamSynchronize(st));
            std::vector<Node4> tn;
            host_tlas(hlo, hhi, tlas_sah, tlas_slot, tn);
            counter = tlas_slot + uint32_t(tn.size());
            if (counter > out.node_capacity)
                return hipErrorOutOfMemory;
            CHECK(hipMemcpyAsync(out.nodes + tlas_slot, tn.data(), tn.size() * sizeof(Node4), hipMemcpyHostToDevice,
                                 st));
        } else {
            counter = blas_nodes + 1;
            CHECK(hipMemcpyAsync(d_counter, &counter, 4, hipMemcpyHostToDevice, st));
            CHECK(build_set(st, sc, tl, sb, tlas_sah, params.builder, 1, out.nodes, tlas_slot, d_counter,
                            out.node_capacity));
            CHECK(hipMemcpyAsync(&counter, d_counter, 4, hipMemcpyDeviceToHost, st));
            CHECK(hipStreamSynchronize(st));
            if (counter > out.node_capacity)
                return hipErrorOutOfMemory;
        }
        out.tlas_root = tlas_slot;
    }
    out.node_count = counter;
    CHECK(hipStreamSynchronize(st));
    return hipSuccess;
}

}  // namespace ptk
