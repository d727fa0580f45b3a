// Traversal and the wavefront stages of the path tracer.
//
// One batch of paths runs as: generate (camera rays, pathtrace.rg.hlsl:16-35) -> per
// segment: extend (closest hit of every queued ray) -> shade (miss / closest-hit logic of
// pathtrace.ms.hlsl and pathtrace.ch.hlsl, Russian roulette of pathtrace.rg.hlsl:53, and
// compaction of continuing paths and shadow rays into queues) -> shadow (any hit, adds
// the unoccluded next-event contribution) -> accumulate (running mean per pixel in sample
// order, pathtrace.rg.hlsl:61-67).
//
// Traversal: two-level 4-wide BVH, one lane per ray, near-to-far child order, triangle
// leaves tested on the spot, a short per-lane stack in LDS that spills its oldest entries to
// global memory, and the watertight ray/triangle test of Woop, Benthin and Wald (2013).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include <hipcub/hipcub.hpp>

#include "device.h"

namespace ptk {

__constant__ uint32_t c_sobol_bytes[2][4][256];
__constant__ float c_srgb[256];

hipError_t upload_sobol_table()
{
    // Joe-Kuo direction numbers for dimensions 0-3 (polynomials 1, x+1, x^2+x+1, x^3+x+1).
    uint32_t v[4][32];
    for (int b = 0; b < 32; b++)
        v[0][b] = 1u << (31 - b);
    const int deg[3] = {1, 2, 3};
    const uint32_t poly[3] = {0, 1, 1};
    const uint32_t minit[3][3] = {{1, 0, 0}, {1, 3, 0}, {1, 3, 1}};
    for (int d = 1; d < 4; d++) {
        const int s = deg[d - 1];
        uint32_t m[32];
        for (int k = 0; k < s; k++)
            m[k] = minit[d - 1][k];
        for (int k = s; k < 32; k++) {
            uint32_t val = m[k - s] ^ (m[k - s] << s);
            for (int j = 1; j < s; j++)
                if ((poly[d - 1] >> (s - 1 - j)) & 1u)
                    val ^= m[k - j] << j;
            m[k] = val;
        }
        for (int k = 0; k < 32; k++)
            v[d][k] = m[k] << (31 - k);
    }
    // dimensions 2 and 3 as byte tables (the device evaluates dimension 1 in closed form)
    static uint32_t bytes[2][4][256];
    for (int d = 2; d < 4; d++)
        for (int c = 0; c < 4; c++)
            for (int x = 0; x < 256; x++) {
                uint32_t y = 0;
                for (int b = 0; b < 8; b++)
                    if ((x >> b) & 1)
                        y ^= v[d][8 * c + b];
                bytes[d - 2][c][x] = y;
            }
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_sobol_bytes), bytes, sizeof(bytes));
    if (e != hipSuccess)
        return e;
    // sRGB transfer function of 8-bit texels (the decode VK_FORMAT_R8G8B8A8_SRGB applies)
    float lut[256];
    for (int i = 0; i < 256; i++) {
        const double c = i / 255.0;
        lut[i] = float(c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4));
    }
    return hipMemcpyToSymbol(HIP_SYMBOL(c_srgb), lut, sizeof(lut));
}

namespace {

constexpr uint32_t NONE = 0xFFFFFFFFu;

// --- traversal -----------------------------------------------------------------------------

struct RayPre {
    float3 o, d, inv;
    float3 oi;    // o x inv (rounded): slab distances as one fused multiply-add, fma(p, inv, -oi)
    float slack;  // absolute error bound of a near/far pair of those distances
    int kx, ky, kz;
    float sx, sy, sz;
};

__device__ __forceinline__ float comp(float3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// Hardware reciprocal (v_rcp_f32, within 1 ulp) instead of IEEE division (about ten
// instructions each). A ray's constants below only have to be the same for every box and
// triangle it meets (the watertight test shears all vertices with the same sx, sy, sz), not
// correctly rounded; the slab test's conservative factor covers the extra ulp of 1/d.
__device__ __forceinline__ float rcp1(float x) { return __builtin_amdgcn_rcpf(x); }

__device__ __forceinline__ void precompute(RayPre& r, float3 o, float3 d)
{
    r.o = o;
    r.d = d;
    const float ex = fabsf(d.x) < 1e-30f ? copysignf(1e-30f, d.x) : d.x;
    const float ey = fabsf(d.y) < 1e-30f ? copysignf(1e-30f, d.y) : d.y;
    const float ez = fabsf(d.z) < 1e-30f ? copysignf(1e-30f, d.z) : d.z;
    r.inv = f3(rcp1(ex), rcp1(ey), rcp1(ez));
    r.oi = f3(o.x * r.inv.x, o.y * r.inv.y, o.z * r.inv.z);
    // fma(p, inv, -oi) = (p - o) inv within 3u relative (the factor below covers it) plus u x |oi|
    // absolute (oi's rounding, u = 2^-24); a pair within 2u max|oi|, the slack is twice that
    r.slack = fmaxf(fmaxf(fabsf(r.oi.x), fabsf(r.oi.y)), fabsf(r.oi.z)) * 0x1p-22f;
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = 0;
    if (ay > ax)
        kz = 1;
    if (az > fmaxf(ax, ay))
        kz = 2;
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    if (comp(d, kz) < 0.0f) {
        const int t = kx;
        kx = ky;
        ky = t;
    }
    r.kx = kx;
    r.ky = ky;
    r.kz = kz;
    r.sz = rcp1(comp(d, kz));
    r.sx = comp(d, kx) * r.sz;
    r.sy = comp(d, ky) * r.sz;
}

// Element i of a scene array by a 32-bit byte offset from its (uniform) base: the loads take
// the scalar-base form, with no 64-bit address arithmetic per lane. The scene validation keeps
// the node and triangle-record arrays under 4 GiB.
template <typename T>
__device__ __forceinline__ const T& at32(const T* base, uint32_t i)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + i * uint32_t(sizeof(T)));
}

// Watertight test; on a hit in [tmin, tmax] returns t and the barycentric weights of v1, v2.
__device__ __forceinline__ bool intersect_tri(const RayPre& r, float3 v0, float3 v1, float3 v2, float tmin, float tmax,
                                              float& t, float& b1, float& b2)
{
    const float3 A = v0 - r.o, B = v1 - r.o, C = v2 - r.o;
    const float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
    const float Ax = comp(A, r.kx) - r.sx * Akz, Ay = comp(A, r.ky) - r.sy * Akz;
    const float Bx = comp(B, r.kx) - r.sx * Bkz, By = comp(B, r.ky) - r.sy * Bkz;
    const float Cx = comp(C, r.kx) - r.sx * Ckz, Cy = comp(C, r.ky) - r.sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {  // edge case: recompute the edge functions in fp64
        U = float(double(Cx) * double(By) - double(Cy) * double(Bx));
        V = float(double(Ax) * double(Cy) - double(Ay) * double(Cx));
        W = float(double(Bx) * double(Ay) - double(By) * double(Ax));
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f))
        return false;
    const float det = U + V + W;
    if (det == 0.0f)
        return false;
    const float T = U * (r.sz * Akz) + V * (r.sz * Bkz) + W * (r.sz * Ckz);
    const float rcp = rcp1(det);
    t = T * rcp;
    if (!(t >= tmin && t < tmax))
        return false;
    b1 = V * rcp;
    b2 = W * rcp;
    return true;
}

// Rows of a lane's LDS column after its stack ring: state of the ray that traversal reads
// rarely, kept out of registers (the ray in world space, read when an instance is entered or
// left; the closest hit's barycentrics, instance and primitive, written when a triangle is hit).
constexpr uint32_t COLD_ROWS = 10;
enum ColdRow : uint32_t { COLD_WO = 0, COLD_WD = 3, COLD_U = 6, COLD_V = 7, COLD_INST = 8, COLD_PRIM = 9 };

// Per-lane traversal stack: the newest D entries in an LDS ring (this lane's column of the
// block's array of D + COLD_ROWS rows), older ones spilled to global memory.
template <uint32_t D>
struct StackT {
    static_assert((D & (D - 1)) == 0, "LDS stack depth is a power of two");
    uint32_t* lds;    // this lane's column: entry k at lds[k * TRACE_BLOCK]
    uint32_t* spill;  // this thread's spill area: entry k at spill[k * spill_stride]
    uint32_t spill_stride;
    uint32_t spill_max;  // spill entries in use (at most SPILL_STACK)
    uint32_t sp;
    bool overflow;

    __device__ __forceinline__ void push(uint32_t v)
    {
        if (sp >= D) {
            const uint32_t deep = sp - D;
            if (deep >= spill_max) {
                overflow = true;
                return;
            }
            spill[size_t(deep) * spill_stride] = lds[(sp & (D - 1)) * TRACE_BLOCK];
        }
        lds[(sp & (D - 1)) * TRACE_BLOCK] = v;
        sp++;
    }
    __device__ __forceinline__ uint32_t pop()
    {
        sp--;
        const uint32_t slot = (sp & (D - 1)) * TRACE_BLOCK;
        const uint32_t v = lds[slot];
        if (sp >= D)
            lds[slot] = spill[size_t(sp - D) * spill_stride];
        return v;
    }
    __device__ __forceinline__ void put(uint32_t row, uint32_t v) { lds[(D + row) * TRACE_BLOCK] = v; }
    __device__ __forceinline__ uint32_t get(uint32_t row) const { return lds[(D + row) * TRACE_BLOCK]; }
    __device__ __forceinline__ void put3(uint32_t row, float3 v)
    {
        put(row, float_as_u32(v.x));
        put(row + 1, float_as_u32(v.y));
        put(row + 2, float_as_u32(v.z));
    }
    __device__ __forceinline__ float3 get3(uint32_t row) const
    {
        return f3(u32_as_float(get(row)), u32_as_float(get(row + 1)), u32_as_float(get(row + 2)));
    }
};
using Stack = StackT<LDS_STACK>;

struct HitOut {
    float t, u, v;
    uint32_t inst, prim;
};

// Fetch counts of one lane's rays (COUNT = true): node lines, triangle records, node lines of
// the TLAS, instances entered.
struct FetchCount {
    uint32_t nodes = 0, tris = 0, tlas_nodes = 0, entries = 0;
};

// One ray's traversal: closest hit (ANY = false) or any hit (ANY = true) in [tmin, tmax]. For
// any-hit rays the transmittance through transparent instances is kept in trans
// (shadowray.ah.hlsl); ALPHA = false compiles that out for scenes without transparent instances
// (every hit occludes). start() sets a ray up; each step() visits one internal node (testing
// the triangle leaves among its children at once) and picks the next one, entering and leaving
// instances, until it returns true: the traversal is over, found() tells whether it hit and
// result() gives the hit. The world-space ray and the closest hit's attributes live in the
// lane's LDS rows (ColdRow); the any-hit record is only made when the ray ends. LEAF1 compiles
// the triangle loop of a leaf out for hierarchies whose BLAS leaves each hold one triangle (the
// host validation finds the largest leaf).
template <bool ANY, bool COUNT, typename Stk, bool ALPHA, bool LEAF1 = false>
struct Walker {
    RayPre r;
    float tmin, best, opacity, trans;
    uint32_t inst, blas_sp, node;
    HitOut hit;  // any hit

    // false: nothing to traverse (an empty scene); the ray is then finished without a hit
    __device__ __forceinline__ bool start(const DeviceScene& s, float3 o, float3 d, float t0, float t1, Stk& stk)
    {
        hit.inst = NONE;
        trans = 1.0f;
        stk.put3(COLD_WO, o);
        stk.put3(COLD_WD, d);
        if (!ANY)
            stk.put(COLD_INST, NONE);
        tmin = t0;
        best = t1;
        inst = NONE;
        blas_sp = NONE;
        opacity = 1.0f;
        node = s.tlas_root;
        stk.sp = 0;
        if (node == CHILD_EMPTY)
            return false;
        precompute(r, o, d);
        return true;
    }

    __device__ __forceinline__ bool found(const Stk& stk) const
    {
        return (ANY ? hit.inst : stk.get(COLD_INST)) != NONE;
    }
    __device__ __forceinline__ HitOut result(const Stk& stk) const
    {
        if (ANY)
            return hit;
        return HitOut{best, u32_as_float(stk.get(COLD_U)), u32_as_float(stk.get(COLD_V)), stk.get(COLD_INST),
                      stk.get(COLD_PRIM)};
    }

    __device__ __forceinline__ bool step(const DeviceScene& s, Stk& stk, FetchCount* fc)
    {
        // Conservative slab far distance: each slab distance fma(p, inv, -oi), inv within 1 ulp, is
        // within 3u of exact (u = 2^-24) plus u x |oi| (RayPre::slack), so a near/far pair within
        // 6u relative plus the slack; 1 + 2^-20 = 1 + 16u.
        const float cons = 1.00000095367431640625f;
        // ---- visit an internal node
        if (COUNT) {
            fc->nodes++;
            fc->tlas_nodes += inst == NONE ? 1u : 0u;
        }
        // The near and far plane rows of each axis are picked by the load address (the ray's
        // direction sign: row hi first when the ray runs towards -axis), from a 32-bit offset off
        // the scene's node array.
        const char* nb = reinterpret_cast<const char*>(s.nodes);
        const uint32_t nbase = node * uint32_t(sizeof(Node4));
        const uint32_t ox = (float_as_u32(r.inv.x) >> 31) << 4, oy = 32u + ((float_as_u32(r.inv.y) >> 31) << 4),
                       oz = 64u + ((float_as_u32(r.inv.z) >> 31) << 4);
        auto row = [&](uint32_t off) { return *reinterpret_cast<const float4*>(nb + (nbase + off)); };
        const float4 nx = row(ox), fx = row(ox ^ 16u), ny = row(oy), fy = row(oy ^ 16u), nz = row(oz), fz = row(oz ^ 16u);
        const uint4 ch = *reinterpret_cast<const uint4*>(nb + (nbase + 96u));
        float tn[4];
        uint32_t cc[4] = {ch.x, ch.y, ch.z, ch.w};
        const float nxa[4] = {nx.x, nx.y, nx.z, nx.w}, fxa[4] = {fx.x, fx.y, fx.z, fx.w};
        const float nya[4] = {ny.x, ny.y, ny.z, ny.w}, fya[4] = {fy.x, fy.y, fy.z, fy.w};
        const float nza[4] = {nz.x, nz.y, nz.z, nz.w}, fza[4] = {fz.x, fz.y, fz.z, fz.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float t0 = fmaxf(fmaxf(fmaf(nxa[k], r.inv.x, -r.oi.x), fmaf(nya[k], r.inv.y, -r.oi.y)),
                                   fmaxf(fmaf(nza[k], r.inv.z, -r.oi.z), tmin));
            const float t1 = fmaf(fminf(fminf(fmaf(fxa[k], r.inv.x, -r.oi.x), fmaf(fya[k], r.inv.y, -r.oi.y)),
                                        fminf(fmaf(fza[k], r.inv.z, -r.oi.z), best)),
                                  cons, r.slack);
            const bool h = cc[k] != CHILD_EMPTY && t0 <= t1;
            tn[k] = h ? t0 : INFINITY;
            if (!h)
                cc[k] = NONE;
        }
        // sort the (up to) four hits near to far (optimal 5-comparator network)
        auto cmpx = [&](int a, int b) {
            if (ANY) {  // selects (measured: any-hit rays -2.5 %, closest-hit rays +1.3 %)
                const bool sw = tn[b] < tn[a];
                const float lo = fminf(tn[a], tn[b]), hi = fmaxf(tn[a], tn[b]);
                const uint32_t ca = cc[a], cb = cc[b];
                tn[a] = lo;
                tn[b] = hi;
                cc[a] = sw ? cb : ca;
                cc[b] = sw ? ca : cb;
            } else if (tn[b] < tn[a]) {
                const float tt = tn[a];
                tn[a] = tn[b];
                tn[b] = tt;
                const uint32_t cu = cc[a];
                cc[a] = cc[b];
                cc[b] = cu;
            }
        };
        cmpx(0, 1);
        cmpx(2, 3);
        cmpx(0, 2);
        cmpx(1, 3);
        cmpx(1, 2);
        // triangle leaves first, near to far (they shorten `best` for the boxes after them)
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t c = cc[k];
            if (c == NONE || tn[k] > best)
                continue;
            if ((c & CHILD_LEAF) && inst != NONE) {
                // triangle leaf of the current BLAS
                const uint32_t first = c & LEAF_FIRST_MASK;
                const uint32_t count = LEAF1 ? 1u : ((c >> 27) & 0xFu) + 1u;
                for (uint32_t q = 0; q < count; q++) {
                    if (COUNT)
                        fc->tris++;
                    const TriRec tr = at32(s.tris, first + q);
                    float t, b1, b2;
                    if (!intersect_tri(r, xyz(tr.v0), xyz(tr.v1), xyz(tr.v2), tmin, best, t, b1, b2))
                        continue;
                    if (ANY) {
                        if (ALPHA && opacity < 1.0f) {
                            float alpha = opacity;
                            if (alpha < 0.0f) {
                                // base-colour map alpha at the hit (shadowray.ah.hlsl:21-33)
                                const InstanceData& id = s.instances[inst];
                                const uint32_t pr = float_as_u32(tr.v0.w), ix = id.index_index + pr * 3;
                                const float2 ua = vertex_uv(s, id.vertex_index + s.indices[ix]);
                                const float2 ub = vertex_uv(s, id.vertex_index + s.indices[ix + 1]);
                                const float2 uc = vertex_uv(s, id.vertex_index + s.indices[ix + 2]);
                                const float wa = 1.0f - b1 - b2;
                                const float2 uv = make_float2(wa * ua.x + b1 * ub.x + b2 * uc.x,
                                                              wa * ua.y + b1 * ub.y + b2 * uc.y);
                                alpha = material_base_colour(s, s.materials[id.material_index], uv).w;
                            }
                            trans *= (1.0f - alpha);
                            if (trans < 1e-5f) {
                                hit = HitOut{t, b1, b2, inst, float_as_u32(tr.v0.w)};
                                return true;
                            }
                            continue;
                        }
                        hit = HitOut{t, b1, b2, inst, float_as_u32(tr.v0.w)};
                        return true;  // occluded
                    }
                    best = t;
                    stk.put(COLD_U, float_as_u32(b1));
                    stk.put(COLD_V, float_as_u32(b2));
                    stk.put(COLD_INST, inst);
                    stk.put(COLD_PRIM, float_as_u32(tr.v0.w));
                }
            }
        }
        // then the nodes (or TLAS leaves) still in reach, far to near: every one but the
        // nearest goes on the stack, so they are popped near to far
        uint32_t next = NONE;
#pragma unroll
        for (int k = 3; k >= 0; k--) {
            const uint32_t c = cc[k];
            if (c == NONE || tn[k] > best || ((c & CHILD_LEAF) && inst != NONE))
                continue;
            if (next != NONE)
                stk.push(next);
            next = c;
        }

        // ---- choose the next internal node (entering instances as needed)
        while (true) {
            if (next == NONE) {
                if (stk.sp == blas_sp) {  // the current BLAS is finished: back to world space
                    inst = NONE;
                    blas_sp = NONE;
                    precompute(r, stk.get3(COLD_WO), stk.get3(COLD_WD));
                }
                if (stk.sp == 0)
                    return true;  // nothing left to visit
                next = stk.pop();
            }
            if ((next & CHILD_LEAF) && inst == NONE) {  // TLAS leaf: enter the instance
                const uint2 entry = at32(s.tlas_entries, next & ~CHILD_LEAF);
                const uint32_t id = entry.x, root = entry.y;
                next = NONE;
                if (root == CHILD_EMPTY)
                    continue;
                const InstXform x = at32(s.inst_inv, id);
                const float3 wo = stk.get3(COLD_WO), wd = stk.get3(COLD_WD);
                const float3 oo = f3(x.r0.x * wo.x + x.r0.y * wo.y + x.r0.z * wo.z + x.r0.w,
                                     x.r1.x * wo.x + x.r1.y * wo.y + x.r1.z * wo.z + x.r1.w,
                                     x.r2.x * wo.x + x.r2.y * wo.y + x.r2.z * wo.z + x.r2.w);
                const float3 od = f3(x.r0.x * wd.x + x.r0.y * wd.y + x.r0.z * wd.z,
                                     x.r1.x * wd.x + x.r1.y * wd.y + x.r1.z * wd.z,
                                     x.r2.x * wd.x + x.r2.y * wd.y + x.r2.z * wd.z);
                precompute(r, oo, od);
                inst = id;
                blas_sp = stk.sp;
                if (COUNT)
                    fc->entries++;
                if (ANY && ALPHA)
                    opacity = s.inst_opacity[id];
                // continue with the entered BLAS node (its root, or a subtree of it) as an internal node
                next = root;
            }
            break;
        }
        node = next;
        return false;
    }
};

template <bool ANY, bool COUNT = false, typename Stk, bool ALPHA = true, bool LEAF1 = false>
__device__ bool traverse(const DeviceScene& s, float3 wo, float3 wd, float tmin, float tmax, Stk& stk, HitOut& hit,
                         float& trans, FetchCount* fc = nullptr)
{
    Walker<ANY, COUNT, Stk, ALPHA, LEAF1> w;
    if (w.start(s, wo, wd, tmin, tmax, stk))
        while (!w.step(s, stk, fc)) {
        }
    hit = w.result(stk);
    trans = w.trans;
    return w.found(stk);
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// A batch's rays carry their scene; queue entries of one view, and so of one scene, stay
// mostly together. True when every active lane of the wave has the scene of the first active
// lane: the callers then index the scene table by that wave-uniform (scalar) value, so the
// scene record is read with scalar loads and held in SGPRs for the whole traversal, as for a
// single scene.
__device__ __forceinline__ bool same_scene(uint32_t lane_sid, uint32_t first)
{
    return __ballot(lane_sid != first) == 0;
}

// Positions in NC queues for the threads of a block that want one, with one atomic per
// queue and block: device-scope atomics on one address serialise across the whole chip, so
// the shade and refill stages aggregate over the block instead of the wave. Every thread of
// the block must call it (block-uniform control flow).
template <int NC>
__device__ __forceinline__ void block_append(const bool (&want)[NC], uint32_t* const (&counter)[NC],
                                             uint32_t (&pos)[NC])
{
    __shared__ uint32_t s_wave[NC][TRACE_BLOCK / 64];
    __shared__ uint32_t s_base[NC];
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
    unsigned long long mask[NC];
#pragma unroll
    for (int k = 0; k < NC; k++) {
        mask[k] = __ballot(want[k]);
        if (lane == 0)
            s_wave[k][wave] = uint32_t(__popcll(mask[k]));
    }
    __syncthreads();
    if (threadIdx.x < NC) {
        const int k = int(threadIdx.x);
        uint32_t total = 0;
        for (uint32_t w = 0; w < waves; w++) {
            const uint32_t c = s_wave[k][w];
            s_wave[k][w] = total;
            total += c;
        }
        s_base[k] = total ? atomicAdd(counter[k], total) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NC; k++)
        pos[k] = s_base[k] + s_wave[k][wave] + uint32_t(__popcll(mask[k] & ((1ull << lane) - 1ull)));
    __syncthreads();  // the next call reuses s_wave / s_base
}

__device__ __forceinline__ uint32_t block_append1(bool want, uint32_t* counter)
{
    uint32_t pos[1];
    block_append<1>({want}, {counter}, pos);
    return pos[0];
}

// --- kernels -------------------------------------------------------------------------------

// Segment of a job path: the last segment whose first path is <= p.
__device__ __forceinline__ uint32_t find_segment(const JobDev* job, uint32_t p)
{
    uint32_t lo = 0, hi = job->nseg;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (job->seg[mid].path_offset <= p)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Starts path p of the job in slot `slot` as entry pos of queue par: camera ray of its view
// (pathtrace.rg.hlsl:18-33) and a fresh payload.
// Pixel of the u-th path of a sample: paths run through 8x8-pixel tiles (64 consecutive
// paths, one wave, share a tile), tiles left to right within bands of 8 rows, bands top to
// bottom; a band's last tile is narrower and the last band shorter when the image sides are
// not multiples of 8. A bijection of [0, width x height) onto the pixels.
__device__ __forceinline__ uint2 tile_pixel(uint32_t u, uint32_t width, uint32_t height)
{
    const uint32_t band = u / (8u * width);
    const uint32_t rows = min(8u, height - 8u * band);
    const uint32_t rem = u - band * 8u * width;
    const uint32_t col = rem / (rows * 8u);
    const uint32_t cw = min(8u, width - 8u * col);
    const uint32_t within = rem - col * rows * 8u;
    return make_uint2(8u * col + within % cw, 8u * band + within / cw);
}

__device__ __forceinline__ void start_path(const ViewDev* views, const JobDev* job, uint32_t slot, const Wavefront& wf,
                                           int par, uint32_t pos, uint32_t p, float4* ro = nullptr,
                                           float4* rd = nullptr)
{
    const SegDev sg = job->seg[find_segment(job, p)];
    const ViewDev& vw = views[sg.view];
    const uint32_t local = p - sg.path_offset;
    // pixel-major: the segment's samples of one pixel are consecutive paths, pixels in the
    // tile order, so a wave traces nearly the same ray 64 times (or a few pixels of one tile
    // when the segment has fewer samples)
    const uint32_t sample = sg.first_sample + local % sg.samples;
    const uint2 pxy = tile_pixel(local / sg.samples, vw.width, vw.height);
    const uint32_t px = pxy.x, py = pxy.y;
    Sampler smp;
    smp.start(px, py, sample, 0);
    const float2 up = smp.s2();
    const float2 ul = smp.s2();
    float3 o, d;
    camera_ray(vw.cam, px, py, vw.width, vw.height, up, ul, o, d);
    const float4 o4 = make_float4(o.x, o.y, o.z, vw.cam.near_clip), d4 = make_float4(d.x, d.y, d.z, vw.cam.far_clip);
    wf.q_o[par][pos] = o4;
    wf.q_d[par][pos] = d4;
    if (ro) {
        *ro = o4;
        *rd = d4;
    }
    // segment 0; its throughput (1, pdf -1) and radiance (0, no training record) are implied,
    // not written: the shade stage reads them only for later segments
    wf.q_ps[par][pos] = make_uint4(p, (py << 16) | px, sample, (sg.view << 16) | (slot << 15));
    wf.q_scene[par][pos] = vw.scene;
}

// Which new paths an iteration's refill appends to the next queue (behind its n_cont
// continuing entries): the older job slot's first, then the other's. k_refill_commit stores the
// plan at CTR(13) when it advances the counters; the entries are then written by k_refill, or,
// in a pipelined iteration, by the next iteration's traversal launch, which traces them at once.
struct RefillPlan {
    uint32_t n_cont, first;   // continuing entries, the older slot
    uint32_t next[2], n[2];   // per slot: first new path id, new paths
};

__device__ __forceinline__ RefillPlan stored_plan(const Wavefront& wf)
{
    const uint32_t* w = wf.counters + CTR(13);
    RefillPlan r;
    r.n_cont = w[0];
    r.first = w[1];
    r.next[0] = w[2];
    r.next[1] = w[3];
    r.n[0] = w[4];
    r.n[1] = w[5];
    return r;
}

// Starts the new path planned for entry pos (>= n_cont) of queue par.
__device__ __forceinline__ void issue_planned(const ViewDev* views, const JobDev* j0, const JobDev* j1,
                                              const RefillPlan& r, const Wavefront& wf, int par, uint32_t pos,
                                              float4* ro = nullptr, float4* rd = nullptr)
{
    // (selects, not indexing by the slot: the plan stays in registers)
    const uint32_t k = pos - r.n_cont;
    const uint32_t n_first = r.first ? r.n[1] : r.n[0];
    const bool older = k < n_first;
    const uint32_t s = older ? r.first : 1u - r.first;
    const uint32_t p = (s ? r.next[1] : r.next[0]) + (older ? k : k - n_first);
    start_path(views, s ? j1 : j0, s, wf, par, pos, p, ro, rd);
}

// Queue entries of a traversal launch in chunks of 64 consecutive entries, which the waves of
// each XCD take in order from their XCD's counter (zeroed before the launch; blocks are dealt
// to the 8 XCDs round-robin). Chunk k of an XCD is its k-th chunk of the rounds of the
// launch: every round of gridDim.x x TRACE_BLOCK entries gives each XCD one contiguous share,
// so neighbouring rays share an XCD's L2 and all XCDs trace one window of the queue at a time
// (shared BVH nodes in the Infinity Cache), while a wave that drew slow rays (grazing ones
// visit hundreds of nodes) takes fewer chunks instead of holding up the end of the launch.
// Each XCD's chunks cover its shares exactly once, and every wave leaves once its XCD's next
// chunk starts past the count.
template <typename F>
__device__ __forceinline__ void for_each_chunk(uint32_t* ctr, uint32_t count, F&& body)
{
    const uint32_t grid = gridDim.x;
    const bool split = grid % 8u == 0u;
    const uint32_t xcd = split ? blockIdx.x % 8u : 0u;
    const uint32_t share = (split ? grid / 8u : grid) * (TRACE_BLOCK / 64u);  // chunks of an XCD per round
    const uint32_t round = grid * TRACE_BLOCK;
    uint32_t* const c = ctr + xcd * 16u;  // the XCDs' counters 64 bytes apart
    auto first = [&](uint32_t k) { return (k / share) * round + (xcd * share + k % share) * 64u; };
    const uint32_t lane = lane_id();
    uint32_t k = 0;
    if (lane == 0)
        k = atomicAdd(c, 1u);
    uint32_t base = first(__builtin_amdgcn_readfirstlane(k));
    while (base < count) {
        uint32_t next = 0;
        if (lane == 0)
            next = atomicAdd(c, 1u);  // the next chunk, fetched while this one is traced
        const uint32_t i = base + lane;
        if (i < count)
            body(i);
        base = first(__builtin_amdgcn_readfirstlane(next));
    }
}

// Closest hits of the extension rays of queue q (count at CTR(q); chunk counters at CTR(10)).
// With `issue` (the views and job slots given), the entries the stored refill plan appended are
// started here: their camera rays are written for the shade stage and traced without being read.
template <bool COUNT, bool MULTI, uint32_t D, bool LEAF1>
__device__ __forceinline__ void extend_queue(const DeviceScene& s0, const DeviceScene* scenes, const Wavefront& wf,
                                             int q, StackT<D>& stk, FetchCount& fc, const ViewDev* views = nullptr,
                                             const JobDev* j0 = nullptr, const JobDev* j1 = nullptr)
{
    const uint32_t count = wf.counters[CTR(q)];
    const RefillPlan plan = views ? stored_plan(wf) : RefillPlan{count, 0, {0, 0}, {0, 0}};
    for_each_chunk(&wf.counters[CTR(10)], count, [&](uint32_t i) {
        float4 o, d;
        if (i >= plan.n_cont)
            issue_planned(views, j0, j1, plan, wf, q, i, &o, &d);
        else
            o = wf.q_o[q][i], d = wf.q_d[q][i];
        HitOut h;
        float trans;
        const uint32_t sid = MULTI ? wf.q_scene[q][i] : 0u, first = __builtin_amdgcn_readfirstlane(sid);
        if (MULTI && !same_scene(sid, first))
            traverse<false, COUNT, StackT<D>, true, LEAF1>(scenes[sid], xyz(o), xyz(d), o.w, d.w, stk, h, trans, &fc);
        else
            traverse<false, COUNT, StackT<D>, true, LEAF1>(MULTI ? scenes[first] : s0, xyz(o), xyz(d), o.w, d.w, stk, h,
                                                           trans, &fc);
        wf.hit[i] = make_float4(h.t, h.u, h.v, u32_as_float(h.inst));
        wf.hit_prim[i] = h.prim;
    });
}

// Any hits of the shadow rays the shade stage of parity qs queued (count at CTR(2); chunk
// counters at CTR(11)): an unoccluded ray adds its contribution, times the transmittance, where
// its path keeps its radiance (the next queue's entry, or the ended path's slot).
template <bool COUNT, bool MULTI, uint32_t D, bool ALPHA, bool LEAF1>
__device__ __forceinline__ void shadow_queue(const DeviceScene& s0, const DeviceScene* scenes, const Wavefront& wf,
                                             int qs, StackT<D>& stk, FetchCount& fc)
{
    const uint32_t count = wf.counters[CTR(2)];
    auto add = [&](uint32_t i, float trans) {
        const uint32_t tgt = wf.s_target[i];
        float4* dst = (tgt & TARGET_ENDED) ? &wf.rad_out[tgt & ~TARGET_ENDED] : &wf.q_rad[1 - qs][tgt];
        const float4 c = wf.s_c[i];
        float4 r = *dst;
        r.x += c.x * trans;
        r.y += c.y * trans;
        r.z += c.z * trans;
        *dst = r;
    };
    for_each_chunk(&wf.counters[CTR(11)], count, [&](uint32_t i) {
        const float4 o = wf.s_o[i], d = wf.s_d[i];
        HitOut h;
        float trans;
        const uint32_t sid = MULTI ? wf.s_scene[i] : 0u, first = __builtin_amdgcn_readfirstlane(sid);
        bool occluded;
        if (MULTI && !same_scene(sid, first))
            occluded = traverse<true, COUNT, StackT<D>, ALPHA, LEAF1>(scenes[sid], xyz(o), xyz(d), 0.0f, o.w, stk, h,
                                                                      trans, &fc);
        else
            occluded = traverse<true, COUNT, StackT<D>, ALPHA, LEAF1>(MULTI ? scenes[first] : s0, xyz(o), xyz(d), 0.0f,
                                                                      o.w, stk, h, trans, &fc);
        if (!occluded && trans >= 1e-5f)
            add(i, trans);
    });
}

// Traversal kernels keep at most 80 registers (6 waves per SIMD, a few bytes spilled): more
// waves in flight hide more of the dependent node fetches (measured: +3% over 5 waves; 8 waves
// spill too much). With a 32-entry LDS stack a block's LDS (42 KiB) lets only 3 waves per SIMD
// in, and the register budget is set for those.
constexpr int trace_waves(uint32_t d) { return d > 16 ? 3 : 6; }
template <bool COUNT, bool MULTI, uint32_t D, bool LEAF1>
__global__ __launch_bounds__(TRACE_BLOCK) __attribute__((amdgpu_waves_per_eu(trace_waves(D)))) void k_extend(DeviceScene s0, const DeviceScene* scenes, Wavefront wf, int q)
{
    __shared__ uint32_t lds[(D + COLD_ROWS) * TRACE_BLOCK];
    const uint32_t gtid = blockIdx.x * TRACE_BLOCK + threadIdx.x;
    StackT<D> stk{&lds[threadIdx.x], wf.spill + gtid, wf.spill_threads, wf.spill_depth, 0, false};
    FetchCount fc;
    extend_queue<COUNT, MULTI, D, LEAF1>(s0, scenes, wf, q, stk, fc);
    if (stk.overflow)
        atomicOr(&wf.counters[CTR(3)], 1u);
    if (COUNT && (fc.nodes | fc.tris)) {
        atomicAdd(&wf.totals[2], (unsigned long long)fc.nodes);
        atomicAdd(&wf.totals[3], (unsigned long long)fc.tris);
        atomicAdd(&wf.totals[6], (unsigned long long)fc.tlas_nodes);
        atomicAdd(&wf.totals[7], (unsigned long long)fc.entries);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        atomicAdd(&wf.totals[0], (unsigned long long)wf.counters[CTR(q)]);
}

template <bool COUNT, bool MULTI, uint32_t D, bool ALPHA, bool LEAF1>
__global__ __launch_bounds__(TRACE_BLOCK) __attribute__((amdgpu_waves_per_eu(trace_waves(D)))) void k_shadow(
    DeviceScene s0, const DeviceScene* scenes, Wavefront wf, int q)
{
    __shared__ uint32_t lds[(D + COLD_ROWS) * TRACE_BLOCK];
    const uint32_t gtid = blockIdx.x * TRACE_BLOCK + threadIdx.x;
    StackT<D> stk{&lds[threadIdx.x], wf.spill + gtid, wf.spill_threads, wf.spill_depth, 0, false};
    FetchCount fc;
    shadow_queue<COUNT, MULTI, D, ALPHA, LEAF1>(s0, scenes, wf, q, stk, fc);
    if (stk.overflow)
        atomicOr(&wf.counters[CTR(3)], 1u);
    if (COUNT && (fc.nodes | fc.tris)) {
        atomicAdd(&wf.totals[4], (unsigned long long)fc.nodes);
        atomicAdd(&wf.totals[5], (unsigned long long)fc.tris);
        atomicAdd(&wf.totals[8], (unsigned long long)fc.tlas_nodes);
        atomicAdd(&wf.totals[9], (unsigned long long)fc.entries);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        atomicAdd(&wf.totals[1], (unsigned long long)wf.counters[CTR(2)]);
}

// The previous iteration's shadow rays (when `consume`), then this iteration's extension rays,
// in one launch: waves that run out of shadow chunks go straight on to extension chunks, so the
// end of the shadow work leaves no part of the chip idle. The shadow rays add to the radiance
// of queue q's entries, which the extension part does not touch.
template <bool MULTI, uint32_t D, bool ALPHA, bool LEAF1>
__global__ __launch_bounds__(TRACE_BLOCK) __attribute__((amdgpu_waves_per_eu(trace_waves(D)))) void k_trace_both(
    DeviceScene s0, const DeviceScene* scenes, Wavefront wf, int q, const ViewDev* views, const JobDev* j0,
    const JobDev* j1, int carry)
{
    __shared__ uint32_t lds[(D + COLD_ROWS) * TRACE_BLOCK];
    const uint32_t gtid = blockIdx.x * TRACE_BLOCK + threadIdx.x;
    StackT<D> stk{&lds[threadIdx.x], wf.spill + gtid, wf.spill_threads, wf.spill_depth, 0, false};
    FetchCount fc;
    if (carry) {  // the previous iteration was pipelined: its shadow rays and its new paths
        shadow_queue<false, MULTI, D, ALPHA, LEAF1>(s0, scenes, wf, 1 - q, stk, fc);
        extend_queue<false, MULTI, D, LEAF1>(s0, scenes, wf, q, stk, fc, views, j0, j1);
    } else {
        extend_queue<false, MULTI, D, LEAF1>(s0, scenes, wf, q, stk, fc);
    }
    if (stk.overflow)
        atomicOr(&wf.counters[CTR(3)], 1u);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        atomicAdd(&wf.totals[0], (unsigned long long)wf.counters[CTR(q)]);
        if (carry)
            atomicAdd(&wf.totals[1], (unsigned long long)wf.counters[CTR(2)]);
    }
}

// Empties the shadow queue once its rays are traced (count and the XCDs' chunk counters).
__global__ void k_clear_shadow(Wavefront wf)
{
    if (threadIdx.x < 8u)
        wf.counters[CTR(11) + threadIdx.x * 16u] = 0;
    if (threadIdx.x == 0)
        wf.counters[CTR(2)] = 0;
}

// Miss / closest-hit / Russian roulette of one segment for every entry of queue q. Continuing
// paths move, with their state, to the front of queue 1-q; an ended path's radiance goes to
// rad_out; next-event rays go to the shadow queue with the place their contribution is added.
// WPE: waves per SIMD the register allocation must allow (2: up to 256 registers per lane).
template <bool MULTI, int WPE, bool CACHE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE))) void k_shade(
    DeviceScene s0, const DeviceScene* scenes, const ViewDev* views,
                                               const JobDev* job, Wavefront wf, int q, CacheDev cd,
                                               const uint32_t* order)
{
    // CACHE = false: the plain tracer's variant, with the radiance cache's code compiled out
    const bool collect = CACHE && cd.collect;
    const uint32_t count = wf.counters[CTR(q)];
    // Whole blocks take part in the compaction. A block takes chunks of 256 entries from its
    // XCD's counter, the traversal launches' scheme by block (chunk k of an XCD is its share of
    // round k / share), so a block that drew hits (shading, next-event estimation) takes fewer
    // chunks than one that drew misses. Every block leaves once its XCD's next chunk starts
    // past the count.
    __shared__ uint32_t chunk_next;
    const uint32_t grid = gridDim.x;
    const bool split = grid % 8u == 0u;
    const uint32_t xcd = split ? blockIdx.x % 8u : 0u;
    const uint32_t share = split ? grid / 8u : grid;
    uint32_t* const chunk_ctr = &wf.counters[CTR(12) + xcd * 16u];
    auto first = [&](uint32_t k) { return ((k / share) * grid + xcd * share + k % share) * 256u; };
    if (threadIdx.x == 0)
        chunk_next = atomicAdd(chunk_ctr, 1u);
    __syncthreads();
    for (uint32_t k = chunk_next; first(k) < count; k = chunk_next) {
        __syncthreads();  // every thread has read chunk_next
        uint32_t k_after = 0;
        if (threadIdx.x == 0)
            k_after = atomicAdd(chunk_ctr, 1u);  // the next chunk, fetched while this one is shaded
        const uint32_t i = first(k) + threadIdx.x;
        const bool valid = i < count;
        bool cont = false, shadow = false;
        bool want_train = false, want_query = false;
        float qf = 1.0f;  // weight of the cache's prediction
        float cr[13];
        float3 cT = f3(0, 0, 0), cL = f3(0, 0, 0);
        uint4 ps = make_uint4(0, 0, 0, 0);  // path state of the entry
        float4 thr_v = make_float4(0, 0, 0, 0), rad_v = make_float4(0, 0, 0, 0);
        uint32_t rec = NONE;  // training record of the path
        float3 next_o = f3(0, 0, 0), next_d = f3(0, 0, 0);
        float3 sh_o = f3(0, 0, 0), sh_d = f3(0, 0, 0), sh_c = f3(0, 0, 0);
        float sh_tmax = 0.0f;
        uint32_t segment = 0;
        uint32_t sid = 0;
        if (valid) {
            const uint32_t qi = order ? order[i] : i;  // material-sorted order when given
            // Loads are grouped by the round trip that resolves their addresses, and every
            // group is issued before the data of the previous one are used: (1) the queue
            // entry (ray, path state, hit); (2) the hit instance; (3) view, material, triangle
            // indices and the light table's first level; (4) vertices and the chosen light bin;
            // then, for next-event estimation, (5-7) an emitting object's triangle table,
            // emitter instance, vertices and emission.
            ps = wf.q_ps[q][qi];  // job path, py << 16 | px, sample, view << 16 | segment
            const bool camera_ray = (ps.w & 0x7FFFu) == 0;  // a new path: throughput 1, no radiance yet
            const float4 tp = camera_ray ? make_float4(1.0f, 1.0f, 1.0f, -1.0f) : wf.q_thr[q][qi];
            const float4 rr = camera_ray ? make_float4(0.0f, 0.0f, 0.0f, u32_as_float(NONE)) : wf.q_rad[q][qi];
            const float4 hh = wf.hit[qi];
            const uint32_t prim = wf.hit_prim[qi];
            const float4 ro4 = wf.q_o[q][qi], rd4 = wf.q_d[q][qi];
            if (MULTI)
                sid = wf.q_scene[q][qi];
            const DeviceScene& s = MULTI ? scenes[sid] : s0;
            const uint32_t hinst = float_as_u32(hh.w);
            rec = float_as_u32(rr.w);
            const ViewDev& vw = views[ps.w >> 16];
            float3 T = xyz(tp);
            float brdf_pdf = tp.w;
            float3 L = xyz(rr);
            const float3 ro = xyz(ro4), rd = xyz(rd4);
            const uint32_t px = ps.y & 0xFFFFu, py = ps.y >> 16, sample = ps.z;
            segment = ps.w & 0x7FFFu;
            bool terminated = false;

            if (hinst == NONE) {
                L = L + T * s.envcol;  // pathtrace.ms.hlsl:7-11
                terminated = true;
            } else {
                // (2) the hit instance: vertex / index / emitter / material indices, normal matrix
                const float4* ip = reinterpret_cast<const float4*>(s.instances + hinst);
                const uint4 ih = *reinterpret_cast<const uint4*>(ip);
                const float4 nm0 = ip[5], nm1 = ip[6], nm2 = ip[7];
                Sampler smp;
                smp.start(px, py, sample, segment + 1);
                // draws in the order the path consumes them: the alpha test, then the light
                Sampler s_alpha = smp;
                const float u_alpha = s_alpha.s1();
                Sampler s_light = s_alpha;
                const float lu0 = s_light.s1();
                const float2 lu12 = s_light.s2();
                // (3)
                const uint32_t max_bounces = vw.max_bounces, vw_width = vw.width;
                const MaterialRec mat = s.materials[ih.w];
                const uint32_t ib = ih.y + prim * 3;
                const uint32_t ia = s.indices[ib], ibb = s.indices[ib + 1], ic = s.indices[ib + 2];
                const uint32_t lc = s.light_count;
                float lux = lu0;
                AliasBin<7> lbin;
                if (lc > 0)
                    lbin = load_alias_bin<7>(s.light + size_t(alias_pick(lc, lux)) * PT_LIGHT_BIN_WORDS);
                // (4)
                const float4* vp = s.vertices;
                const float4 a0 = vp[size_t(ih.x + ia) * 2], a1 = vp[size_t(ih.x + ia) * 2 + 1];
                const float4 b0 = vp[size_t(ih.x + ibb) * 2], b1 = vp[size_t(ih.x + ibb) * 2 + 1];
                const float4 c0 = vp[size_t(ih.x + ic) * 2], c1 = vp[size_t(ih.x + ic) * 2 + 1];
                if (lc > 0 && alias_follow(lbin.q, lux))
                    lbin = load_alias_bin<7>(s.light + size_t(lbin.alias) * PT_LIGHT_BIN_WORDS);
                const float t = hh.x, bu = hh.y, bv = hh.z;
                const bool final_seg = segment >= max_bounces;
                const float3 pa = xyz(a0), pb = xyz(b0), pc = xyz(c0);
                const float3 na = f3(a0.w, a1.x, a1.y), nb = f3(b0.w, b1.x, b1.y), nc = f3(c0.w, c1.x, c1.y);
                const float2 uva = make_float2(a1.z, a1.w), uvb = make_float2(b1.z, b1.w), uvc = make_float2(c1.z, c1.w);
                // get_vertex_attributes (geometry.hlsli:39-48)
                const float wa = 1.0f - bu - bv;
                const float3 obj_n = normalize3(wa * na + bu * nb + bv * nc);
                const float2 uv = make_float2(wa * uva.x + bu * uvb.x + bv * uvc.x, wa * uva.y + bu * uvb.y + bv * uvc.y);
                const float3 emission = material_emission(s, mat, uv);
                if (maxc3(emission) > 0.0f) {
                    // ObjectRayDirection(): the world->object transform applied to the direction
                    const InstXform x = s.inst_inv[hinst];
                    const float3 odir = f3(x.r0.x * rd.x + x.r0.y * rd.y + x.r0.z * rd.z,
                                           x.r1.x * rd.x + x.r1.y * rd.y + x.r1.z * rd.z,
                                           x.r2.x * rd.x + x.r2.y * rd.y + x.r2.z * rd.z);
                    if (dot3(odir, obj_n) < 0.0f) {
                        if (brdf_pdf < 0.0f) {
                            L = L + T * emission;
                        } else {
                            const InstanceData& in = s.instances[hinst];
                            const float3 wab = xform_vec(in, pb - pa);
                            const float3 wac = xform_vec(in, pc - pa);
                            const float3 sn = cross3(wab, wac);
                            const float area = sqrtf(dot3(sn, sn)) * 0.5f;
                            const float dist2 = t * t;
                            const float cosine = fabsf(dot3(rd, normalize3(sn)));
                            float pdf_l = light_pdf(s.light, ih.z, prim);
                            pdf_l *= dist2 / (area * cosine);
                            if (pdf_l > 0.0f && dist2 > 0.0f) {
                                const float w = brdf_pdf / (brdf_pdf + pdf_l);
                                L = L + T * (w * emission);
                            }
                        }
                    }
                }
                const float4 base = final_seg ? make_float4(0, 0, 0, 1) : material_base_colour(s, mat, uv);
                if (final_seg) {
                    terminated = true;
                } else if ((smp = s_alpha, base.w < u_alpha)) {
                    // stochastic transparency: continue straight through (ch.hlsl:279-283)
                    next_o = ro + rd * (t + ORIGIN_OFFSET);
                    next_d = rd;
                } else {
                    BrdfInput bi;
                    bi.base_colour = xyz(base);
                    float roughness;
                    float3 mapped_n;
                    material_rmn(s, mat, uv, roughness, bi.metalness, mapped_n);
                    const float rough_param = roughness;
                    bi.specular = material_specular(s, mat, uv);
                    bi.sheen = material_sheen(s, mat, uv);
                    const float anisotropy = mat.cc.z;
                    roughness = fmaxf(0.00001f, roughness * roughness);
                    const float aspect = sqrtf(1.0f - 0.9f * fabsf(anisotropy));
                    bi.alpha = make_float2(roughness / aspect, roughness * aspect);
                    if (anisotropy < 0.0f)
                        bi.alpha = make_float2(bi.alpha.y, bi.alpha.x);
                    // get_vertex_vectors (geometry.hlsli:50-68)
                    const float2 duv_ab = make_float2(uvb.x - uva.x, uvb.y - uva.y);
                    const float2 duv_ac = make_float2(uvc.x - uva.x, uvc.y - uva.y);
                    const float3 ab = pb - pa, ac = pc - pa;
                    const float3 true_n = cross3(ab, ac);
                    const float denom = duv_ab.x * duv_ac.y - duv_ac.x * duv_ab.y;
                    float3 tangent;
                    if (fabsf(denom) < 1e-5f) {
                        tangent = cross3(obj_n, fabsf(obj_n.y) > 1e-5f ? f3(1, 0, 0) : f3(0, 1, 0));
                    } else {
                        tangent = (duv_ac.y * ab - duv_ab.y * ac) / denom;
                        tangent = tangent - dot3(tangent, obj_n) * obj_n;
                    }
                    const float3 bitangent = cross3(obj_n, tangent);
                    float3 wx = normalize3(xform_rows(nm0, nm1, nm2, tangent));
                    float3 wy = normalize3(xform_rows(nm0, nm1, nm2, bitangent));
                    float3 wn = normalize3(xform_rows(nm0, nm1, nm2, obj_n));
                    if (dot3(wn, rd) > 0.0f) {
                        wx = -wx;
                        wy = -wy;
                        wn = -wn;
                    }
                    float3 twn = normalize3(xform_rows(nm0, nm1, nm2, true_n));
                    if (dot3(rd, twn) > 0.0f)
                        twn = -twn;
                    if (mapped_n.z < 0.999f) {
                        // normal map: the shading frame becomes (mx, my, mn) expressed in the
                        // interpolated frame (ch.hlsl:321-326)
                        float3 mx, my;
                        onb(mapped_n, mx, my);
                        const float3 fx = mx.x * wx + mx.y * wy + mx.z * wn;
                        const float3 fy = my.x * wx + my.y * wy + my.z * wn;
                        const float3 fn = mapped_n.x * wx + mapped_n.y * wy + mapped_n.z * wn;
                        wx = fx;
                        wy = fy;
                        wn = fn;
                    }
                    const float3 nrd = -rd;
                    const float3 o = normalize3(f3(dot3(wx, nrd), dot3(wy, nrd), dot3(wn, nrd)));
                    bool cached = false;
                    if (o.z > 0.0f && collect && segment == cd.segment) {
                        // The cache's vertex (the second by default): a training path records the cache input here
                        // (its target is completed when the path ends); any other path ends
                        // here with the cache's prediction of the radiance leaving it.
                        const uint32_t pix = py * vw_width + px;
                        const float3 hp = ro + rd * t;
                        cr[0] = hp.x, cr[1] = hp.y, cr[2] = hp.z;
                        cr[3] = -rd.x, cr[4] = -rd.y, cr[5] = -rd.z;
                        cr[6] = wn.x, cr[7] = wn.y, cr[8] = wn.z;
                        cr[9] = rough_param;
                        cr[10] = base.x, cr[11] = base.y, cr[12] = base.z;
                        cT = T;
                        cL = L;
                        if (hash_u32(pix ^ (sample * 0x9E3779B9u)) % cd.train_stride == 0) {
                            want_train = true;
                        } else if (cd.query && (cd.max_pdf <= 0.0f || brdf_pdf < cd.max_pdf)) {
                            want_query = true;
                            const float uc = float(hash_u32(hash_u32(pix ^ 0xA511E9B3u) + sample) >> 8) * 0x1p-24f;
                            if (uc < cd.cont_prob) {
                                // control variate: the path goes on with weight 1/p and the
                                // prediction is added with weight 1 - 1/p (unbiased)
                                qf = 1.0f - 1.0f / cd.cont_prob;
                                T = T * (1.0f / cd.cont_prob);
                            } else {
                                cached = true;
                                terminated = true;
                            }
                        }
                    }
                    if (o.z <= 0.0f) {
                        terminated = true;
                    } else if (!cached) {
                        const float3 xp = ro + rd * t + twn * ORIGIN_OFFSET;
                        // next-event estimation (ch.hlsl:159-225): the shadow ray is queued
                        if (lc > 0) {
                            smp = s_light;
                            // (5) an emitting object (payload: type 3, table offset, triangles,
                            // first instance, instances): its triangle table
                            const bool object_light = lbin.w[0] == 3u;
                            const uint32_t* etab = s.light + lbin.w[1];
                            float eux = lux;
                            AliasBin<3> ebin;
                            if (object_light)
                                ebin = load_alias_bin<3>(etab + size_t(alias_pick(lbin.w[2], eux)) * PT_EMITTER_BIN_WORDS);
                            // (6) the table's alias and the emitter instance
                            uint4 eh = make_uint4(0, 0, 0, 0);
                            float4 em0 = make_float4(0, 0, 0, 0), em1 = em0, em2 = em0;
                            if (object_light) {
                                if (alias_follow(ebin.q, eux))
                                    ebin = load_alias_bin<3>(etab + size_t(ebin.alias) * PT_EMITTER_BIN_WORDS);
                                const uint32_t einst = lbin.w[3] + min(lbin.w[4] - 1, uint32_t(eux * float(lbin.w[4])));
                                const float4* ep = reinterpret_cast<const float4*>(s.instances + einst);
                                eh = *reinterpret_cast<const uint4*>(ep);
                                em0 = ep[1], em1 = ep[2], em2 = ep[3];
                            }
                            // (7) the emitter triangle's vertices and emission
                            float4 ev0 = make_float4(0, 0, 0, 0), ev1 = ev0, ev2 = ev0;
                            float4 e_emission = ev0, e_rmnb = ev0;
                            if (object_light) {
                                ev0 = vp[size_t(eh.x + ebin.w[0]) * 2];
                                ev1 = vp[size_t(eh.x + ebin.w[1]) * 2];
                                ev2 = vp[size_t(eh.x + ebin.w[2]) * 2];
                                e_emission = s.materials[eh.w].emission;
                                e_rmnb = s.materials[eh.w].rmnb;
                            }
                            float3 light_dir;
                            float light_dist, lpdf;
                            bool is_delta;
                            const float3 le =
                                light_sample_at(s, lbin, ebin, eh, em0, em1, em2, ev0, ev1, ev2, e_emission, e_rmnb,
                                                lu12, xp, light_dir, light_dist, lpdf, is_delta);
                            if (maxc3(le) > 0.0f) {
                                float3 li = f3(dot3(wx, light_dir), dot3(wy, light_dir), dot3(wn, light_dir));
                                if (li.z > 0.0f) {
                                    li = normalize3(li);
                                    const float3 m = normalize3(li + o);
                                    float pdf_b;
                                    const float3 brdf = evaluate_brdf(bi, li, o, m, pdf_b);
                                    float3 contrib;
                                    if (is_delta) {
                                        contrib = le * brdf * li.z / lpdf;
                                    } else {
                                        const float w = lpdf / (lpdf + pdf_b);
                                        contrib = w * le * brdf * li.z / lpdf;
                                    }
                                    shadow = true;
                                    sh_o = xp;
                                    sh_d = light_dir;
                                    sh_tmax = light_dist - 1e-5f;
                                    sh_c = T * contrib;
                                }
                            }
                        }
                        float3 m;
                        float w, pdf;
                        const float3 dir_i = sample_direction(o, bi.alpha, bi.metalness, smp, m, w, pdf);
                        if (dir_i.z * o.z <= 0.0f) {
                            terminated = true;
                        } else {
                            float pdf_brdf;
                            const float3 brdf = evaluate_brdf(bi, dir_i, o, m, pdf_brdf);
                            const float3 thr = w * brdf * dir_i.z / pdf;
                            brdf_pdf = pdf_brdf;
                            T = T * thr;
                            next_o = xp;
                            next_d = normalize3(wx * dir_i.x + wy * dir_i.y + wn * dir_i.z);
                        }
                    }
                }
                // Russian roulette after segments >= 1 (pathtrace.rg.hlsl:53, pathtrace.hlsli:118-129)
                if (!terminated && segment >= 1) {
                    const float mt = maxc3(T);
                    const float u = smp.s1();
                    if (mt < 1.0f) {
                        const float qq = fmaxf(0.0f, 1.0f - mt);
                        if (u < qq)
                            terminated = true;
                        else
                            T = T * (1.0f / (1.0f - qq));
                    }
                }
            }
            cont = !terminated;
            thr_v = make_float4(T.x, T.y, T.z, brdf_pdf);
            rad_v = make_float4(L.x, L.y, L.z, 0.0f);
        }
        // a recorded path that ends here: its training record is completed after the shadow
        // rays of this iteration have added their share
        const bool ended = valid && !cont;
        const bool ended_rec = ended && (rec != NONE || want_train);
        const uint32_t slot = (ps.w >> 15) & 1u;  // job slot of the path
        // ended paths are counted per job slot: a job is complete when all its paths ended
        uint32_t apos[7] = {0, 0, 0, 0, 0, 0, 0};
        if (collect) {  // uniform: training records, cache queries and ended records are appended too
            block_append<7>({cont, shadow, want_train, want_query, ended_rec, ended && slot == 0u, ended && slot == 1u},
                            {&wf.counters[CTR(1 - q)], &wf.counters[CTR(2)], &cd.counters[0], &cd.counters[1],
                             &wf.counters[CTR(4)], &wf.counters[CTR(7)], &wf.counters[CTR(8)]},
                            apos);
        } else {
            uint32_t p4[4];
            block_append<4>({cont, shadow, ended && slot == 0u, ended && slot == 1u},
                            {&wf.counters[CTR(1 - q)], &wf.counters[CTR(2)], &wf.counters[CTR(7)],
                             &wf.counters[CTR(8)]},
                            p4);
            apos[0] = p4[0], apos[1] = p4[1];
        }
        // the path's radiance slot: its job slot's half of rad_out
        const uint32_t ring = slot * wf.rad_out_capacity + ps.x;
        const uint32_t pos = apos[0], spos = apos[1], tpos = apos[2], qpos = apos[3];
        if (want_train)
            rec = tpos < cd.record_capacity ? tpos : NONE;
        if (cont) {
            wf.q_o[1 - q][pos] = make_float4(next_o.x, next_o.y, next_o.z, 0.0f);
            wf.q_d[1 - q][pos] = make_float4(next_d.x, next_d.y, next_d.z, INFINITY);
            wf.q_ps[1 - q][pos] = make_uint4(ps.x, ps.y, ps.z, ps.w + 1u);  // next segment
            wf.q_thr[1 - q][pos] = thr_v;
            wf.q_rad[1 - q][pos] = make_float4(rad_v.x, rad_v.y, rad_v.z, u32_as_float(rec));
            if (MULTI)
                wf.q_scene[1 - q][pos] = sid;
        } else if (valid) {
            wf.rad_out[ring] = rad_v;
        }
        const uint32_t target = cont ? pos : (TARGET_ENDED | ring);
        if (ended_rec)
            wf.train_end[apos[4]] = make_uint2(ring, rec);
        if (shadow) {
            wf.s_target[spos] = target;
            wf.s_o[spos] = make_float4(sh_o.x, sh_o.y, sh_o.z, sh_tmax);
            wf.s_d[spos] = make_float4(sh_d.x, sh_d.y, sh_d.z, 0.0f);
            wf.s_c[spos] = make_float4(sh_c.x, sh_c.y, sh_c.z, 0.0f);
            if (MULTI)
                wf.s_scene[spos] = sid;
        }
        if (collect) {
            // training record: inputs, throughput and radiance so far (target filled at the end)
            if (want_train && tpos < cd.record_capacity) {
                float* r = cd.records + size_t(tpos) * MLP_REC_FLOATS;
                for (int k = 0; k < 13; k++)
                    r[k] = cr[k];
                r[13] = cT.x, r[14] = cT.y, r[15] = cT.z;
                r[16] = cL.x, r[17] = cL.y, r[18] = cL.z;
                r[19] = 0.0f;
            }
            if (want_query) {
                float* r = cd.queries + size_t(qpos) * MLP_REC_FLOATS;
                for (int k = 0; k < 13; k++)
                    r[k] = cr[k];
                r[13] = cT.x, r[14] = cT.y, r[15] = cT.z;
                r[19] = qf;
                cd.query_slot[qpos] = target;
            }
        }
        if (threadIdx.x == 0)
            chunk_next = k_after;
        __syncthreads();
    }
}

// Sort keys of the extension hits: (scene, material) of the hit instance; misses after every
// material, entries past the queue's end last. (A counting sort by material with one global
// atomic per distinct key and wave was tried instead of the radix sort: the few hot buckets
// serialise the atomics and it cost more than the shading it saved.)
template <bool MULTI>
__global__ __launch_bounds__(256) void k_hit_keys(DeviceScene s0, const DeviceScene* scenes, Wavefront wf, int q,
                                                  uint32_t n)
{
    const uint32_t count = wf.counters[CTR(q)];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t key = 0xFFFFFFFFu;
        if (i < count) {
            const uint32_t inst = float_as_u32(wf.hit[i].w);
            if (inst == NONE) {
                key = 0xFFFFFFFEu;
            } else {
                const uint32_t sid = MULTI ? wf.q_scene[q][i] : 0u;
                const DeviceScene& s = MULTI ? scenes[sid] : s0;
                key = (sid << 20) | min(s.instances[inst].material_index, 0xFFFFFu);
            }
        }
        wf.sort_keys[0][i] = key;
        wf.sort_vals[0][i] = i;
    }
}

// Ends an iteration: completes the training records of recorded paths that ended in it
// (their radiance is final once the shadow rays and cache queries have added theirs), then
// writes the new paths k_refill_commit planned behind the continuing ones in queue 1-q: the
// remaining paths of the older job slot first, then those of the newer one, consecutive path ids
// into consecutive entries, so no atomics.
__device__ __forceinline__ RefillPlan refill_plan(const JobDev* j0, const JobDev* j1, const Wavefront& wf, int q)
{
    RefillPlan r;
    r.n_cont = wf.counters[CTR(1 - q)];
    r.first = int(j1->seq - j0->seq) < 0 ? 1u : 0u;
    const JobDev* jb[2] = {j0, j1};
    uint32_t room = wf.capacity - r.n_cont;
    for (int k = 0; k < 2; k++) {
        const uint32_t s = k == 0 ? r.first : 1u - r.first;
        r.next[s] = wf.counters[CTR(5 + s)];
        const uint32_t paths = jb[s]->paths;  // a free slot has none
        r.n[s] = paths > r.next[s] ? min(room, paths - r.next[s]) : 0u;
        room -= r.n[s];
    }
    return r;
}

__global__ __launch_bounds__(256) void k_refill(const ViewDev* views, const JobDev* j0, const JobDev* j1, Wavefront wf,
                                                int q, CacheDev cd)
{
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    if (cd.collect) {
        const uint32_t ended = wf.counters[CTR(4)];
        for (uint32_t i = tid; i < ended; i += stride) {
            const uint2 e = wf.train_end[i];
            if (e.y == NONE)
                continue;
            // training target: radiance gathered after the 2nd-bounce vertex, divided by the
            // throughput that reached it
            const float4 lf = wf.rad_out[e.x];
            float* rec = cd.records + size_t(e.y) * MLP_REC_FLOATS;
            const float lv[3] = {lf.x, lf.y, lf.z};
            for (int c = 0; c < 3; c++) {
                const float tc = rec[13 + c];
                rec[13 + c] = tc > 0.0f ? (lv[c] - rec[16 + c]) / tc : 0.0f;
            }
            rec[19] = 1.0f;  // complete
        }
    }
    // the new paths k_refill_commit planned
    const RefillPlan r = stored_plan(wf);
    const uint32_t n_new = r.n[0] + r.n[1];
    for (uint32_t k = tid; k < n_new; k += stride)
        issue_planned(views, j0, j1, r, wf, 1 - q, r.n_cont + k);
}

// Plans the refill (the job's next paths behind the continuing ones, stored at CTR(13)),
// advances the counters by it and leaves the counts the host steers by (next queue, paths ended
// and paths started per slot) in wf.snapshot[q]. The planned entries are written by k_refill or
// by the next pipelined traversal launch.
__global__ void k_refill_commit(const JobDev* j0, const JobDev* j1, Wavefront wf, int q)
{
    if (threadIdx.x == 0) {
        const RefillPlan r = refill_plan(j0, j1, wf, q);
        uint32_t* w = wf.counters + CTR(13);
        w[0] = r.n_cont;
        w[1] = r.first;
        w[2] = r.next[0];
        w[3] = r.next[1];
        w[4] = r.n[0];
        w[5] = r.n[1];
        const uint32_t count = r.n_cont + r.n[0] + r.n[1];
        wf.counters[CTR(1 - q)] = count;
        wf.counters[CTR(5)] = r.next[0] + r.n[0];
        wf.counters[CTR(6)] = r.next[1] + r.n[1];
        uint32_t* snap = wf.snapshot + 8 * q;
        snap[0] = count;
        snap[1] = wf.counters[CTR(7)];
        snap[2] = wf.counters[CTR(8)];
        snap[3] = r.next[0] + r.n[0];
        snap[4] = r.next[1] + r.n[1];
        snap[5] = wf.counters[CTR(3)];  // error flags (traversal stack overflow)
    }
}

// The job in slot s has been accumulated: the slot issues nothing until it is loaded again,
// and its counts restart from zero.
__global__ void k_slot_free(JobDev* job, Wavefront wf, uint32_t s)
{
    if (threadIdx.x == 0) {
        job->paths = 0;
        wf.counters[CTR(5 + s)] = 0;
        wf.counters[CTR(7 + s)] = 0;
    }
}

// A job was loaded into slot s: no paths of it started or ended yet.
__global__ void k_slot_reset(Wavefront wf, uint32_t s)
{
    if (threadIdx.x == 0) {
        wf.counters[CTR(5 + s)] = 0;
        wf.counters[CTR(7 + s)] = 0;
    }
}

// Clears the queue counters an iteration is about to fill.
// Counters an iteration starts from. The shadow queue is kept (clear_shadow = 0) when it still
// holds the previous iteration's rays, which this iteration traces first.
__global__ void k_prepare(Wavefront wf, int q, CacheDev cd, int clear_shadow)
{
    if (threadIdx.x < 8u) {  // the XCDs' chunk counters of the extend, shadow and shade launches
        wf.counters[CTR(10) + threadIdx.x * 16u] = 0;
        wf.counters[CTR(11) + threadIdx.x * 16u] = 0;
        wf.counters[CTR(12) + threadIdx.x * 16u] = 0;
    }
    if (threadIdx.x == 0) {
        wf.counters[CTR(1 - q)] = 0;
        if (clear_shadow)
            wf.counters[CTR(2)] = 0;
        wf.counters[CTR(4)] = 0;
        if (cd.collect)
            cd.counters[1] = 0;
    }
}

// Running mean per pixel in sample order (pathtrace.rg.hlsl:61-67, one sample per dispatch),
// for every segment of the job. A block (one wave) takes 64 consecutive pixels of the tile
// order, whose paths are one contiguous run (each pixel's samples in a row): chunks of ACC_C
// samples of the 64 pixels are read with the lanes along the run (whole 128-byte lines) into
// LDS, and each lane then folds its pixel's samples in order.
constexpr uint32_t ACC_C = 8;

__global__ __launch_bounds__(64) void k_accumulate(const ViewDev* views, const JobDev* job, const float4* rad)
{
    __shared__ float4 buf[64 * (ACC_C + 1)];  // a pixel's chunk per row, padded against bank conflicts
    const uint32_t lane = threadIdx.x;
    const bool signed_radiance = job->signed_radiance != 0;
    for (uint32_t s = 0; s < job->nseg; s++) {
        const SegDev sg = job->seg[s];
        const ViewDev& vw = views[sg.view];
        const uint32_t groups = (vw.npix + 63u) / 64u;
        const float exposure = vw.cam.exposure;
        for (uint32_t g = blockIdx.x; g < groups; g += gridDim.x) {
            const uint32_t npx = min(64u, vw.npix - g * 64u);  // pixels of this group
            uint32_t pix = 0;
            float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (lane < npx) {
                const uint2 pxy = tile_pixel(g * 64u + lane, vw.width, vw.height);
                pix = pxy.y * vw.width + pxy.x;
                acc = vw.accum[pix];
            }
            const float4* src = rad + sg.path_offset + size_t(g) * 64u * sg.samples;  // the group's first path
            for (uint32_t c0 = 0; c0 < sg.samples; c0 += ACC_C) {
                const uint32_t cn = min(ACC_C, sg.samples - c0);
                for (uint32_t e = lane; e < npx * cn; e += 64u) {
                    const uint32_t p = e / cn, k = e - p * cn;
                    buf[p * (ACC_C + 1) + k] = src[size_t(p) * sg.samples + c0 + k];
                }
                __syncthreads();
                if (lane < npx) {
                    for (uint32_t k = 0; k < cn; k++) {
                        const float4 r4 = buf[lane * (ACC_C + 1) + k];
                        float rx, ry, rz;
                        if (signed_radiance) {
                            rx = r4.x * exposure, ry = r4.y * exposure, rz = r4.z * exposure;
                            rx = isfinite(rx) ? rx : 0.0f, ry = isfinite(ry) ? ry : 0.0f, rz = isfinite(rz) ? rz : 0.0f;
                        } else {
                            rx = correct_component(r4.x * exposure);
                            ry = correct_component(r4.y * exposure);
                            rz = correct_component(r4.z * exposure);
                        }
                        const uint32_t si = sg.first_sample + c0 + k;
                        if (si == 0) {
                            acc = make_float4(rx, ry, rz, 1.0f);
                        } else {
                            const float n = float(si + 1u);
                            acc.x = acc.x + (rx - 1.0f * acc.x) / n;
                            acc.y = acc.y + (ry - 1.0f * acc.y) / n;
                            acc.z = acc.z + (rz - 1.0f * acc.z) / n;
                        }
                    }
                }
                __syncthreads();
            }
            if (lane < npx)
                vw.accum[pix] = acc;
        }
    }
}

__global__ __launch_bounds__(TRACE_BLOCK) void k_primary_hits(DeviceScene s, CameraBlock cam, uint32_t width,
                                                              uint32_t height, uint32_t sample_index, pt_hit* hits,
                                                              Wavefront wf)
{
    __shared__ uint32_t lds[(LDS_STACK + COLD_ROWS) * TRACE_BLOCK];
    const uint32_t n = width * height;
    const uint32_t gtid = blockIdx.x * TRACE_BLOCK + threadIdx.x;
    Stack stk{&lds[threadIdx.x], wf.spill + gtid, wf.spill_threads, wf.spill_depth, 0, false};
    for (uint32_t p = gtid; p < n; p += gridDim.x * TRACE_BLOCK) {
        const uint32_t px = p % width, py = p / width;
        Sampler smp;
        smp.start(px, py, sample_index, 0);
        const float2 up = smp.s2();
        const float2 ul = smp.s2();
        float3 o, d;
        camera_ray(cam, px, py, width, height, up, ul, o, d);
        HitOut h;
        float trans;
        traverse<false>(s, o, d, cam.near_clip, cam.far_clip, stk, h, trans);
        hits[p] = h.inst == NONE ? pt_hit{0.0f, 0.0f, 0.0f, PT_NO_HIT, PT_NO_HIT} : pt_hit{h.t, h.u, h.v, h.inst, h.prim};
        if (stk.overflow)
            atomicOr(&wf.counters[CTR(9)], 1u);
    }
}

__global__ __launch_bounds__(TRACE_BLOCK) void k_trace_rays(DeviceScene s, const pt_ray* rays, uint32_t n, pt_hit* hits,
                                                            int any_hit, Wavefront wf)
{
    __shared__ uint32_t lds[(LDS_STACK + COLD_ROWS) * TRACE_BLOCK];
    const uint32_t gtid = blockIdx.x * TRACE_BLOCK + threadIdx.x;
    Stack stk{&lds[threadIdx.x], wf.spill + gtid, wf.spill_threads, wf.spill_depth, 0, false};
    for (uint32_t p = gtid; p < n; p += gridDim.x * TRACE_BLOCK) {
        const pt_ray r = rays[p];
        HitOut h;
        float trans;
        bool found;
        const float3 o = f3(r.origin[0], r.origin[1], r.origin[2]), d = f3(r.dir[0], r.dir[1], r.dir[2]);
        if (any_hit)
            found = traverse<true>(s, o, d, r.tmin, r.tmax, stk, h, trans);
        else
            found = traverse<false>(s, o, d, r.tmin, r.tmax, stk, h, trans);
        hits[p] = found ? pt_hit{h.t, h.u, h.v, h.inst, h.prim} : pt_hit{0.0f, 0.0f, 0.0f, PT_NO_HIT, PT_NO_HIT};
        if (stk.overflow)
            atomicOr(&wf.counters[CTR(9)], 1u);
    }
}

inline unsigned blocks_for(uint64_t n, unsigned block, unsigned cap)
{
    const uint64_t g = (n + block - 1) / block;
    return unsigned(std::max<uint64_t>(1, std::min<uint64_t>(g, cap)));
}

}  // namespace

hipError_t wavefront_alloc(Wavefront& wf, uint32_t capacity, uint32_t rad_out_paths, uint32_t spill_threads)
{
    wavefront_free(wf);
    wf.capacity = capacity;
    wf.rad_out_capacity = rad_out_paths;
    wf.spill_threads = spill_threads;
    const size_t c = std::max<uint32_t>(capacity, 1);
    hipError_t e = hipSuccess;
    auto A = [&](auto*& p, size_t bytes) {
        if (e == hipSuccess)
            e = hipMalloc((void**)&p, bytes);
    };
    A(wf.s_scene, c * 4);
    for (int k = 0; k < 2; k++) {
        A(wf.q_o[k], c * 16);
        A(wf.q_d[k], c * 16);
        A(wf.q_ps[k], c * 16);
        A(wf.q_thr[k], c * 16);
        A(wf.q_rad[k], c * 16);
        A(wf.q_scene[k], c * 4);
    }
    A(wf.hit, c * 16);
    A(wf.hit_prim, c * 4);
    A(wf.s_target, c * 4);
    A(wf.s_o, c * 16);
    A(wf.s_d, c * 16);
    A(wf.s_c, c * 16);
    A(wf.train_end, c * 8);
    A(wf.rad_out, size_t(std::max<uint32_t>(rad_out_paths, 1)) * 2 * 16);  // one half per job slot
    A(wf.counters, size_t(CTR(16)) * 4);
    A(wf.snapshot, 16 * 4);
    A(wf.totals, TOTALS * 8);
    A(wf.spill, size_t(spill_threads) * SPILL_STACK * 4);
    if (e == hipSuccess)
        e = hipHostMalloc((void**)&wf.host_counts, 16 * sizeof(uint32_t), hipHostMallocDefault);
    for (int k = 0; k < 2 && e == hipSuccess; k++)
        e = hipEventCreateWithFlags(&wf.count_ready[k], hipEventDisableTiming);
    if (e == hipSuccess)
        e = hipMemset(wf.counters, 0, size_t(CTR(16)) * 4);
    if (e == hipSuccess)
        e = hipMemset(wf.totals, 0, TOTALS * 8);
    return e;
}

void wavefront_free(Wavefront& wf)
{
    void* ptrs[] = {wf.q_o[0],      wf.q_o[1],      wf.q_d[0],      wf.q_d[1],      wf.q_ps[0],
                    wf.q_ps[1],     wf.q_thr[0],    wf.q_thr[1],    wf.q_rad[0],    wf.q_rad[1],
                    wf.q_scene[0],  wf.q_scene[1],  wf.hit,         wf.hit_prim,    wf.s_target,
                    wf.s_o,         wf.s_d,         wf.s_c,         wf.s_scene,     wf.train_end,
                    wf.rad_out,     wf.counters,    wf.totals,      wf.spill,       wf.snapshot,    wf.sort_keys[0],
                    wf.sort_keys[1], wf.sort_vals[0], wf.sort_vals[1], wf.sort_temp};
    for (void* p : ptrs)
        if (p)
            (void)hipFree(p);
    if (wf.host_counts)
        (void)hipHostFree(wf.host_counts);
    for (int k = 0; k < 2; k++)
        if (wf.count_ready[k])
            (void)hipEventDestroy(wf.count_ready[k]);
    wf = Wavefront{};
}

namespace {

typedef void (*ExtendFn)(DeviceScene, const DeviceScene*, Wavefront, int);
typedef void (*ShadowFn)(DeviceScene, const DeviceScene*, Wavefront, int);

// Traversal kernel variants by fetch counting, scene table, LDS stack depth (8, 16 or 32
// entries per lane), alpha-weighted shadow rays and single-triangle leaves (the counting
// kernels keep the general leaf loop).
template <uint32_t D>
ExtendFn pick_extend_d(bool count, bool multi, bool leaf1)
{
    if (count)
        return multi ? k_extend<true, true, D, false> : k_extend<true, false, D, false>;
    return multi ? (leaf1 ? k_extend<false, true, D, true> : k_extend<false, true, D, false>)
                 : (leaf1 ? k_extend<false, false, D, true> : k_extend<false, false, D, false>);
}
template <uint32_t D, bool ALPHA>
ShadowFn pick_shadow_da(bool count, bool multi, bool leaf1)
{
    if (count)
        return multi ? k_shadow<true, true, D, ALPHA, false> : k_shadow<true, false, D, ALPHA, false>;
    return multi ? (leaf1 ? k_shadow<false, true, D, ALPHA, true> : k_shadow<false, true, D, ALPHA, false>)
                 : (leaf1 ? k_shadow<false, false, D, ALPHA, true> : k_shadow<false, false, D, ALPHA, false>);
}
template <uint32_t D>
ShadowFn pick_shadow_d(bool count, bool multi, bool alpha, bool leaf1)
{
    return alpha ? pick_shadow_da<D, true>(count, multi, leaf1) : pick_shadow_da<D, false>(count, multi, leaf1);
}
ExtendFn pick_extend(bool count, bool multi, uint32_t depth, bool leaf1)
{
    return depth == 8    ? pick_extend_d<8>(count, multi, leaf1)
           : depth == 32 ? pick_extend_d<32>(count, multi, leaf1)
                         : pick_extend_d<16>(count, multi, leaf1);
}
ShadowFn pick_shadow(bool count, bool multi, uint32_t depth, bool alpha, bool leaf1)
{
    return depth == 8    ? pick_shadow_d<8>(count, multi, alpha, leaf1)
           : depth == 32 ? pick_shadow_d<32>(count, multi, alpha, leaf1)
                         : pick_shadow_d<16>(count, multi, alpha, leaf1);
}
typedef void (*BothFn)(DeviceScene, const DeviceScene*, Wavefront, int, const ViewDev*, const JobDev*, const JobDev*,
                       int);
template <uint32_t D, bool ALPHA>
BothFn pick_both_da(bool multi, bool leaf1)
{
    return multi ? (leaf1 ? k_trace_both<true, D, ALPHA, true> : k_trace_both<true, D, ALPHA, false>)
                 : (leaf1 ? k_trace_both<false, D, ALPHA, true> : k_trace_both<false, D, ALPHA, false>);
}
template <uint32_t D>
BothFn pick_both_d(bool multi, bool alpha, bool leaf1)
{
    return alpha ? pick_both_da<D, true>(multi, leaf1) : pick_both_da<D, false>(multi, leaf1);
}
BothFn pick_both(bool multi, uint32_t depth, bool alpha, bool leaf1)
{
    return depth == 8    ? pick_both_d<8>(multi, alpha, leaf1)
           : depth == 32 ? pick_both_d<32>(multi, alpha, leaf1)
                         : pick_both_d<16>(multi, alpha, leaf1);
}

using ShadeFn = void (*)(DeviceScene, const DeviceScene*, const ViewDev*, const JobDev*, Wavefront, int, CacheDev,
                        const uint32_t*);

// Register budget of the shade stage: PT_SHADE_WAVES (2..6 waves per SIMD; tuning) or 2, which
// lets the compiler use what the code needs without spilling: the plain tracer's variant takes
// 162 registers (3 waves per SIMD), the radiance cache's 185 (2 waves). Forcing 4 waves spills
// the plain variant and ran its stage 8 % slower.
int shade_waves()
{
    static const int w = [] {
        const char* env = getenv("PT_SHADE_WAVES");
        return env ? std::max(2, std::min(6, atoi(env))) : 2;
    }();
    return w;
}

template <bool MULTI, bool CACHE>
ShadeFn pick_shade_m()
{
    switch (shade_waves()) {
    case 2:
        return k_shade<MULTI, 2, CACHE>;
    case 3:
        return k_shade<MULTI, 3, CACHE>;
    case 5:
        return k_shade<MULTI, 5, CACHE>;
    case 6:
        return k_shade<MULTI, 6, CACHE>;
    default:
        return k_shade<MULTI, 4, CACHE>;
    }
}

ShadeFn pick_shade(bool multi, bool cache)
{
    return multi ? (cache ? pick_shade_m<true, true>() : pick_shade_m<true, false>())
                 : (cache ? pick_shade_m<false, true>() : pick_shade_m<false, false>());
}

// Blocks of 256 threads of the shade stage resident on the whole chip at once: a
// grid-stride launch of exactly that many finishes in one round, with no partial round.
unsigned shade_grid(bool multi, bool cache)
{
    static unsigned g[4] = {0, 0, 0, 0};
    unsigned& v = g[(multi ? 1 : 0) + (cache ? 2 : 0)];
    if (v == 0) {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
            cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(pick_shade(multi, cache)), 256,
                                                         0) != hipSuccess ||
            per_cu < 1)
            per_cu = 2;
        v = unsigned(cus) * unsigned(per_cu);
    }
    return v;
}

// One wavefront iteration with queue parity q: clear the counters, extend, shade, shadow,
// cache queries, refill, and copy the counts the host steers by to pinned memory.
// Pipelined (mode.pipelined): the shade stage's shadow rays are traced in the next iteration,
// in one launch with its extension rays, and the new paths the refill plans are started by that
// launch too (k_trace_both; mode.consume when the previous iteration left both). mode.flush: a
// non-pipelined iteration first traces the shadow rays and writes the new paths a pipelined one
// left.
hipError_t launch_iteration(const RenderIO& io, const CacheDev& cd, int q, unsigned shade_blocks, IterMode mode)
{
    Wavefront& wf = *io.wf;
    hipStream_t st = io.st;
    KernelTimer& timer = *io.timer;
    const bool multi = io.d_scenes != nullptr;
    const DeviceScene* tab = io.d_scenes;
    hipError_t e;
    const dim3 tg(io.grid_blocks), tb(TRACE_BLOCK);
    const bool keep_shadow = mode.pipelined || mode.flush;
    hipLaunchKernelGGL(k_prepare, dim3(1), dim3(64), 0, st, wf, q, cd, keep_shadow ? 0 : 1);
    if (mode.flush) {  // what the pipelined iteration before left: its shadow rays, its new paths
        hipLaunchKernelGGL(pick_shadow(false, multi, io.lds_stack, io.alpha_shadows, io.leaf1), tg, tb, 0, st, io.scene, tab, wf,
                           1 - q);
        hipLaunchKernelGGL(k_clear_shadow, dim3(1), dim3(64), 0, st, wf);
        CacheDev none = cd;
        none.collect = 0;
        hipLaunchKernelGGL(k_refill, dim3(shade_blocks), dim3(256), 0, st, io.d_views, io.d_job, io.d_job + 1, wf,
                           1 - q, none);
    }
    if (mode.pipelined) {
        hipLaunchKernelGGL(pick_both(multi, io.lds_stack, io.alpha_shadows, io.leaf1), tg, tb, 0, st, io.scene, tab, wf, q,
                           io.d_views, io.d_job, io.d_job + 1, mode.consume ? 1 : 0);
        hipLaunchKernelGGL(k_clear_shadow, dim3(1), dim3(64), 0, st, wf);
    } else {
        if ((e = timer.mark(st, KERNEL_EXTEND, true)) != hipSuccess)
            return e;
        hipLaunchKernelGGL(pick_extend(io.count_fetches, multi, io.lds_stack, io.leaf1), tg, tb, 0, st, io.scene, tab, wf, q);
        if ((e = timer.mark(st, KERNEL_EXTEND, false)) != hipSuccess)
            return e;
    }
    const uint32_t* order = nullptr;
    if (io.sort_hits) {
        // shade in (scene, material) order: radix sort of the hit keys over the whole pool
        if (multi)
            hipLaunchKernelGGL(k_hit_keys<true>, dim3(shade_blocks), dim3(256), 0, st, io.scene, tab, wf, q, wf.capacity);
        else
            hipLaunchKernelGGL(k_hit_keys<false>, dim3(shade_blocks), dim3(256), 0, st, io.scene, tab, wf, q,
                               wf.capacity);
        hipcub::DoubleBuffer<uint32_t> dk(wf.sort_keys[0], wf.sort_keys[1]);
        hipcub::DoubleBuffer<uint32_t> dv(wf.sort_vals[0], wf.sort_vals[1]);
        size_t bytes = wf.sort_temp_bytes;
        if ((e = hipcub::DeviceRadixSort::SortPairs(wf.sort_temp, bytes, dk, dv, int(wf.capacity), 0, 32, st)) !=
            hipSuccess)
            return e;
        order = dv.Current();
    }
    if ((e = timer.mark(st, KERNEL_SHADE, true)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(pick_shade(multi, cd.collect != 0), dim3(shade_blocks), dim3(256), 0, st, io.scene, tab,
                       io.d_views, io.d_job,
                       wf, q, cd, order);
    if ((e = timer.mark(st, KERNEL_SHADE, false)) != hipSuccess)
        return e;
    if (!mode.pipelined) {
        if ((e = timer.mark(st, KERNEL_SHADOW, true)) != hipSuccess)
            return e;
        hipLaunchKernelGGL(pick_shadow(io.count_fetches, multi, io.lds_stack, io.alpha_shadows,
                                       io.leaf1), tg, tb, 0, st,
                           io.scene, tab, wf, q);
        if ((e = timer.mark(st, KERNEL_SHADOW, false)) != hipSuccess)
            return e;
    }
    if (cd.query) {
        if ((e = timer.mark(st, KERNEL_CACHE, true)) != hipSuccess)
            return e;
        if ((e = cache_query(*io.cache, &io.cache->counters[1], wf.q_rad[1 - q], wf.rad_out, io.grid_blocks, st)) !=
            hipSuccess)
            return e;
        if ((e = timer.mark(st, KERNEL_CACHE, false)) != hipSuccess)
            return e;
    }
    if ((e = timer.mark(st, KERNEL_REFILL, true)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_refill_commit, dim3(1), dim3(64), 0, st, io.d_job, io.d_job + 1, wf, q);
    if (!mode.pipelined)
        hipLaunchKernelGGL(k_refill, dim3(shade_blocks), dim3(256), 0, st, io.d_views, io.d_job, io.d_job + 1, wf, q,
                           cd);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    if ((e = timer.mark(st, KERNEL_REFILL, false)) != hipSuccess)
        return e;
    return hipMemcpyAsync(wf.host_counts + 8 * q, wf.snapshot + 8 * q, 8 * 4, hipMemcpyDeviceToHost, st);
}

// Launch configuration an iteration graph depends on.
uint64_t graph_signature(const RenderIO& io, const CacheDev& cd, IterMode mode)
{
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](uint64_t v) { h = (h ^ v) * 1099511628211ull; };
    mix((mode.pipelined ? 1u : 0u) | (mode.consume ? 2u : 0u) | (mode.flush ? 4u : 0u));
    mix(io.count_fetches);
    mix(io.lds_stack);
    mix(io.alpha_shadows);
    mix(io.leaf1);
    mix(io.sort_hits);
    mix(reinterpret_cast<uintptr_t>(io.scene.nodes));
    mix(reinterpret_cast<uintptr_t>(io.d_scenes));
    mix(reinterpret_cast<uintptr_t>(io.d_views));
    mix(reinterpret_cast<uintptr_t>(io.d_job));
    mix(reinterpret_cast<uintptr_t>(io.wf->q_ps[0]));
    mix(io.wf->capacity);
    mix(io.grid_blocks);
    mix(io.cache ? reinterpret_cast<uintptr_t>(io.cache->params) : 0);
    mix(cd.collect);
    mix(cd.query);
    uint32_t cp;
    memcpy(&cp, &cd.cont_prob, 4);
    mix(cp);
    mix(cd.train_stride);
    mix(cd.segment);
    mix(float_as_u32(cd.max_pdf));
    mix(reinterpret_cast<uintptr_t>(cd.records));
    mix(reinterpret_cast<uintptr_t>(cd.queries));
    return h | 1ull;
}

// Runs one iteration through its captured graph (captured on first use).
hipError_t graph_iteration(const RenderIO& io, const CacheDev& cd, int q, unsigned shade_blocks, IterMode mode)
{
    IterGraphs& g = *io.graphs;
    const uint64_t sig = graph_signature(io, cd, mode);
    hipError_t e;
    if (!g.exec[q] || g.sig[q] != sig) {
        if (g.exec[q]) {
            (void)hipGraphExecDestroy(g.exec[q]);
            g.exec[q] = nullptr;
        }
        if ((e = hipStreamBeginCapture(io.st, hipStreamCaptureModeThreadLocal)) != hipSuccess)
            return e;
        hipError_t le = launch_iteration(io, cd, q, shade_blocks, mode);
        hipGraph_t graph = nullptr;
        e = hipStreamEndCapture(io.st, &graph);
        if (le != hipSuccess || e != hipSuccess) {
            if (graph)
                (void)hipGraphDestroy(graph);
            return le != hipSuccess ? le : e;
        }
        e = hipGraphInstantiate(&g.exec[q], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            g.exec[q] = nullptr;
            return e;
        }
        g.sig[q] = sig;
    }
    return hipGraphLaunch(g.exec[q], io.st);
}

// Cuts the work into jobs of at most wf.rad_out_capacity paths (and, when the cache trains
// between frames, at most max_samples samples of one view per job).
struct JobPlan {
    std::vector<SegDev> segs;
    uint64_t paths = 0, pix = 0;
};

std::vector<JobPlan> plan_jobs(const std::vector<ViewWork>& work, uint32_t cap, uint32_t max_samples)
{
    std::vector<JobPlan> jobs;
    JobPlan cur;
    auto close = [&]() {
        if (!cur.segs.empty())
            jobs.push_back(cur);
        cur = JobPlan{};
    };
    for (const ViewWork& w : work) {
        uint32_t left = w.spp, sample = w.sample_index;
        while (left > 0) {
            uint64_t fit = (cap - cur.paths) / w.npix;
            fit = std::min<uint64_t>(fit, left);
            if (max_samples)
                fit = std::min<uint64_t>(fit, max_samples);
            if (fit == 0 || cur.segs.size() == MAX_SEGS) {
                close();
                continue;
            }
            SegDev sg{};
            sg.view = w.view;
            sg.first_sample = sample;
            sg.samples = uint32_t(fit);
            sg.path_offset = uint32_t(cur.paths);
            sg.pix_offset = uint32_t(cur.pix);
            cur.segs.push_back(sg);
            cur.paths += fit * w.npix;
            cur.pix += w.npix;
            left -= uint32_t(fit);
            sample += uint32_t(fit);
            if (max_samples)
                close();  // one frame per job: the cache trains in between
        }
    }
    close();
    return jobs;
}

}  // namespace

hipError_t wavefront_sort_buffers(Wavefront& wf)
{
    if (wf.sort_temp)
        return hipSuccess;
    const size_t c = std::max<uint32_t>(wf.capacity, 1);
    hipError_t e = hipSuccess;
    for (int k = 0; k < 2 && e == hipSuccess; k++) {
        e = hipMalloc(&wf.sort_keys[k], c * 4);
        if (e == hipSuccess)
            e = hipMalloc(&wf.sort_vals[k], c * 4);
    }
    if (e != hipSuccess)
        return e;
    hipcub::DoubleBuffer<uint32_t> dk(wf.sort_keys[0], wf.sort_keys[1]);
    hipcub::DoubleBuffer<uint32_t> dv(wf.sort_vals[0], wf.sort_vals[1]);
    size_t bytes = 0;
    if ((e = hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, dk, dv, int(wf.capacity), 0, 32)) != hipSuccess)
        return e;
    wf.sort_temp_bytes = bytes;
    return hipMalloc(&wf.sort_temp, std::max<size_t>(bytes, 16));
}

hipError_t render_views(const RenderIO& io, const std::vector<ViewWork>& work, bool drain)
{
    Wavefront& wf = *io.wf;
    StreamState& S = *io.stream;
    hipStream_t st = io.st;
    hipError_t e;
    if (io.sort_hits && (e = wavefront_sort_buffers(wf)) != hipSuccess)
        return e;
    CacheState* cs = io.cache && io.cache->configured ? io.cache : nullptr;
    const bool train_between = cs && cs->train_steps > 0 && cs->frame_spp > 0;
    const bool train_inflight = cs && cs->train_steps > 0 && cs->frame_spp == 0;
    for (const ViewWork& w : work)
        if (w.npix == 0 || w.npix > wf.rad_out_capacity)
            return hipErrorInvalidValue;
    auto jobs = plan_jobs(work, wf.rad_out_capacity, train_between ? cs->frame_spp : 0);
    // The radiance cache records, trains and counts per job: its jobs run one at a time.
    const int slots = cs ? 1 : 2;
    if (cs)
        drain = true;
    // grid-stride stages: the blocks that fill the chip in one round, not one per pool slot
    const unsigned shade_blocks = blocks_for(wf.capacity, 256, shade_grid(io.d_scenes != nullptr, cs != nullptr));
    CacheDev cd{};
    if (cs) {
        cd.records = cs->records;
        cd.counters = cs->counters;
        cd.record_capacity = cs->record_capacity;
        cd.queries = cs->queries;
        cd.query_slot = cs->query_slot;
        cd.train_stride = std::max<uint32_t>(1, cs->train_stride);
        cd.collect = 1;
        cd.segment = cs->query_vertex - 1;  // the path segment that arrives at the cache's vertex
        cd.query = (cs->query && cs->steps_done > 0) ? 1u : 0u;  // an untrained cache is not queried
        cd.cont_prob = cs->cont_prob;
        cd.max_pdf = cs->query_max_pdf;
    }
    // (the library radix sort is launched directly: it may record memsets)
    const bool graphs = io.graphs && !io.timer->enabled && !train_inflight && !io.sort_hits;
    // Shadow rays traced with the next iteration's extension rays: in the production path (the
    // profiled pass times every stage on its own; fetch counting and the cache, whose training
    // targets need every path's radiance complete when it ends, trace them in their iteration)
    const bool pipelined = graphs && !io.count_fetches && !cs;
    size_t next_job = 0;
    auto load = [&](int s) -> hipError_t {
        const JobPlan& plan = jobs[next_job++];
        JobDev& jh = io.h_job[s];
        jh.nseg = uint32_t(plan.segs.size());
        for (size_t k = 0; k < plan.segs.size(); k++)
            jh.seg[k] = plan.segs[k];
        jh.paths = uint32_t(plan.paths);
        jh.pix_total = uint32_t(plan.pix);
        jh.seq = ++S.seq;
        jh.signed_radiance = (cd.query && cd.cont_prob > 0.0f) ? 1u : 0u;
        const size_t bytes = offsetof(JobDev, seg) + plan.segs.size() * sizeof(SegDev);
        hipError_t le = hipMemcpyAsync(io.d_job + s, &jh, bytes, hipMemcpyHostToDevice, st);
        if (le != hipSuccess)
            return le;
        hipLaunchKernelGGL(k_slot_reset, dim3(1), dim3(64), 0, st, wf, uint32_t(s));
        if (cs && cd.collect) {  // a job's training records start afresh
            const uint32_t zero[2] = {0, 0};
            if ((le = hipMemcpyAsync(cd.counters, zero, 8, hipMemcpyHostToDevice, st)) != hipSuccess)
                return le;
        }
        S.slot[s] = StreamState::Slot{true, jh.paths, jh.seq, S.launched, 0, 0};
        return hipGetLastError();
    };
    // The job in slot s has ended: accumulate it, free the slot (its paths are read before a
    // later job's paths can be written there: stream order), train between frames.
    auto finish = [&](int s) -> hipError_t {
        const uint32_t pix_total = io.h_job[s].pix_total;
        hipLaunchKernelGGL(k_accumulate, dim3(blocks_for(pix_total, 64, 65535u * 4)), dim3(64), 0, st, io.d_views,
                           io.d_job + s, wf.rad_out + size_t(s) * wf.rad_out_capacity);
        hipError_t le = hipGetLastError();
        if (le != hipSuccess)
            return le;
        hipLaunchKernelGGL(k_slot_free, dim3(1), dim3(64), 0, st, io.d_job + s, wf, uint32_t(s));
        if ((le = hipGetLastError()) != hipSuccess)
            return le;
        S.slot[s].busy = false;
        if (train_between) {
            // online training on the path suffixes this job recorded
            uint32_t n_rec = 0;
            if ((le = hipMemcpyAsync(&n_rec, &cs->counters[0], 4, hipMemcpyDeviceToHost, st)) != hipSuccess)
                return le;
            if ((le = hipStreamSynchronize(st)) != hipSuccess)
                return le;
            n_rec = std::min(n_rec, cs->record_capacity);
            cs->records_total += n_rec;
            cs->last_records = n_rec;
            if ((le = cache_train_records(*cs, n_rec, st)) != hipSuccess)
                return le;
        }
        return hipSuccess;
    };
    while (true) {
        // jobs into free slots (one slot with the cache)
        for (int s = 0; s < slots && next_job < jobs.size(); s++)
            if (!S.slot[s].busy)
                if ((e = load(s)) != hipSuccess)
                    return e;
        if (S.idle())
            break;  // nothing in flight and nothing left to do
        if (!drain && next_job == jobs.size()) {
            // every job of this call is loaded; return once all its paths have started
            bool started = true;
            for (int s = 0; s < 2; s++)
                started &= !S.slot[s].busy || S.slot[s].issued == S.slot[s].paths;
            if (started)
                break;
        }
        // One iteration. With an empty pool it only refills, so a stream starts by itself.
        IterMode mode;
        mode.pipelined = pipelined;
        mode.consume = pipelined && S.shadow_pending;
        mode.flush = !pipelined && S.shadow_pending;
        e = graphs ? graph_iteration(io, cd, S.q, shade_blocks, mode)
                   : launch_iteration(io, cd, S.q, shade_blocks, mode);
        if (e != hipSuccess)
            return e;
        S.shadow_pending = pipelined;
        if (train_inflight) {
            if ((e = io.timer->mark(st, KERNEL_CACHE, true)) != hipSuccess)
                return e;
            if ((e = cache_train_inflight(*cs, st)) != hipSuccess)
                return e;
            if ((e = io.timer->mark(st, KERNEL_CACHE, false)) != hipSuccess)
                return e;
        }
        if ((e = hipEventRecord(wf.count_ready[S.q], st)) != hipSuccess)
            return e;
        S.q = 1 - S.q;
        if (++S.launched < 2)
            continue;
        // the counts of the iteration before the one just queued (q now names its parity)
        if ((e = hipEventSynchronize(wf.count_ready[S.q])) != hipSuccess)
            return e;
        const uint64_t counted = S.launched - 2;
        const uint32_t* snap = wf.host_counts + 8 * S.q;
        for (int s = 0; s < 2; s++) {
            // a snapshot taken before a job was loaded still holds its slot's previous job
            if (!S.slot[s].busy || counted < S.slot[s].first)
                continue;
            S.slot[s].ended = snap[1 + s];
            S.slot[s].issued = snap[3 + s];
        }
        if (snap[5] && counted >= S.flags_from)  // earlier snapshots predate the last clear
            S.overflow = true;
        // finished jobs in the order they were loaded
        for (int round = 0; round < 2; round++) {
            int oldest = -1;
            for (int s = 0; s < 2; s++)
                if (S.slot[s].busy && (oldest < 0 || int(S.slot[s].seq - S.slot[oldest].seq) < 0))
                    oldest = s;
            if (oldest < 0 || S.slot[oldest].ended != S.slot[oldest].paths)
                break;
            if ((e = finish(oldest)) != hipSuccess)
                return e;
        }
    }
    return hipSuccess;
}

hipError_t render_samples(Ctx& ctx, uint32_t spp, bool drain)
{
    hipError_t e;
    ViewDev& v = *ctx.h_view;
    v.cam = ctx.cam;
    v.accum = ctx.accum;
    v.width = ctx.width;
    v.height = ctx.height;
    v.npix = ctx.width * ctx.height;
    v.scene = 0;
    v.max_bounces = ctx.max_bounces;
    if ((e = hipMemcpyAsync(ctx.d_views, ctx.h_view, sizeof(v), hipMemcpyHostToDevice, ctx.stream)) != hipSuccess)
        return e;
    RenderIO io{ctx.stream,       &ctx.wf,     ctx.grid_blocks, ctx.count_fetches,
                &ctx.timer,       ctx.scene,   nullptr,         ctx.d_views,
                ctx.d_job,        ctx.h_job,   &ctx.stream_state, ctx.use_graphs ? &ctx.graphs : nullptr,
                &ctx.cache,       ctx.lds_stack,   ctx.sort_hits};
    io.alpha_shadows = ctx.transparent;
    io.leaf1 = ctx.single_tri_leaves;
    return render_views(io, spp ? std::vector<ViewWork>{ViewWork{0, v.npix, ctx.sample_index, spp}}
                                : std::vector<ViewWork>{},
                        drain);
}

hipError_t drain_stream(Ctx& ctx)
{
    if (ctx.stream_state.idle())
        return hipSuccess;
    RenderIO io{ctx.stream,       &ctx.wf,     ctx.grid_blocks, ctx.count_fetches,
                &ctx.timer,       ctx.scene,   nullptr,         ctx.d_views,
                ctx.d_job,        ctx.h_job,   &ctx.stream_state, ctx.use_graphs ? &ctx.graphs : nullptr,
                &ctx.cache,       ctx.lds_stack,   ctx.sort_hits};
    io.alpha_shadows = ctx.transparent;
    io.leaf1 = ctx.single_tri_leaves;
    return render_views(io, {}, true);
}

hipError_t primary_hits(Ctx& ctx, uint32_t sample_index, pt_hit* d_hits)
{
    hipLaunchKernelGGL(k_primary_hits, dim3(ctx.grid_blocks), dim3(TRACE_BLOCK), 0, ctx.stream, ctx.scene, ctx.cam,
                       ctx.width, ctx.height, sample_index, d_hits, ctx.wf);
    return hipGetLastError();
}

hipError_t trace_rays(Ctx& ctx, const pt_ray* d_rays, uint32_t count, pt_hit* d_hits, bool any_hit)
{
    if (count == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_trace_rays, dim3(ctx.grid_blocks), dim3(TRACE_BLOCK), 0, ctx.stream, ctx.scene, d_rays, count,
                       d_hits, any_hit ? 1 : 0, ctx.wf);
    return hipGetLastError();
}

}  // namespace ptk
