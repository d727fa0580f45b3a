// CPU oracle of the path-tracing hot path. TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Each function cites the reference shader lines it restates. Shading arithmetic is fp32,
// as in the shaders; intersection is the watertight ray/triangle test in fp64.
#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <vector>

namespace {

// ------------------------------------------------------------------------------------------
// small vector types (fp32)

struct f2 {
    float x, y;
};
struct f3 {
    float x, y, z;
};
struct f4 {
    float x, y, z, w;
};
inline f3 mk(float x, float y, float z) { return {x, y, z}; }
inline f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
inline f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline f3 operator*(float s, f3 a) { return {a.x * s, a.y * s, a.z * s}; }
inline f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline f3 cross(f3 a, f3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline f3 normalize(f3 v) { return v * (1.0f / std::sqrt(dot(v, v))); }
inline float maxc(f3 v) { return std::max(std::max(v.x, v.y), v.z); }
inline float lerpf(float a, float b, float t) { return a + t * (b - a); }
inline f3 lerp3(f3 a, f3 b, f3 t) { return {lerpf(a.x, b.x, t.x), lerpf(a.y, b.y, t.y), lerpf(a.z, b.z, t.z)}; }
inline float sq(float x) { return x * x; }
inline float as_float(uint32_t u)
{
    float f;
    memcpy(&f, &u, 4);
    return f;
}

const float PI_F = 3.141592654f;  // util.hlsli:4
const float JUST_BELOW_ONE = as_float(0x3F7FFFFFu);
const float ORIGIN_OFFSET = 1e-5f;  // pathtrace.hlsli:6

// ------------------------------------------------------------------------------------------
// hashing and the Sobol/Owen sampler (util.hlsli:8-27, sample.hlsli:31-166)

inline uint32_t hash_combine(uint32_t seed, uint32_t x) { return seed ^ (x + (seed << 6) + (seed >> 2)); }

inline uint32_t hash_u32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x85ebca6bu;
    x ^= x >> 13;
    x *= 0xc2b2ae35u;
    x ^= x >> 16;
    return x;
}

inline uint32_t reverse_bits(uint32_t x)
{
    uint32_t r = 0;
    for (int i = 0; i < 32; i++)
        r |= ((x >> i) & 1u) << (31 - i);
    return r;
}

inline uint32_t lk_permutation(uint32_t x, uint32_t seed)
{
    x += seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return x;
}

inline uint32_t owen_scramble(uint32_t x, uint32_t seed) { return reverse_bits(lk_permutation(reverse_bits(x), seed)); }

// Sobol direction numbers of dimensions 0-3 (sample.hlsli:31-67): generated from the
// primitive polynomials x+1, x^2+x+1, x^3+x+1 with initial m = 1 | 1,3 | 1,3,1
// (Joe & Kuo); dimension 0 is the van der Corput sequence.
struct SobolTable {
    uint32_t v[4][32];
    SobolTable()
    {
        for (int b = 0; b < 32; b++)
            v[0][b] = 1u << (31 - b);
        const int deg[3] = {1, 2, 3};
        const uint32_t poly[3] = {0, 1, 1};  // interior coefficients a
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
    }
};
const SobolTable kSobol;

inline uint32_t sobol(uint32_t index, uint32_t dim)
{
    uint32_t x = 0;
    for (int b = 0; b < 32; b++)
        if ((index >> b) & 1u)
            x ^= kSobol.v[dim][b];
    return x;
}

struct Sampler {
    uint32_t index, seed, dimension, shuffled;

    void start(uint32_t px, uint32_t py, uint32_t sample_index, uint32_t s)
    {
        seed = hash_combine(s, hash_combine(px, py));
        index = sample_index;
        dimension = 4;
        shuffled = 0;
    }
    void pad()
    {
        dimension = 0;
        seed = hash_u32(seed);
        shuffled = owen_scramble(index, seed);
    }
    float value(uint32_t d) const
    {
        const uint32_t x = owen_scramble(sobol(shuffled, d), hash_combine(seed, d));
        return float(x) * 0x1p-32f;
    }
    float s1()
    {
        if (dimension >= 4)
            pad();
        float r = value(dimension);
        dimension += 1;
        return std::min(r, JUST_BELOW_ONE);
    }
    f2 s2()
    {
        if (dimension + 1 >= 4)
            pad();
        f2 r{value(dimension), value(dimension + 1)};
        dimension += 2;
        return {std::min(r.x, JUST_BELOW_ONE), std::min(r.y, JUST_BELOW_ONE)};
    }
};

// ------------------------------------------------------------------------------------------
// sampling helpers (sample.hlsli:193-246)

f3 cosine_weighted_rand_dir(f2 u)
{
    const float r = std::sqrt(u.x);
    return {r * std::cos(2.0f * PI_F * u.y), r * std::sin(2.0f * PI_F * u.y), std::sqrt(1.0f - u.x)};
}

f3 random_barycentric(f2 u)
{
    const float beta = 1.0f - std::sqrt(u.x);
    const float gamma = (1.0f - beta) * u.y;
    return {1.0f - beta - gamma, beta, gamma};
}

f2 sample_concentric_disk(f2 u)
{
    const float a = 2.0f * u.x - 1.0f;
    const float b = 2.0f * u.y - 1.0f;
    float r, phi;
    if (a * a > b * b) {
        r = a;
        phi = (PI_F / 4.0f) * (b / a);
    } else {
        r = b;
        phi = (PI_F / 2.0f) - (PI_F / 4.0f) * (a / b);
    }
    return {r * std::cos(phi), r * std::sin(phi)};
}

f3 sample_visible_micronormal(f3 v, f2 alpha, f2 u)
{
    v = normalize(mk(alpha.x * v.x, alpha.y * v.y, v.z));
    const f3 x = (v.z < 0.99999f) ? normalize(cross(mk(0, 0, 1), v)) : mk(1, 0, 0);
    const f3 y = cross(v, x);
    const float a = 1.0f / (1.0f + v.z);
    const float r = std::sqrt(u.x);
    const float phi = (u.y < a) ? u.y / a * PI_F : PI_F + (u.y - a) / (1.0f - a) * PI_F;
    const float p1 = r * std::cos(phi);
    const float p2 = r * std::sin(phi) * ((u.y < a) ? 1.0f : v.z);
    const f3 m = p1 * x + p2 * y + std::sqrt(std::max(0.0f, 1.0f - p1 * p1 - p2 * p2)) * v;
    return normalize(mk(alpha.x * m.x, alpha.y * m.y, std::max(1e-6f, m.z)));
}

// ------------------------------------------------------------------------------------------
// BSDF (pathtrace.ch.hlsl:22-157, util.hlsli:29-64)

struct BrdfInput {
    f3 base_colour;
    f2 alpha;
    float metalness;
    f3 specular;
    f3 sheen;
    float clearcoat;
    float clearcoat_alpha;
};

inline float cos_theta(f3 v) { return v.z; }
inline float cos2_theta(f3 v) { return sq(v.z); }
inline float sin2_theta(f3 v) { return std::max(0.0f, 1.0f - cos2_theta(v)); }
inline float sin_theta(f3 v) { return std::sqrt(sin2_theta(v)); }
inline float tan2_theta(f3 v) { return sin2_theta(v) / cos2_theta(v); }

float lambda_fn(f3 m, f2 alpha)
{
    const float tan2 = tan2_theta(m);
    if (std::isinf(tan2))
        return 0.0f;
    const float s = sin_theta(m);
    const float cp = s == 0.0f ? 1.0f : std::clamp(m.x / s, -1.0f, 1.0f);
    const float sp = s == 0.0f ? 0.0f : std::clamp(m.y / s, -1.0f, 1.0f);
    const float a2 = sq(cp * alpha.x) + sq(sp * alpha.y);
    return (std::sqrt(1.0f + a2 * tan2) - 1.0f) / 2.0f;
}

float g1_fn(f3 v, f2 alpha) { return 1.0f / (1.0f + lambda_fn(v, alpha)); }
float masking(f3 o, f3 i, f2 alpha) { return 1.0f / (1.0f + lambda_fn(o, alpha) + lambda_fn(i, alpha)); }

float distribution(f3 m, f2 alpha)
{
    const float tan2 = tan2_theta(m);
    if (std::isinf(tan2))
        return 0.0f;
    const float cos4 = sq(cos2_theta(m));
    const float s = sin_theta(m);
    const float cp = s == 0.0f ? 1.0f : std::clamp(m.x / s, -1.0f, 1.0f);
    const float sp = s == 0.0f ? 0.0f : std::clamp(m.y / s, -1.0f, 1.0f);
    const float e = tan2 * (sq(cp / alpha.x) + sq(sp / alpha.y));
    return 1.0f / (PI_F * alpha.x * alpha.y * cos4 * sq(1.0f + e));
}

inline float diffuse_sample_probability(float metalness) { return lerpf(0.5f, 0.0f, metalness); }

f3 evaluate_brdf(const BrdfInput& in, f3 i, f3 o, f3 m, float& pdf)
{
    const f3 f0 = lerp3(in.specular * 0.08f, in.base_colour, mk(in.metalness, in.metalness, in.metalness));
    const float cos_d = dot(m, o);
    const float base = 1.0f - cos_d;
    const float f = base * base * base * base * base;
    const float g = masking(o, i, in.alpha);
    const float d = distribution(m, in.alpha);
    const f3 fr = lerp3(f0, mk(1, 1, 1), mk(f, f, f));
    // d * lerp(f0, 1, f) * g / (4 cos_i cos_o), evaluated left to right as the shader does
    const f3 spec_exact = mk(d * fr.x * g, d * fr.y * g, d * fr.z * g) / (4.0f * cos_theta(i) * cos_theta(o));
    const f3 sheen = in.sheen * f;
    const f3 diffuse = (in.base_colour + sheen) * (1.0f - in.metalness) / PI_F * (1.0f - f);
    // Clearcoat is computed by the shader but never added to the result (ch.hlsl:93-114).
    const float diffuse_prob = diffuse_sample_probability(in.metalness);
    const float pdf_diffuse = cos_theta(i) / PI_F;
    const float pdf_m = (g1_fn(o, in.alpha) / std::fabs(cos_theta(o))) * d * std::fabs(cos_d);
    const float pdf_gtr = pdf_m / (4.0f * std::fabs(cos_d));
    pdf = pdf_diffuse * diffuse_prob + pdf_gtr * (1.0f - diffuse_prob);
    return spec_exact + diffuse;
}

f3 sample_direction(f3 o, f2 alpha, float metalness, Sampler& smp, f3& m, float& w, float& pdf)
{
    const float diffuse_prob = diffuse_sample_probability(metalness);
    const bool sample_diffuse = smp.s1() < diffuse_prob;
    const f2 u = smp.s2();
    f3 i;
    if (sample_diffuse) {
        i = cosine_weighted_rand_dir(u);
        m = normalize(o + i);
    } else {
        m = sample_visible_micronormal(o, alpha, u);
        i = 2.0f * dot(o, m) * m - o;
    }
    float pdf_diffuse = cos_theta(i) / PI_F;
    const float cos_d = dot(m, o);
    const float pdf_m = (g1_fn(o, alpha) / std::fabs(cos_theta(o))) * distribution(m, alpha) * std::fabs(cos_d);
    float pdf_gtr = pdf_m / (4.0f * std::fabs(cos_d));
    pdf_diffuse *= diffuse_prob;
    pdf_gtr *= (1.0f - diffuse_prob);
    pdf = sample_diffuse ? pdf_diffuse : pdf_gtr;
    w = pdf / (pdf_diffuse + pdf_gtr);
    return i;
}

// ------------------------------------------------------------------------------------------
// scene views and the acceleration structure

struct Vtx {
    f3 pos, nrm;
    f2 uv;
};

struct Instance {  // InstanceData, geometry.hlsli:4-11
    uint32_t vertex_index, index_index, emitter_index, material_index;
    float m[4][4];   // row-major M (object -> world)
    float n[3][4];   // row r = column r of M^-1: read as the normal matrix (A^-1)^T
};

struct Material {
    f4 base_colour, emission, rmnb, specular, sheen, cc;
};

struct AABBd {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    void grow(const double* p)
    {
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    }
    void grow(const AABBd& b)
    {
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    double area() const
    {
        double d[3];
        for (int k = 0; k < 3; k++)
            d[k] = std::max(0.0, hi[k] - lo[k]);
        return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
    }
};

struct BvhNode {
    AABBd box;
    int32_t left = -1, right = -1;  // children; leaf when left < 0
    uint32_t first = 0, count = 0;  // leaf primitive range in `order`
};

struct Bvh {
    std::vector<BvhNode> nodes;
    std::vector<uint32_t> order;

    void build(const std::vector<AABBd>& boxes)
    {
        const uint32_t n = uint32_t(boxes.size());
        order.resize(n);
        for (uint32_t i = 0; i < n; i++)
            order[i] = i;
        nodes.clear();
        nodes.reserve(2 * size_t(std::max(n, 1u)));
        std::vector<double> cen(size_t(n) * 3);
        for (uint32_t i = 0; i < n; i++)
            for (int k = 0; k < 3; k++)
                cen[size_t(i) * 3 + k] = 0.5 * (boxes[i].lo[k] + boxes[i].hi[k]);
        nodes.emplace_back();
        if (n == 0)
            return;
        struct Item {
            uint32_t node, first, count;
        };
        std::vector<Item> stack{{0, 0, n}};
        while (!stack.empty()) {
            Item it = stack.back();
            stack.pop_back();
            BvhNode& nd = nodes[it.node];
            AABBd cb;
            for (uint32_t i = 0; i < it.count; i++) {
                nd.box.grow(boxes[order[it.first + i]]);
                cb.grow(&cen[size_t(order[it.first + i]) * 3]);
            }
            if (it.count <= 2) {
                nd.first = it.first;
                nd.count = it.count;
                continue;
            }
            // binned SAH, 16 bins per axis
            const int B = 16;
            double best_cost = 1e300;
            int best_axis = -1, best_split = -1;
            for (int ax = 0; ax < 3; ax++) {
                const double ext = cb.hi[ax] - cb.lo[ax];
                if (ext <= 0)
                    continue;
                AABBd bb[B];
                uint32_t bc[B] = {};
                for (uint32_t i = 0; i < it.count; i++) {
                    uint32_t p = order[it.first + i];
                    int b = int((cen[size_t(p) * 3 + ax] - cb.lo[ax]) / ext * B);
                    b = std::min(std::max(b, 0), B - 1);
                    bb[b].grow(boxes[p]);
                    bc[b]++;
                }
                double la[B];
                uint32_t lc[B];
                AABBd acc;
                uint32_t cnt = 0;
                for (int b = 0; b < B; b++) {
                    if (bc[b])
                        acc.grow(bb[b]);
                    cnt += bc[b];
                    la[b] = cnt ? acc.area() : 0;
                    lc[b] = cnt;
                }
                AABBd racc;
                uint32_t rc = 0;
                for (int b = B - 1; b > 0; b--) {
                    if (bc[b])
                        racc.grow(bb[b]);
                    rc += bc[b];
                    if (lc[b - 1] == 0 || rc == 0)
                        continue;
                    double cost = la[b - 1] * lc[b - 1] + racc.area() * rc;
                    if (cost < best_cost) {
                        best_cost = cost;
                        best_axis = ax;
                        best_split = b;
                    }
                }
            }
            uint32_t mid;
            if (best_axis < 0) {
                mid = it.count / 2;  // all centroids equal: split in the middle
            } else {
                const double ext = cb.hi[best_axis] - cb.lo[best_axis];
                auto beg = order.begin() + it.first, end = beg + it.count;
                auto pivot = std::partition(beg, end, [&](uint32_t p) {
                    int b = int((cen[size_t(p) * 3 + best_axis] - cb.lo[best_axis]) / ext * B);
                    b = std::min(std::max(b, 0), B - 1);
                    return b < best_split;
                });
                mid = uint32_t(pivot - beg);
                if (mid == 0 || mid == it.count)
                    mid = it.count / 2;
            }
            const uint32_t l = uint32_t(nodes.size());
            nodes.emplace_back();
            nodes.emplace_back();
            nodes[it.node].left = int32_t(l);
            nodes[it.node].right = int32_t(l + 1);
            stack.push_back({l, it.first, mid});
            stack.push_back({l + 1, it.first + mid, it.count - mid});
        }
    }
};

struct RayD {
    double o[3], d[3], inv[3];
    double tmin, tmax;
};

inline bool slab(const AABBd& b, const RayD& r, double tmax)
{
    double t0 = r.tmin, t1 = tmax;
    for (int k = 0; k < 3; k++) {
        double a = (b.lo[k] - r.o[k]) * r.inv[k];
        double c = (b.hi[k] - r.o[k]) * r.inv[k];
        if (std::isnan(a) || std::isnan(c)) {  // 0 * inf: origin on the slab plane
            if (r.o[k] < b.lo[k] || r.o[k] > b.hi[k])
                return false;
            continue;
        }
        if (a > c)
            std::swap(a, c);
        t0 = std::max(t0, a);
        t1 = std::min(t1, c);
    }
    return t0 <= t1 * (1.0 + 1e-12) + 1e-300;
}

// Watertight ray/triangle intersection (Woop, Benthin, Wald 2013) in fp64.
inline bool tri_hit(const RayD& r, const double* v0, const double* v1, const double* v2, double& t, double& b1,
                    double& b2)
{
    int kz = 0;
    double ad = std::fabs(r.d[0]);
    if (std::fabs(r.d[1]) > ad) {
        kz = 1;
        ad = std::fabs(r.d[1]);
    }
    if (std::fabs(r.d[2]) > ad)
        kz = 2;
    int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
    if (r.d[kz] < 0)
        std::swap(kx, ky);
    const double Sx = r.d[kx] / r.d[kz], Sy = r.d[ky] / r.d[kz], Sz = 1.0 / r.d[kz];
    const double A[3] = {v0[0] - r.o[0], v0[1] - r.o[1], v0[2] - r.o[2]};
    const double Bv[3] = {v1[0] - r.o[0], v1[1] - r.o[1], v1[2] - r.o[2]};
    const double C[3] = {v2[0] - r.o[0], v2[1] - r.o[1], v2[2] - r.o[2]};
    const double Ax = A[kx] - Sx * A[kz], Ay = A[ky] - Sy * A[kz];
    const double Bx = Bv[kx] - Sx * Bv[kz], By = Bv[ky] - Sy * Bv[kz];
    const double Cx = C[kx] - Sx * C[kz], Cy = C[ky] - Sy * C[kz];
    const double U = Cx * By - Cy * Bx;
    const double V = Ax * Cy - Ay * Cx;
    const double W = Bx * Ay - By * Ax;
    if ((U < 0 || V < 0 || W < 0) && (U > 0 || V > 0 || W > 0))
        return false;
    const double det = U + V + W;
    if (det == 0.0)
        return false;
    const double T = U * Sz * A[kz] + V * Sz * Bv[kz] + W * Sz * C[kz];
    t = T / det;
    b1 = V / det;
    b2 = W / det;
    return true;
}

struct MeshAccel {
    Bvh bvh;
    std::vector<double> tri;  // 9 doubles per triangle (object space)
    AABBd box;
};

struct HitRec {
    double t;
    float u, v;
    uint32_t inst, prim;
};

}  // namespace

struct oracle_scene {
    std::vector<Vtx> vertices;
    std::vector<uint32_t> indices;
    std::vector<Instance> instances;
    std::vector<Material> materials;
    std::vector<uint32_t> light;
    uint32_t light_count = 0;
    std::vector<pt_mesh_desc> meshes;
    std::vector<uint32_t> instance_mesh;
    f3 envcol{0, 0, 0};

    std::vector<pt_texture_desc> tex;
    std::vector<uint8_t> texels;
    float srgb_lut[256];

    std::vector<MeshAccel> mesh_accel;
    std::vector<double> inst_inv;  // 12 doubles per instance: world->object 3x4 (row-major)

    // --- textures (material.hlsli:23-106): Texture2D.SampleLevel(sampler, uv, 0) with the
    // reference's sampler (src/graphics/sampler.cpp:8-22): bilinear, REPEAT, mip 0. Texel
    // centres at (i + 0.5) / size; sRGB textures decode each texel before filtering.
    static uint32_t bits_of(float f)
    {
        uint32_t u;
        memcpy(&u, &f, 4);
        return u;
    }
    f4 texel(const pt_texture_desc& d, uint32_t x, uint32_t y) const
    {
        const uint8_t* p = texels.data() + d.offset + (size_t(y) * d.width + x) * d.channels;
        const float k = 1.0f / 255.0f;
        if (d.channels == 4) {
            if (d.srgb)
                return {srgb_lut[p[0]], srgb_lut[p[1]], srgb_lut[p[2]], float(p[3]) * k};
            return {float(p[0]) * k, float(p[1]) * k, float(p[2]) * k, float(p[3]) * k};
        }
        if (d.channels == 2)
            return {float(p[0]) / 255.0f, float(p[1]) / 255.0f, 0.0f, 1.0f};
        return {float(p[0]) / 255.0f, 0.0f, 0.0f, 1.0f};
    }
    f4 sample_tex(uint32_t t, f2 uv) const
    {
        const pt_texture_desc& d = tex.at(t);
        const float u = uv.x - std::floor(uv.x), v = uv.y - std::floor(uv.y);
        const float x = u * float(d.width) - 0.5f, y = v * float(d.height) - 0.5f;
        const float fx = std::floor(x), fy = std::floor(y);
        const float ax = x - fx, ay = y - fy;
        const int64_t ix = int64_t(fx), iy = int64_t(fy);
        const uint32_t x0 = uint32_t((ix % int64_t(d.width) + d.width) % d.width);
        const uint32_t y0 = uint32_t((iy % int64_t(d.height) + d.height) % d.height);
        const uint32_t x1 = (x0 + 1) % d.width, y1 = (y0 + 1) % d.height;
        const f4 a = texel(d, x0, y0), b = texel(d, x1, y0), c = texel(d, x0, y1), e = texel(d, x1, y1);
        const float w00 = (1.0f - ax) * (1.0f - ay), w10 = ax * (1.0f - ay), w01 = (1.0f - ax) * ay, w11 = ax * ay;
        return {w00 * a.x + w10 * b.x + w01 * c.x + w11 * e.x, w00 * a.y + w10 * b.y + w01 * c.y + w11 * e.y,
                w00 * a.z + w10 * b.z + w01 * c.z + w11 * e.z, w00 * a.w + w10 * b.w + w01 * c.w + w11 * e.w};
    }
    f4 base_colour_at(const Material& m, f2 uv) const  // get_base_colour
    {
        if (bits_of(m.rmnb.w) & 1u)
            return sample_tex(bits_of(m.base_colour.x), uv);
        return m.base_colour;
    }
    f3 emission_at(const Material& m, f2 uv) const  // get_emission
    {
        if (bits_of(m.rmnb.w) & 2u) {
            const f4 t = sample_tex(bits_of(m.emission.x), uv);
            return mk(t.x, t.y, t.z) * m.emission.w;
        }
        return mk(m.emission.x, m.emission.y, m.emission.z) * m.emission.w;
    }
    f3 specular_at(const Material& m, f2 uv) const  // get_specular (material.hlsli:70-78)
    {
        if (bits_of(m.rmnb.w) & (1u << 5)) {
            const f4 t = sample_tex(bits_of(m.specular.x), uv);
            return mk(t.x, t.y, t.z);
        }
        return mk(m.specular.x, m.specular.y, m.specular.z);
    }
    f3 sheen_at(const Material& m, f2 uv) const  // get_sheen (material.hlsli:80-88)
    {
        if (bits_of(m.rmnb.w) & (1u << 7)) {
            const f4 t = sample_tex(bits_of(m.sheen.x), uv);
            return mk(t.x, t.y, t.z);
        }
        return mk(m.sheen.x, m.sheen.y, m.sheen.z);
    }
    float clearcoat_at(const Material& m, f2 uv) const  // get_clearcoat_anisotropy (material.hlsli:90-102)
    {
        return (bits_of(m.rmnb.w) & (1u << 9)) ? sample_tex(bits_of(m.cc.x), uv).x : m.cc.x;
    }
    // get_roughness_metalness_normal
    void rmn_at(const Material& m, f2 uv, float& roughness, float& metalness, f3& normal) const
    {
        const uint32_t b = bits_of(m.rmnb.w);
        if (b & (1u << 4)) {
            const f4 t = sample_tex(bits_of(m.rmnb.x), uv);
            roughness = t.y;
            metalness = t.z;
        } else {
            roughness = (b & (1u << 2)) ? sample_tex(bits_of(m.rmnb.x), uv).x : m.rmnb.x;
            metalness = (b & (1u << 3)) ? sample_tex(bits_of(m.rmnb.y), uv).x : m.rmnb.y;
        }
        if (b & (1u << 12)) {
            const f4 t = sample_tex(bits_of(m.rmnb.z), uv);
            normal = normalize(mk(t.x * 2.0f - 1.0f, t.y * 2.0f - 1.0f, t.z * 2.0f - 1.0f));
        } else {
            normal = mk(0.0f, 0.0f, 1.0f);
        }
    }
    Bvh top;
    std::vector<AABBd> inst_box;

    Vtx vertex(uint32_t i) const { return vertices[i]; }

    void build()
    {
        mesh_accel.resize(meshes.size());
        for (size_t mi = 0; mi < meshes.size(); mi++) {
            const pt_mesh_desc& md = meshes[mi];
            MeshAccel& ma = mesh_accel[mi];
            ma.tri.resize(size_t(md.tri_count) * 9);
            std::vector<AABBd> boxes(md.tri_count);
            for (uint32_t t = 0; t < md.tri_count; t++) {
                for (int c = 0; c < 3; c++) {
                    const Vtx& v = vertices[md.vertex_start + indices[md.index_start + t * 3 + c]];
                    double p[3] = {v.pos.x, v.pos.y, v.pos.z};
                    for (int k = 0; k < 3; k++)
                        ma.tri[size_t(t) * 9 + c * 3 + k] = p[k];
                    boxes[t].grow(p);
                }
                ma.box.grow(boxes[t]);
            }
            ma.bvh.build(boxes);
        }
        inst_inv.resize(instances.size() * 12);
        inst_box.resize(instances.size());
        for (size_t i = 0; i < instances.size(); i++) {
            const Instance& in = instances[i];
            double a[3][3], t[3];
            for (int r = 0; r < 3; r++) {
                for (int c = 0; c < 3; c++)
                    a[r][c] = in.m[r][c];
                t[r] = in.m[r][3];
            }
            const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                               a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                               a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
            double inv[3][3];
            inv[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / det;
            inv[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det;
            inv[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det;
            inv[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) / det;
            inv[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det;
            inv[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
            inv[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / det;
            inv[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
            inv[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
            for (int r = 0; r < 3; r++) {
                for (int c = 0; c < 3; c++)
                    inst_inv[i * 12 + r * 4 + c] = inv[r][c];
                inst_inv[i * 12 + r * 4 + 3] = -(inv[r][0] * t[0] + inv[r][1] * t[1] + inv[r][2] * t[2]);
            }
            const AABBd& mb = mesh_accel[instance_mesh[i]].box;
            for (int c = 0; c < 8; c++) {
                double p[3] = {(c & 1) ? mb.hi[0] : mb.lo[0], (c & 2) ? mb.hi[1] : mb.lo[1],
                               (c & 4) ? mb.hi[2] : mb.lo[2]};
                double w[3];
                for (int r = 0; r < 3; r++)
                    w[r] = a[r][0] * p[0] + a[r][1] * p[1] + a[r][2] * p[2] + t[r];
                inst_box[i].grow(w);
            }
            for (int k = 0; k < 3; k++) {  // pad for fp32 transform rounding
                const double pad = 1e-6 * (std::fabs(inst_box[i].lo[k]) + std::fabs(inst_box[i].hi[k]) + 1.0);
                inst_box[i].lo[k] -= pad;
                inst_box[i].hi[k] += pad;
            }
        }
        top.build(inst_box);
    }

    // Closest (any == false) or first (any == true) hit along a world ray in [tmin, tmax].
    // For any-hit rays on transparent instances, returns the shadow transmittance in *trans.
    bool trace(const float* o, const float* d, float tmin, float tmax, bool any, HitRec& hit, float* trans) const
    {
        RayD wr;
        for (int k = 0; k < 3; k++) {
            wr.o[k] = o[k];
            wr.d[k] = d[k];
            wr.inv[k] = 1.0 / wr.d[k];
        }
        wr.tmin = tmin;
        wr.tmax = tmax;
        bool found = false;
        double best = double(tmax);
        hit.t = best;
        if (trans)
            *trans = 1.0f;
        if (instances.empty())
            return false;
        std::vector<int32_t> stack{0};
        std::vector<int32_t> bstack;
        while (!stack.empty()) {
            const BvhNode& nd = top.nodes[stack.back()];
            stack.pop_back();
            if (!slab(nd.box, wr, best))
                continue;
            if (nd.left >= 0) {
                stack.push_back(nd.left);
                stack.push_back(nd.right);
                continue;
            }
            for (uint32_t k = 0; k < nd.count; k++) {
                const uint32_t ii = top.order[nd.first + k];
                const double* iv = &inst_inv[size_t(ii) * 12];
                RayD r;
                for (int q = 0; q < 3; q++) {
                    r.o[q] = iv[q * 4 + 0] * wr.o[0] + iv[q * 4 + 1] * wr.o[1] + iv[q * 4 + 2] * wr.o[2] + iv[q * 4 + 3];
                    r.d[q] = iv[q * 4 + 0] * wr.d[0] + iv[q * 4 + 1] * wr.d[1] + iv[q * 4 + 2] * wr.d[2];
                    r.inv[q] = 1.0 / r.d[q];
                }
                r.tmin = wr.tmin;
                r.tmax = wr.tmax;
                const MeshAccel& ma = mesh_accel[instance_mesh[ii]];
                const Material& mat = materials[instances[ii].material_index];
                uint32_t bits;
                memcpy(&bits, &mat.rmnb.w, 4);
                const bool transparent = (bits & (1u << 13)) != 0;
                if (ma.bvh.nodes.empty() || ma.tri.empty())
                    continue;
                bstack.assign(1, 0);
                while (!bstack.empty()) {
                    const BvhNode& bn = ma.bvh.nodes[bstack.back()];
                    bstack.pop_back();
                    if (!slab(bn.box, r, best))
                        continue;
                    if (bn.left >= 0) {
                        bstack.push_back(bn.left);
                        bstack.push_back(bn.right);
                        continue;
                    }
                    for (uint32_t q = 0; q < bn.count; q++) {
                        const uint32_t p = ma.bvh.order[bn.first + q];
                        const double* tv = &ma.tri[size_t(p) * 9];
                        double t, b1, b2;
                        if (!tri_hit(r, tv, tv + 3, tv + 6, t, b1, b2))
                            continue;
                        if (!(t >= double(tmin) && t <= best))
                            continue;
                        if (any && transparent && trans) {
                            // shadowray.ah.hlsl:12-40: T *= 1 - alpha (of the base colour at the
                            // hit), accept when T < 1e-5.
                            const uint32_t ix = instances[ii].index_index + p * 3;
                            const Vtx& a0 = vertices[instances[ii].vertex_index + indices[ix]];
                            const Vtx& a1 = vertices[instances[ii].vertex_index + indices[ix + 1]];
                            const Vtx& a2 = vertices[instances[ii].vertex_index + indices[ix + 2]];
                            const float fb1 = float(b1), fb2 = float(b2), fwa = 1.0f - fb1 - fb2;
                            const f2 huv = {fwa * a0.uv.x + fb1 * a1.uv.x + fb2 * a2.uv.x,
                                            fwa * a0.uv.y + fb1 * a1.uv.y + fb2 * a2.uv.y};
                            *trans *= (1.0f - base_colour_at(mat, huv).w);
                            if (*trans < 1e-5f) {
                                hit = {t, float(b1), float(b2), ii, p};
                                return true;
                            }
                            continue;
                        }
                        best = t;
                        hit = {t, float(b1), float(b2), ii, p};
                        found = true;
                        if (any)
                            return true;
                    }
                }
            }
        }
        return found;
    }
};

namespace {

// ------------------------------------------------------------------------------------------
// light sampling (lightsample.hlsli:7-143)

uint32_t sample_alias_table(const std::vector<uint32_t>& buf, uint32_t bin_count, uint32_t bin_size,
                            uint32_t table_offset, float& u, float& p)
{
    uint32_t i = std::min(bin_count - 1, uint32_t(u * float(bin_count)));
    uint32_t bo = table_offset + i * bin_size;
    u = std::min((u * float(bin_count)) - float(i), JUST_BELOW_ONE);
    const float q = as_float(buf[bo + 1]);
    if (u < q) {
        u = std::min(u / q, JUST_BELOW_ONE);
    } else {
        i = buf[bo + 2];
        bo = table_offset + i * bin_size;
        u = std::min((u - q) / (1.0f - q), JUST_BELOW_ONE);
    }
    p = as_float(buf[bo + 0]);
    return bo + 3;
}

float light_pdf(const oracle_scene& s, uint32_t emitter_index, uint32_t prim)
{
    if (emitter_index == PT_NO_EMITTER)
        return 0.0f;
    uint32_t bo = emitter_index * PT_LIGHT_BIN_WORDS;
    float p = as_float(s.light[bo + 0]);
    const uint32_t type = s.light[bo + 3];
    if (type == 3) {
        const uint32_t table_offset = s.light[bo + 4];
        bo = table_offset + prim * PT_EMITTER_BIN_WORDS;
        p *= as_float(s.light[bo + 0]);
    }
    return p;
}

f3 xform_point(const Instance& in, f3 p)
{
    return {in.m[0][0] * p.x + in.m[0][1] * p.y + in.m[0][2] * p.z + in.m[0][3],
            in.m[1][0] * p.x + in.m[1][1] * p.y + in.m[1][2] * p.z + in.m[1][3],
            in.m[2][0] * p.x + in.m[2][1] * p.y + in.m[2][2] * p.z + in.m[2][3]};
}

f3 xform_vec3(const Instance& in, f3 v)
{
    return {in.m[0][0] * v.x + in.m[0][1] * v.y + in.m[0][2] * v.z,
            in.m[1][0] * v.x + in.m[1][1] * v.y + in.m[1][2] * v.z,
            in.m[2][0] * v.x + in.m[2][1] * v.y + in.m[2][2] * v.z};
}

f3 xform_normal(const Instance& in, f3 v)
{
    return {in.n[0][0] * v.x + in.n[0][1] * v.y + in.n[0][2] * v.z,
            in.n[1][0] * v.x + in.n[1][1] * v.y + in.n[1][2] * v.z,
            in.n[2][0] * v.x + in.n[2][1] * v.y + in.n[2][2] * v.z};
}

inline f3 emission_of(const Material& m) { return mk(m.emission.x, m.emission.y, m.emission.z) * m.emission.w; }

f3 sample_light(const oracle_scene& s, f3 u, f3 pt, f3& light_dir, float& light_dist, float& pdf, bool& is_delta,
               uint32_t flags = 0)
{
    uint32_t bo = sample_alias_table(s.light, s.light_count, PT_LIGHT_BIN_WORDS, 0, u.x, pdf);
    const uint32_t type = s.light[bo + 0];
    if (type == 3) {
        const uint32_t table_offset = s.light[bo + 1];
        const uint32_t tri_count = s.light[bo + 2];
        const uint32_t inst_off = s.light[bo + 3];
        const uint32_t inst_count = s.light[bo + 4];
        float p;
        bo = sample_alias_table(s.light, tri_count, PT_EMITTER_BIN_WORDS, table_offset, u.x, p);
        pdf *= p;
        const uint32_t inst = inst_off + std::min(inst_count - 1, uint32_t(u.x * float(inst_count)));
        const Instance& in = s.instances[inst];
        const Vtx va = s.vertex(in.vertex_index + s.light[bo + 0]);
        const Vtx vb = s.vertex(in.vertex_index + s.light[bo + 1]);
        const Vtx vc = s.vertex(in.vertex_index + s.light[bo + 2]);
        const f3 pa = xform_point(in, va.pos), pb = xform_point(in, vb.pos), pc = xform_point(in, vc.pos);
        const f3 bary = random_barycentric({u.y, u.z});
        const f3 on_tri = bary.x * pa + bary.y * pb + bary.z * pc;
        light_dir = on_tri - pt;
        light_dist = 1.0f;
        const f3 tn = cross(pb - pa, pc - pa);
        const float area = std::sqrt(dot(tn, tn)) * 0.5f;
        const float cosine = std::fabs(dot(normalize(light_dir), normalize(tn)));
        const float dist2 = dot(light_dir, light_dir);
        pdf *= ((flags & ORACLE_PLANT_NO_DIST2) ? 1.0f : dist2) / (area * cosine);
        is_delta = false;
        if (!std::isfinite(pdf))
            return mk(0, 0, 0);
        const f2 luv = {bary.x * va.uv.x + bary.y * vb.uv.x + bary.z * vc.uv.x,
                        bary.x * va.uv.y + bary.y * vb.uv.y + bary.z * vc.uv.y};
        return s.emission_at(s.materials[in.material_index], luv);
    } else if (type == 1) {
        const f3 lp = mk(as_float(s.light[bo + 1]), as_float(s.light[bo + 2]), as_float(s.light[bo + 3]));
        light_dir = lp - pt;
        light_dist = 1.0f;
        is_delta = true;
        return (1.0f / dot(light_dir, light_dir)) *
               mk(as_float(s.light[bo + 4]), as_float(s.light[bo + 5]), as_float(s.light[bo + 6]));
    } else if (type == 2) {
        light_dir = mk(as_float(s.light[bo + 1]), as_float(s.light[bo + 2]), as_float(s.light[bo + 3]));
        light_dist = INFINITY;
        is_delta = true;
        return mk(as_float(s.light[bo + 4]), as_float(s.light[bo + 5]), as_float(s.light[bo + 6]));
    }
    light_dist = 0.0f;
    is_delta = true;
    return mk(0, 0, 0);
}

// ------------------------------------------------------------------------------------------
// the path (pathtrace.rg.hlsl, pathtrace.ch.hlsl, pathtrace.ms.hlsl, pathtrace.hlsli)

struct Payload {
    f3 throughput;
    bool final_segment;
    f3 radiance;
    float brdf_pdf;  // -1 before the first bounce
    Sampler smp;
    f3 next_origin, next_dir;
    uint32_t flags;  // ORACLE_NO_NEE / ORACLE_NO_RR / ORACLE_PLANT_* (oracle.h)
};

struct Counters {
    uint64_t ext = 0, shadow = 0;
};

void camera_ray(const pt_camera& c, uint32_t px, uint32_t py, uint32_t w, uint32_t h, f2 u_pixel, f2 u_lens, f3& origin,
                f3& dir)
{
    const float fx = float(px) + u_pixel.x, fy = float(py) + u_pixel.y;
    const float nx = (fx / float(w)) * 2.0f - 1.0f;
    const float ny = (fy / float(h)) * 2.0f - 1.0f;
    const float* P = c.inv_proj;
    f3 target = {P[0] * nx + P[1] * ny + P[2] * 1.0f + P[3] * 1.0f, P[4] * nx + P[5] * ny + P[6] * 1.0f + P[7] * 1.0f,
                 P[8] * nx + P[9] * ny + P[10] * 1.0f + P[11] * 1.0f};
    f3 ro = {0, 0, 0}, rd = target;
    if (c.lens_radius > 0.0f) {
        const f2 disk = sample_concentric_disk(u_lens);
        const f2 lens = {c.lens_radius * disk.x, c.lens_radius * disk.y};
        const float t_focus = c.focus_dist / -rd.z;
        const f3 fp = ro + t_focus * rd;
        ro = {lens.x, lens.y, 0.0f};
        rd = fp - ro;
    }
    const float* V = c.inv_view;
    origin = {V[0] * ro.x + V[1] * ro.y + V[2] * ro.z + V[3], V[4] * ro.x + V[5] * ro.y + V[6] * ro.z + V[7],
              V[8] * ro.x + V[9] * ro.y + V[10] * ro.z + V[11]};
    dir = normalize(mk(V[0] * rd.x + V[1] * rd.y + V[2] * rd.z, V[4] * rd.x + V[5] * rd.y + V[6] * rd.z,
                       V[8] * rd.x + V[9] * rd.y + V[10] * rd.z));
}

f3 light_contribution(const oracle_scene& s, const BrdfInput& in, f3 o, f3 x, const f3 fw[3], Sampler& smp,
                      Counters& cnt, uint32_t flags)
{
    if (s.light_count == 0)
        return mk(0, 0, 0);
    const float u0 = smp.s1();
    const f2 u12 = smp.s2();
    f3 light_dir;
    float light_dist, pdf;
    bool is_delta;
    f3 emission = sample_light(s, mk(u0, u12.x, u12.y), x, light_dir, light_dist, pdf, is_delta, flags);
    if (maxc(emission) <= 0.0f)
        return mk(0, 0, 0);
    f3 i = {dot(fw[0], light_dir), dot(fw[1], light_dir), dot(fw[2], light_dir)};
    if (i.z <= 0.0f)
        return mk(0, 0, 0);
    cnt.shadow++;
    HitRec h;
    float trans = 1.0f;
    const float o3[3] = {x.x, x.y, x.z}, d3[3] = {light_dir.x, light_dir.y, light_dir.z};
    const bool occluded = s.trace(o3, d3, 0.0f, light_dist - 1e-5f, true, h, &trans);
    if (occluded || trans < 1e-5f)
        return mk(0, 0, 0);
    emission = emission * trans;
    i = normalize(i);
    const f3 m = normalize(i + o);
    float pdf_brdf;
    const f3 brdf = evaluate_brdf(in, i, o, m, pdf_brdf);
    if (is_delta)
        return emission * brdf * cos_theta(i) / pdf;
    const float w = pdf / (pdf + pdf_brdf);
    return w * emission * brdf * cos_theta(i) / pdf;
}

void closest_hit(const oracle_scene& s, Payload& pl, f3 ro, f3 rd, const HitRec& h, Counters& cnt)
{
    const Instance& in = s.instances[h.inst];
    const Material& mat = s.materials[in.material_index];
    const uint32_t ib = in.index_index + h.prim * 3;
    const Vtx va = s.vertex(in.vertex_index + s.indices[ib + 0]);
    const Vtx vb = s.vertex(in.vertex_index + s.indices[ib + 1]);
    const Vtx vc = s.vertex(in.vertex_index + s.indices[ib + 2]);
    const float t = float(h.t);
    // get_vertex_attributes (geometry.hlsli:39-48)
    const float alpha_b = 1.0f - h.u - h.v;
    const f3 obj_n = normalize(alpha_b * va.nrm + h.u * vb.nrm + h.v * vc.nrm);
    const f2 uv = {alpha_b * va.uv.x + h.u * vb.uv.x + h.v * vc.uv.x, alpha_b * va.uv.y + h.u * vb.uv.y + h.v * vc.uv.y};

    const f3 emission = s.emission_at(mat, uv);
    // ObjectRayDirection() = world->object linear part applied to the world direction.
    const double* iv = &s.inst_inv[size_t(h.inst) * 12];
    const f3 obj_dir = {float(iv[0] * rd.x + iv[1] * rd.y + iv[2] * rd.z), float(iv[4] * rd.x + iv[5] * rd.y + iv[6] * rd.z),
                        float(iv[8] * rd.x + iv[9] * rd.y + iv[10] * rd.z)};
    if (maxc(emission) > 0.0f && dot(obj_dir, obj_n) < 0.0f) {
        if (pl.brdf_pdf < 0.0f || (pl.flags & ORACLE_NO_NEE)) {
            pl.radiance = pl.radiance + pl.throughput * emission;
        } else {
            const f3 wab = xform_vec3(in, vb.pos - va.pos);
            const f3 wac = xform_vec3(in, vc.pos - va.pos);
            const f3 sn = cross(wab, wac);
            const float area = std::sqrt(dot(sn, sn)) * 0.5f;
            const float dist2 = t * t;
            const float cosine = std::fabs(dot(rd, normalize(sn)));
            float pdf_l = light_pdf(s, in.emitter_index, h.prim);
            pdf_l *= dist2 / (area * cosine);
            if (pdf_l > 0.0f && dist2 > 0.0f) {
                const float w = ((pl.flags & ORACLE_PLANT_MIS_SWAP) ? pdf_l : pl.brdf_pdf) / (pl.brdf_pdf + pdf_l);
                pl.radiance = pl.radiance + pl.throughput * (w * emission);
            }
        }
    }
    if (pl.final_segment)
        return;

    const f4 base = s.base_colour_at(mat, uv);
    if (base.w < pl.smp.s1()) {
        pl.next_origin = ro + rd * (t + ORIGIN_OFFSET);
        pl.next_dir = rd;
        return;
    }
    BrdfInput bi;
    bi.base_colour = mk(base.x, base.y, base.z);
    float roughness;
    f3 mapped_n;
    s.rmn_at(mat, uv, roughness, bi.metalness, mapped_n);
    bi.specular = s.specular_at(mat, uv);
    bi.clearcoat = s.clearcoat_at(mat, uv);
    const float cc_rough = mat.cc.y;
    const float anisotropy = mat.cc.z;
    bi.clearcoat_alpha = lerpf(0.001f, 0.1f, cc_rough * cc_rough);
    bi.sheen = s.sheen_at(mat, uv);
    roughness = std::max(0.00001f, roughness * roughness);
    const float aspect = std::sqrt(1.0f - 0.9f * std::fabs(anisotropy));
    bi.alpha = {roughness / aspect, roughness * aspect};
    if (anisotropy < 0.0f)
        bi.alpha = {bi.alpha.y, bi.alpha.x};

    // get_vertex_vectors (geometry.hlsli:50-68)
    const f2 duv_ab = {vb.uv.x - va.uv.x, vb.uv.y - va.uv.y};
    const f2 duv_ac = {vc.uv.x - va.uv.x, vc.uv.y - va.uv.y};
    const f3 ab = vb.pos - va.pos, ac = vc.pos - va.pos;
    const f3 true_n = cross(ab, ac);
    const float denom = duv_ab.x * duv_ac.y - duv_ac.x * duv_ab.y;
    f3 tangent;
    if (std::fabs(denom) < 1e-5f) {
        tangent = cross(obj_n, std::fabs(obj_n.y) > 1e-5f ? mk(1, 0, 0) : mk(0, 1, 0));
    } else {
        tangent = (duv_ac.y * ab - duv_ab.y * ac) / denom;
        tangent = tangent - dot(tangent, obj_n) * obj_n;
    }
    const f3 bitangent = cross(obj_n, tangent);

    f3 wx = normalize(xform_normal(in, tangent));
    f3 wy = normalize(xform_normal(in, bitangent));
    f3 wn = normalize(xform_normal(in, obj_n));
    if (dot(wn, rd) > 0.0f) {
        wx = -wx;
        wy = -wy;
        wn = -wn;
    }
    f3 twn = normalize(xform_normal(in, true_n));
    if (dot(rd, twn) > 0.0f)
        twn = -twn;
    if (mapped_n.z < 0.999f) {
        // from_world_space = float3x3(mx, my, mn) * float3x3(wx, wy, wn) (ch.hlsl:321-326), with
        // the orthonormal basis of Duff et al. 2017 (util.hlsli:57-64)
        const float sg = mapped_n.z < 0.0f ? -1.0f : 1.0f;
        const float a = -1.0f / (sg + mapped_n.z);
        const float b = mapped_n.x * mapped_n.y * a;
        const f3 mx = mk(1.0f + sg * mapped_n.x * mapped_n.x * a, sg * b, -sg * mapped_n.x);
        const f3 my = mk(b, sg + mapped_n.y * mapped_n.y * a, -mapped_n.y);
        const f3 fx = mx.x * wx + mx.y * wy + mx.z * wn;
        const f3 fy = my.x * wx + my.y * wy + my.z * wn;
        const f3 fn = mapped_n.x * wx + mapped_n.y * wy + mapped_n.z * wn;
        wx = fx;
        wy = fy;
        wn = fn;
    }
    const f3 fw[3] = {wx, wy, wn};  // rows of from_world_space
    const f3 nrd = -rd;
    const f3 o = normalize(mk(dot(wx, nrd), dot(wy, nrd), dot(wn, nrd)));
    if (o.z <= 0.0f) {
        pl.final_segment = true;
        return;
    }
    const f3 x = ro + rd * t + twn * ORIGIN_OFFSET;
    if (!(pl.flags & ORACLE_NO_NEE))  // SAMPLE_LIGHTS (pathtrace.ch.hlsl:8,338-340)
        pl.radiance = pl.radiance + pl.throughput * light_contribution(s, bi, o, x, fw, pl.smp, cnt, pl.flags);

    f3 m;
    float w, pdf;
    const f3 i = sample_direction(o, bi.alpha, bi.metalness, pl.smp, m, w, pdf);
    if (i.z * o.z <= 0.0f) {
        pl.final_segment = true;
        return;
    }
    float pdf_brdf;
    const f3 brdf = evaluate_brdf(bi, i, o, m, pdf_brdf);
    const f3 thr = w * brdf * cos_theta(i) / pdf;
    pl.brdf_pdf = pdf_brdf;
    pl.throughput = pl.throughput * thr;
    pl.next_origin = x;
    pl.next_dir = normalize(wx * i.x + wy * i.y + wn * i.z);
}

bool trace_segment(const oracle_scene& s, Payload& pl, f3 ro, f3 rd, float tmin, float tmax, Counters& cnt)
{
    cnt.ext++;
    HitRec h;
    const float o3[3] = {ro.x, ro.y, ro.z}, d3[3] = {rd.x, rd.y, rd.z};
    if (!s.trace(o3, d3, tmin, tmax, false, h, nullptr)) {
        pl.radiance = pl.radiance + pl.throughput * s.envcol;  // pathtrace.ms.hlsl:7-11
        pl.final_segment = true;
        return false;
    }
    closest_hit(s, pl, ro, rd, h, cnt);
    return true;
}

bool russian_roulette(Payload& pl)
{
    const float mt = maxc(pl.throughput);
    const float u = pl.smp.s1();
    if (mt < 1.0f) {
        const float q = std::max(0.0f, 1.0f - mt);
        if (u < q)
            return true;
        pl.throughput = pl.throughput * (1.0f / ((pl.flags & ORACLE_PLANT_RR_Q) ? q : 1.0f - q));
    }
    return false;
}

inline float correct(float x) { return (x < 0.0f || !std::isfinite(x)) ? 0.0f : x; }

f3 trace_path(const oracle_scene& s, const pt_camera& c, uint32_t px, uint32_t py, uint32_t w, uint32_t h, uint32_t si,
              uint32_t max_bounces, Counters& cnt, uint32_t flags)
{
    Payload pl;
    pl.flags = flags;
    pl.radiance = {0, 0, 0};
    pl.smp.start(px, py, si, 0);
    const f2 up = pl.smp.s2();
    const f2 ul = pl.smp.s2();
    f3 ro, rd;
    camera_ray(c, px, py, w, h, up, ul, ro, rd);
    pl.smp.start(px, py, si, 1);
    pl.throughput = {1, 1, 1};
    pl.brdf_pdf = -1.0f;
    pl.final_segment = (max_bounces == 0);
    trace_segment(s, pl, ro, rd, c.near_clip, c.far_clip, cnt);
    if (!pl.final_segment) {
        for (uint32_t seg = 1; seg <= max_bounces; seg++) {
            pl.smp.start(px, py, si, seg + 1);
            ro = pl.next_origin;
            rd = pl.next_dir;
            pl.final_segment = seg >= max_bounces;
            trace_segment(s, pl, ro, rd, 0.0f, INFINITY, cnt);
            if (pl.final_segment || (!(pl.flags & ORACLE_NO_RR) && russian_roulette(pl)))
                break;
        }
    }
    const f3 r = pl.radiance * c.exposure;
    return {correct(r.x), correct(r.y), correct(r.z)};
}

oracle_scene* make_scene(const pt_scene_blob* b)
{
    auto s = std::make_unique<oracle_scene>();
    s->vertices.resize(b->vertex_count);
    memcpy(s->vertices.data(), b->vertices, b->vertex_count * 32);
    s->indices.assign(b->indices, b->indices + b->index_count);
    s->instances.resize(b->instance_count);
    memcpy(s->instances.data(), b->instances, size_t(b->instance_count) * 128);
    s->materials.resize(b->material_count);
    memcpy(s->materials.data(), b->materials, size_t(b->material_count) * 96);
    if (b->light_word_count)
        s->light.assign(b->light_words, b->light_words + b->light_word_count);
    s->light_count = b->light_count;
    s->meshes.assign(b->meshes, b->meshes + b->mesh_count);
    s->instance_mesh.assign(b->instance_mesh, b->instance_mesh + b->instance_count);
    s->envcol = {b->envcol[0], b->envcol[1], b->envcol[2]};
    if (b->texture_count)
        s->tex.assign(b->textures, b->textures + b->texture_count);
    if (b->texel_bytes)
        s->texels.assign(b->texels, b->texels + b->texel_bytes);
    for (int i = 0; i < 256; i++) {  // the sRGB transfer function VK_FORMAT_R8G8B8A8_SRGB decodes with
        const double c = i / 255.0;
        s->srgb_lut[i] = float(c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4));
    }
    s->build();
    return s.release();
}

}  // namespace

static_assert(sizeof(Instance) == 128, "instance layout");
static_assert(sizeof(Material) == 96, "material layout");
static_assert(sizeof(Vtx) == 32, "vertex layout");

extern "C" {

int oracle_scene_create(const pt_scene_blob* blob, oracle_scene** out)
{
    if (!blob || !out)
        return PT_EINVAL;
    try {
        *out = make_scene(blob);
        return PT_OK;
    } catch (...) {
        return PT_EINVAL;
    }
}

void oracle_scene_free(oracle_scene* s) { delete s; }

int oracle_render(const oracle_scene* s, const pt_camera* cam, uint32_t width, uint32_t height, uint32_t sample_index,
                  uint32_t spp, uint32_t max_bounces, float* accum, uint32_t rows_step, int threads,
                  oracle_counts* counts)
{
    return oracle_render_ex(s, cam, width, height, sample_index, spp, max_bounces, accum, rows_step, threads, counts,
                            0u);
}

int oracle_render_ex(const oracle_scene* s, const pt_camera* cam, uint32_t width, uint32_t height, uint32_t sample_index,
                     uint32_t spp, uint32_t max_bounces, float* accum, uint32_t rows_step, int threads,
                     oracle_counts* counts, uint32_t flags)
{
    if (!s || !cam || !accum || spp == 0 || width == 0 || height == 0)
        return PT_EINVAL;
    if (rows_step == 0)
        rows_step = 1;
    const int nt = threads > 0 ? threads : omp_get_max_threads();
    std::atomic<uint64_t> ext{0}, sh{0}, paths{0};
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int64_t y = 0; y < int64_t(height); y++) {
        if (uint32_t(y) % rows_step != 0)
            continue;
        Counters c;
        uint64_t np = 0;
        for (uint32_t x = 0; x < width; x++) {
            float* px = accum + (size_t(y) * width + x) * 4;
            for (uint32_t k = 0; k < spp; k++) {
                const uint32_t si = sample_index + k;
                const f3 r = trace_path(*s, *cam, x, uint32_t(y), width, height, si, max_bounces, c, flags);
                np++;
                // pathtrace.rg.hlsl:61-67 with samples == 1 per dispatch
                if (si == 0) {
                    px[0] = r.x;
                    px[1] = r.y;
                    px[2] = r.z;
                    px[3] = 1.0f;
                } else {
                    const float n = float(si + 1);
                    px[0] = px[0] + (r.x - 1.0f * px[0]) / n;
                    px[1] = px[1] + (r.y - 1.0f * px[1]) / n;
                    px[2] = px[2] + (r.z - 1.0f * px[2]) / n;
                }
            }
        }
        ext += c.ext;
        sh += c.shadow;
        paths += np;
    }
    if (counts) {
        counts->extension_rays = ext;
        counts->shadow_rays = sh;
        counts->paths = paths;
    }
    return PT_OK;
}

int oracle_primary(const oracle_scene* s, const pt_camera* cam, uint32_t width, uint32_t height, uint32_t sample_index,
                   pt_hit* hits, pt_ray* rays_out)
{
    if (!s || !cam || !hits)
        return PT_EINVAL;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t y = 0; y < int64_t(height); y++) {
        for (uint32_t x = 0; x < width; x++) {
            Sampler smp;
            smp.start(x, uint32_t(y), sample_index, 0);
            const f2 up = smp.s2();
            const f2 ul = smp.s2();
            f3 ro, rd;
            camera_ray(*cam, x, uint32_t(y), width, height, up, ul, ro, rd);
            const float o3[3] = {ro.x, ro.y, ro.z}, d3[3] = {rd.x, rd.y, rd.z};
            HitRec h;
            pt_hit& out = hits[size_t(y) * width + x];
            if (s->trace(o3, d3, cam->near_clip, cam->far_clip, false, h, nullptr))
                out = {float(h.t), h.u, h.v, h.inst, h.prim};
            else
                out = {0.0f, 0.0f, 0.0f, PT_NO_HIT, PT_NO_HIT};
            if (rays_out) {
                pt_ray& r = rays_out[size_t(y) * width + x];
                r = {{ro.x, ro.y, ro.z}, cam->near_clip, {rd.x, rd.y, rd.z}, cam->far_clip};
            }
        }
    }
    return PT_OK;
}

int oracle_trace_rays(const oracle_scene* s, const pt_ray* rays, uint32_t count, pt_hit* hits, int any_hit)
{
    if (!s || (!rays && count) || (!hits && count))
        return PT_EINVAL;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t k = 0; k < int64_t(count); k++) {
        HitRec h;
        const pt_ray& r = rays[k];
        if (s->trace(r.origin, r.dir, r.tmin, r.tmax, any_hit != 0, h, nullptr))
            hits[k] = {float(h.t), h.u, h.v, h.inst, h.prim};
        else
            hits[k] = {0.0f, 0.0f, 0.0f, PT_NO_HIT, PT_NO_HIT};
    }
    return PT_OK;
}

int oracle_sampler_draw(uint32_t px, uint32_t py, uint32_t sample_index, uint32_t seed, const uint8_t* pattern,
                        uint32_t n_calls, float* out)
{
    Sampler smp;
    smp.start(px, py, sample_index, seed);
    size_t k = 0;
    for (uint32_t c = 0; c < n_calls; c++) {
        if (pattern[c] == 1) {
            out[k++] = smp.s1();
        } else {
            f2 v = smp.s2();
            out[k++] = v.x;
            out[k++] = v.y;
        }
    }
    return PT_OK;
}

int oracle_eval_brdf_many(const float* p, size_t n, const float* i, const float* o, float* rgb_out)
{
    BrdfInput in;
    in.base_colour = mk(p[0], p[1], p[2]);
    in.alpha = {p[3], p[4]};
    in.metalness = p[5];
    in.specular = mk(p[6], p[7], p[8]);
    in.sheen = mk(p[9], p[10], p[11]);
    in.clearcoat = p[12];
    in.clearcoat_alpha = p[13];
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < int64_t(n); k++) {
        const f3 ii = mk(i[3 * k], i[3 * k + 1], i[3 * k + 2]), oo = mk(o[3 * k], o[3 * k + 1], o[3 * k + 2]);
        float pdf;
        const f3 r = evaluate_brdf(in, ii, oo, normalize(ii + oo), pdf);
        rgb_out[3 * k] = r.x;
        rgb_out[3 * k + 1] = r.y;
        rgb_out[3 * k + 2] = r.z;
    }
    return PT_OK;
}

float oracle_light_pdf(const oracle_scene* s, uint32_t emitter_index, uint32_t prim)
{
    return s ? light_pdf(*s, emitter_index, prim) : 0.0f;
}

int oracle_eval_brdf(const float* p, const float* i, const float* o, const float* m, float* rgb_out, float* pdf_out)
{
    BrdfInput in;
    in.base_colour = mk(p[0], p[1], p[2]);
    in.alpha = {p[3], p[4]};
    in.metalness = p[5];
    in.specular = mk(p[6], p[7], p[8]);
    in.sheen = mk(p[9], p[10], p[11]);
    in.clearcoat = p[12];
    in.clearcoat_alpha = p[13];
    float pdf;
    const f3 r = evaluate_brdf(in, mk(i[0], i[1], i[2]), mk(o[0], o[1], o[2]), mk(m[0], m[1], m[2]), pdf);
    rgb_out[0] = r.x;
    rgb_out[1] = r.y;
    rgb_out[2] = r.z;
    *pdf_out = pdf;
    return PT_OK;
}

int oracle_sample_direction(const float* o, const float* alpha, float metalness, const float* u3, float* i_out,
                            float* m_out, float* w_out, float* pdf_out)
{
    // sample_direction with its three uniforms given explicitly.
    const float diffuse_prob = diffuse_sample_probability(metalness);
    const bool sample_diffuse = u3[0] < diffuse_prob;
    const f2 u = {u3[1], u3[2]};
    const f3 ov = mk(o[0], o[1], o[2]);
    const f2 al = {alpha[0], alpha[1]};
    f3 i, m;
    if (sample_diffuse) {
        i = cosine_weighted_rand_dir(u);
        m = normalize(ov + i);
    } else {
        m = sample_visible_micronormal(ov, al, u);
        i = 2.0f * dot(ov, m) * m - ov;
    }
    float pdf_diffuse = cos_theta(i) / PI_F;
    const float cos_d = dot(m, ov);
    const float pdf_m = (g1_fn(ov, al) / std::fabs(cos_theta(ov))) * distribution(m, al) * std::fabs(cos_d);
    float pdf_gtr = pdf_m / (4.0f * std::fabs(cos_d));
    pdf_diffuse *= diffuse_prob;
    pdf_gtr *= (1.0f - diffuse_prob);
    const float pdf = sample_diffuse ? pdf_diffuse : pdf_gtr;
    i_out[0] = i.x;
    i_out[1] = i.y;
    i_out[2] = i.z;
    m_out[0] = m.x;
    m_out[1] = m.y;
    m_out[2] = m.z;
    *w_out = pdf / (pdf_diffuse + pdf_gtr);
    *pdf_out = pdf;
    return PT_OK;
}

int oracle_camera_ray(const pt_camera* cam, uint32_t px, uint32_t py, uint32_t width, uint32_t height,
                      const float* u_pixel, const float* u_lens, float* origin, float* dir)
{
    f3 ro, rd;
    camera_ray(*cam, px, py, width, height, {u_pixel[0], u_pixel[1]}, {u_lens[0], u_lens[1]}, ro, rd);
    origin[0] = ro.x;
    origin[1] = ro.y;
    origin[2] = ro.z;
    dir[0] = rd.x;
    dir[1] = rd.y;
    dir[2] = rd.z;
    return PT_OK;
}

int oracle_sample_light(const oracle_scene* s, const float* point, const float* u3, float* emission, float* light_dir,
                        float* light_dist, float* pdf)
{
    if (!s || s->light_count == 0)
        return PT_ESTATE;
    f3 ld;
    bool delta;
    const f3 e = sample_light(*s, mk(u3[0], u3[1], u3[2]), mk(point[0], point[1], point[2]), ld, *light_dist, *pdf, delta);
    emission[0] = e.x;
    emission[1] = e.y;
    emission[2] = e.z;
    light_dir[0] = ld.x;
    light_dir[1] = ld.y;
    light_dir[2] = ld.z;
    return PT_OK;
}

// The PNG path's ACES tone mapping (shaders/compute/tonemap.cs.hlsl:30-66 with
// src_is_srgb = false and dst_is_srgb = true, as PathTraceDispatcher constructs its
// ToneMapper, src/pathtrace/ptdispatch.cpp:29) and write_png's per-channel
// (uchar)clamp(x * 255, 0, 255) (src/io/imagewrite.cpp:39-42); alpha passes through.
// n RGBA float pixels -> n RGBA8 pixels.
int oracle_tonemap(const float* rgba, size_t n, uint8_t* out)
{
    auto to_srgb = [](float x) {  // tonemap.cs.hlsl:30-32, threshold as the shader writes it
        return x <= 0.00031308f ? 12.92f * x : 1.055f * std::pow(x, 1.0f / 2.4f) - 0.055f;
    };
    auto fit = [](float v) {  // rrt_to_odt_fit, tonemap.cs.hlsl:48-53
        const float a = v * (v + 0.0245786f) - 0.000090537f;
        const float b = v * (0.983729f * v + 0.4329510f) + 0.238081f;
        return a / b;
    };
    auto byte = [](float x) {
        const float y = std::min(std::max(x * 255.0f, 0.0f), 255.0f);
        return uint8_t(y);
    };
    static const float rrt[3][3] = {{0.59719f, 0.35458f, 0.04823f},
                                    {0.07600f, 0.90834f, 0.01566f},
                                    {0.02840f, 0.13383f, 0.83777f}};
    static const float odt[3][3] = {{1.60475f, -0.53108f, -0.07367f},
                                    {-0.10208f, 1.10813f, -0.00605f},
                                    {-0.00327f, -0.07276f, 1.07602f}};
    for (size_t i = 0; i < n; i++) {
        const float* c = rgba + i * 4;
        const float s[3] = {to_srgb(c[0]), to_srgb(c[1]), to_srgb(c[2])};
        float f[3];
        for (int r = 0; r < 3; r++)
            f[r] = fit(rrt[r][0] * s[0] + rrt[r][1] * s[1] + rrt[r][2] * s[2]);
        for (int r = 0; r < 3; r++) {
            const float o = odt[r][0] * f[0] + odt[r][1] * f[1] + odt[r][2] * f[2];
            out[i * 4 + r] = byte(std::min(std::max(o, 0.0f), 1.0f));  // saturate (NaN -> 0)
        }
        out[i * 4 + 3] = byte(c[3]);
    }
    return 0;
}

}  // extern "C"
