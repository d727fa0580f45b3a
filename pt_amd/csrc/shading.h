// Device-side sampler, camera, BSDF and light sampling of the path tracer.
//
// Semantics follow the reference shaders line by line (cited per function) so that, with
// the same seeds, a path makes the same decisions as the reference's megakernel. The
// arithmetic is fp32 as in the shaders.
#pragma once

#include "common.h"

namespace ptk {

constexpr float PI_F = 3.141592654f;  // util.hlsli:4
constexpr float ORIGIN_OFFSET = 1e-5f; // pathtrace.hlsli:6
__device__ __forceinline__ float just_below_one() { return u32_as_float(0x3F7FFFFFu); }

// --- hashing (util.hlsli:8-27) ------------------------------------------------------------
__device__ __forceinline__ uint32_t hash_combine(uint32_t seed, uint32_t x)
{
    return seed ^ (x + (seed << 6) + (seed >> 2));
}

__device__ __forceinline__ uint32_t hash_u32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x85ebca6bu;
    x ^= x >> 13;
    x *= 0xc2b2ae35u;
    x ^= x >> 16;
    return x;
}

// --- Sobol + hash-based Owen scrambling (sample.hlsli:94-166) -----------------------------

// Direction numbers of Sobol dimensions 0-3 (primitive polynomials 1, x+1, x^2+x+1,
// x^3+x+1 with initial numbers 1 | 1,3 | 1,3,1), filled in by the host at library load.
// Sobol dimensions 2 and 3: the XOR of the direction numbers that each byte value selects, per
// byte position of the index (c_sobol_bytes[dim - 2][byte position][byte value]).
extern __constant__ uint32_t c_sobol_bytes[2][4][256];

__device__ __forceinline__ uint32_t lk_permutation(uint32_t x, uint32_t seed)
{
    x += seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return x;
}

__device__ __forceinline__ uint32_t owen_scramble(uint32_t x, uint32_t seed)
{
    return __builtin_bitreverse32(lk_permutation(__builtin_bitreverse32(x), seed));
}

// Sobol point `index` in dimension dim (0..3): the XOR of the direction numbers of the index's
// set bits, evaluated without a loop over the bits.
__device__ __forceinline__ uint32_t sobol(uint32_t index, uint32_t dim)
{
    if (dim == 0)  // van der Corput
        return __builtin_bitreverse32(index);
    if (dim == 1) {
        // Polynomial x + 1: the generator matrix is Pascal's triangle mod 2, so by Lucas' theorem
        // output bit 31 - u is the parity of the index bits at the positions b that contain u
        // (b & u == u): a superset sum over the five bits of the position.
        uint32_t z = index;
        z ^= (z >> 1) & 0x55555555u;
        z ^= (z >> 2) & 0x33333333u;
        z ^= (z >> 4) & 0x0F0F0F0Fu;
        z ^= (z >> 8) & 0x00FF00FFu;
        z ^= z >> 16;
        return __builtin_bitreverse32(z);
    }
    const uint32_t* t = c_sobol_bytes[dim - 2][0];
    return t[index & 255u] ^ t[256u + ((index >> 8) & 255u)] ^ t[512u + ((index >> 16) & 255u)] ^ t[768u + (index >> 24)];
}

struct Sampler {
    uint32_t index, seed, dimension, shuffled;

    __device__ __forceinline__ void start(uint32_t px, uint32_t py, uint32_t sample_index, uint32_t s)
    {
        seed = hash_combine(s, hash_combine(px, py));
        index = sample_index;
        dimension = 4;
        shuffled = 0;
    }
    __device__ __forceinline__ void pad()
    {
        dimension = 0;
        seed = hash_u32(seed);
        shuffled = owen_scramble(index, seed);
    }
    __device__ __forceinline__ float value(uint32_t d) const
    {
        return float(owen_scramble(sobol(shuffled, d), hash_combine(seed, d))) * 0x1p-32f;
    }
    __device__ __forceinline__ float s1()
    {
        if (dimension >= 4)
            pad();
        const float r = value(dimension);
        dimension += 1;
        return fminf(r, just_below_one());
    }
    __device__ __forceinline__ float2 s2()
    {
        if (dimension + 1 >= 4)
            pad();
        const float a = value(dimension), b = value(dimension + 1);
        dimension += 2;
        return make_float2(fminf(a, just_below_one()), fminf(b, just_below_one()));
    }
};

// --- sampling helpers (sample.hlsli:193-246) ----------------------------------------------

__device__ __forceinline__ float3 cosine_weighted_rand_dir(float2 u)
{
    const float r = sqrtf(u.x);
    return f3(r * cosf(2.0f * PI_F * u.y), r * sinf(2.0f * PI_F * u.y), sqrtf(1.0f - u.x));
}

__device__ __forceinline__ float3 random_barycentric(float2 u)
{
    const float beta = 1.0f - sqrtf(u.x);
    const float gamma = (1.0f - beta) * u.y;
    return f3(1.0f - beta - gamma, beta, gamma);
}

__device__ __forceinline__ float2 sample_concentric_disk(float2 u)
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
    return make_float2(r * cosf(phi), r * sinf(phi));
}

__device__ __forceinline__ float3 sample_visible_micronormal(float3 v, float2 alpha, float2 u)
{
    v = normalize3(f3(alpha.x * v.x, alpha.y * v.y, v.z));
    const float3 x = (v.z < 0.99999f) ? normalize3(cross3(f3(0, 0, 1), v)) : f3(1, 0, 0);
    const float3 y = cross3(v, x);
    const float a = 1.0f / (1.0f + v.z);
    const float r = sqrtf(u.x);
    const float phi = (u.y < a) ? u.y / a * PI_F : PI_F + (u.y - a) / (1.0f - a) * PI_F;
    const float p1 = r * cosf(phi);
    const float p2 = r * sinf(phi) * ((u.y < a) ? 1.0f : v.z);
    const float3 m = p1 * x + p2 * y + sqrtf(fmaxf(0.0f, 1.0f - p1 * p1 - p2 * p2)) * v;
    return normalize3(f3(alpha.x * m.x, alpha.y * m.y, fmaxf(1e-6f, m.z)));
}

// --- camera (sample.hlsli:258-280) --------------------------------------------------------

struct CameraBlock {
    float inv_view[16];
    float inv_proj[16];
    float near_clip, far_clip, focus_dist, lens_radius, exposure;
};

__device__ __forceinline__ void camera_ray(const CameraBlock& c, uint32_t px, uint32_t py, uint32_t w, uint32_t h,
                                           float2 u_pixel, float2 u_lens, float3& origin, float3& dir)
{
    const float nx = ((float(px) + u_pixel.x) / float(w)) * 2.0f - 1.0f;
    const float ny = ((float(py) + u_pixel.y) / float(h)) * 2.0f - 1.0f;
    const float* P = c.inv_proj;
    float3 rd = f3(P[0] * nx + P[1] * ny + P[2] * 1.0f + P[3] * 1.0f, P[4] * nx + P[5] * ny + P[6] * 1.0f + P[7] * 1.0f,
                   P[8] * nx + P[9] * ny + P[10] * 1.0f + P[11] * 1.0f);
    float3 ro = f3(0, 0, 0);
    if (c.lens_radius > 0.0f) {
        const float2 disk = sample_concentric_disk(u_lens);
        const float lx = c.lens_radius * disk.x, ly = c.lens_radius * disk.y;
        const float t_focus = c.focus_dist / -rd.z;
        const float3 fp = ro + t_focus * rd;
        ro = f3(lx, ly, 0.0f);
        rd = fp - ro;
    }
    const float* V = c.inv_view;
    origin = f3(V[0] * ro.x + V[1] * ro.y + V[2] * ro.z + V[3], V[4] * ro.x + V[5] * ro.y + V[6] * ro.z + V[7],
                V[8] * ro.x + V[9] * ro.y + V[10] * ro.z + V[11]);
    dir = normalize3(f3(V[0] * rd.x + V[1] * rd.y + V[2] * rd.z, V[4] * rd.x + V[5] * rd.y + V[6] * rd.z,
                        V[8] * rd.x + V[9] * rd.y + V[10] * rd.z));
}

// --- BSDF (pathtrace.ch.hlsl:22-157) ------------------------------------------------------

struct BrdfInput {
    float3 base_colour;
    float2 alpha;
    float metalness;
    float3 specular;
    float3 sheen;
};

__device__ __forceinline__ float sq(float x) { return x * x; }
__device__ __forceinline__ float lerpf(float a, float b, float t) { return a + t * (b - a); }
__device__ __forceinline__ float sin2_theta(float3 v) { return fmaxf(0.0f, 1.0f - sq(v.z)); }

__device__ __forceinline__ float lambda_fn(float3 m, float2 alpha)
{
    const float s2 = sin2_theta(m);
    const float tan2 = s2 / sq(m.z);
    if (isinf(tan2))
        return 0.0f;
    const float s = sqrtf(s2);
    const float cp = s == 0.0f ? 1.0f : fminf(fmaxf(m.x / s, -1.0f), 1.0f);
    const float sp = s == 0.0f ? 0.0f : fminf(fmaxf(m.y / s, -1.0f), 1.0f);
    const float a2 = sq(cp * alpha.x) + sq(sp * alpha.y);
    return (sqrtf(1.0f + a2 * tan2) - 1.0f) / 2.0f;
}

__device__ __forceinline__ float distribution(float3 m, float2 alpha)
{
    const float s2 = sin2_theta(m);
    const float c2 = sq(m.z);
    const float tan2 = s2 / c2;
    if (isinf(tan2))
        return 0.0f;
    const float cos4 = sq(c2);
    const float s = sqrtf(s2);
    const float cp = s == 0.0f ? 1.0f : fminf(fmaxf(m.x / s, -1.0f), 1.0f);
    const float sp = s == 0.0f ? 0.0f : fminf(fmaxf(m.y / s, -1.0f), 1.0f);
    const float e = tan2 * (sq(cp / alpha.x) + sq(sp / alpha.y));
    return 1.0f / (PI_F * alpha.x * alpha.y * cos4 * sq(1.0f + e));
}

__device__ __forceinline__ float diffuse_sample_probability(float metalness) { return lerpf(0.5f, 0.0f, metalness); }

// Clearcoat is computed by the shader but never added (ch.hlsl:93-114), so it is omitted.
__device__ __forceinline__ float3 evaluate_brdf(const BrdfInput& in, float3 i, float3 o, float3 m, float& pdf)
{
    const float3 f0 = f3(lerpf(in.specular.x * 0.08f, in.base_colour.x, in.metalness),
                         lerpf(in.specular.y * 0.08f, in.base_colour.y, in.metalness),
                         lerpf(in.specular.z * 0.08f, in.base_colour.z, in.metalness));
    const float cos_d = dot3(m, o);
    const float base = 1.0f - cos_d;
    const float f = base * base * base * base * base;
    const float lo = lambda_fn(o, in.alpha);
    const float g = 1.0f / (1.0f + lo + lambda_fn(i, in.alpha));
    const float d = distribution(m, in.alpha);
    const float3 fr = f3(lerpf(f0.x, 1.0f, f), lerpf(f0.y, 1.0f, f), lerpf(f0.z, 1.0f, f));
    const float denom = 4.0f * i.z * o.z;
    const float3 spec = f3(d * fr.x * g, d * fr.y * g, d * fr.z * g) / denom;
    const float3 sheen = in.sheen * f;
    const float3 diffuse = (in.base_colour + sheen) * (1.0f - in.metalness) / PI_F * (1.0f - f);
    const float diffuse_prob = diffuse_sample_probability(in.metalness);
    const float pdf_diffuse = i.z / PI_F;
    const float g1o = 1.0f / (1.0f + lo);
    const float pdf_m = (g1o / fabsf(o.z)) * d * fabsf(cos_d);
    const float pdf_gtr = pdf_m / (4.0f * fabsf(cos_d));
    pdf = pdf_diffuse * diffuse_prob + pdf_gtr * (1.0f - diffuse_prob);
    return spec + diffuse;
}

__device__ __forceinline__ float3 sample_direction(float3 o, float2 alpha, float metalness, Sampler& smp, float3& m,
                                                   float& w, float& pdf)
{
    const float diffuse_prob = diffuse_sample_probability(metalness);
    const bool sample_diffuse = smp.s1() < diffuse_prob;
    const float2 u = smp.s2();
    float3 i;
    if (sample_diffuse) {
        i = cosine_weighted_rand_dir(u);
        m = normalize3(o + i);
    } else {
        m = sample_visible_micronormal(o, alpha, u);
        i = 2.0f * dot3(o, m) * m - o;
    }
    float pdf_diffuse = i.z / PI_F;
    const float cos_d = dot3(m, o);
    const float g1o = 1.0f / (1.0f + lambda_fn(o, alpha));
    const float pdf_m = (g1o / fabsf(o.z)) * distribution(m, alpha) * fabsf(cos_d);
    float pdf_gtr = pdf_m / (4.0f * fabsf(cos_d));
    pdf_diffuse *= diffuse_prob;
    pdf_gtr *= (1.0f - diffuse_prob);
    pdf = sample_diffuse ? pdf_diffuse : pdf_gtr;
    w = pdf / (pdf_diffuse + pdf_gtr);
    return i;
}

// --- light sampling (lightsample.hlsli:7-143) ---------------------------------------------

__device__ __forceinline__ uint32_t sample_alias_table(const uint32_t* buf, uint32_t bin_count, uint32_t bin_size,
                                                       uint32_t table_offset, float& u, float& p)
{
    uint32_t i = min(bin_count - 1, uint32_t(u * float(bin_count)));
    uint32_t bo = table_offset + i * bin_size;
    u = fminf((u * float(bin_count)) - float(i), just_below_one());
    const float q = u32_as_float(buf[bo + 1]);
    if (u < q) {
        u = fminf(u / q, just_below_one());
    } else {
        i = buf[bo + 2];
        bo = table_offset + i * bin_size;
        u = fminf((u - q) / (1.0f - q), just_below_one());
    }
    p = u32_as_float(buf[bo + 0]);
    return bo + 3;
}

__device__ __forceinline__ float light_pdf(const uint32_t* light, uint32_t emitter_index, uint32_t prim)
{
    if (emitter_index == PT_NO_EMITTER)
        return 0.0f;
    uint32_t bo = emitter_index * PT_LIGHT_BIN_WORDS;
    float p = u32_as_float(light[bo + 0]);
    if (light[bo + 3] == 3u) {
        const uint32_t table_offset = light[bo + 4];
        bo = table_offset + prim * PT_EMITTER_BIN_WORDS;
        p *= u32_as_float(light[bo + 0]);
    }
    return p;
}

__device__ __forceinline__ float3 xform_point(const InstanceData& in, float3 p)
{
    return f3(in.m[0][0] * p.x + in.m[0][1] * p.y + in.m[0][2] * p.z + in.m[0][3],
              in.m[1][0] * p.x + in.m[1][1] * p.y + in.m[1][2] * p.z + in.m[1][3],
              in.m[2][0] * p.x + in.m[2][1] * p.y + in.m[2][2] * p.z + in.m[2][3]);
}

__device__ __forceinline__ float3 xform_vec(const InstanceData& in, float3 v)
{
    return f3(in.m[0][0] * v.x + in.m[0][1] * v.y + in.m[0][2] * v.z, in.m[1][0] * v.x + in.m[1][1] * v.y + in.m[1][2] * v.z,
              in.m[2][0] * v.x + in.m[2][1] * v.y + in.m[2][2] * v.z);
}

__device__ __forceinline__ float3 xform_normal(const InstanceData& in, float3 v)
{
    return f3(in.n[0][0] * v.x + in.n[0][1] * v.y + in.n[0][2] * v.z, in.n[1][0] * v.x + in.n[1][1] * v.y + in.n[1][2] * v.z,
              in.n[2][0] * v.x + in.n[2][1] * v.y + in.n[2][2] * v.z);
}

__device__ __forceinline__ float3 vertex_pos(const DeviceScene& s, uint32_t v)
{
    return xyz(s.vertices[size_t(v) * 2]);
}

__device__ __forceinline__ float3 emission_of(const MaterialRec& m)
{
    return f3(m.emission.x, m.emission.y, m.emission.z) * m.emission.w;
}

// --- textures (material.hlsli:23-106; sampler: bilinear, REPEAT, mip 0, src/graphics/sampler.cpp:8-22)

extern __constant__ float c_srgb[256];  // sRGB-encoded byte -> linear (VK_FORMAT_R8G8B8A8_SRGB fetch)

enum : uint32_t {
    MAP_BASE_COLOUR = 1u << 0,
    MAP_EMISSION = 1u << 1,
    MAP_ROUGHNESS = 1u << 2,
    MAP_METALNESS = 1u << 3,
    MAP_ROUGHNESS_METALNESS = 1u << 4,
    MAP_SPECULAR = 1u << 5,
    MAP_SHEEN = 1u << 7,
    MAP_NORMAL = 1u << 12,
};

__device__ __forceinline__ uint32_t map_bits(const MaterialRec& m) { return float_as_u32(m.rmnb.w); }

__device__ __forceinline__ float4 texel(const DeviceScene& s, const pt_texture_desc& d, uint32_t x, uint32_t y)
{
    const uint8_t* p = s.texels + d.offset + (size_t(y) * d.width + x) * d.channels;
    if (d.channels == 4) {
        const uint32_t w = *reinterpret_cast<const uint32_t*>(p);  // 4-byte aligned (texel offsets are)
        const uint32_t r = w & 255u, g = (w >> 8) & 255u, b = (w >> 16) & 255u, a = w >> 24;
        const float k = 1.0f / 255.0f;
        if (d.srgb)
            return make_float4(c_srgb[r], c_srgb[g], c_srgb[b], float(a) * k);
        return make_float4(float(r) * k, float(g) * k, float(b) * k, float(a) * k);
    }
    if (d.channels == 2)
        return make_float4(float(p[0]) / 255.0f, float(p[1]) / 255.0f, 0.0f, 1.0f);
    return make_float4(float(p[0]) / 255.0f, 0.0f, 0.0f, 1.0f);
}

// Texture2D.SampleLevel(sampler, uv, 0): texel centres at (i + 0.5) / size, four texels weighted
// bilinearly, coordinates wrapped (REPEAT).
__device__ __forceinline__ float4 sample_texture(const DeviceScene& s, uint32_t t, float2 uv)
{
    const pt_texture_desc d = s.tex[t];
    const float u = uv.x - floorf(uv.x), v = uv.y - floorf(uv.y);
    const float x = u * float(d.width) - 0.5f, y = v * float(d.height) - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float ax = x - fx, ay = y - fy;
    int ix = int(fx), iy = int(fy);
    const uint32_t x0 = ix < 0 ? d.width - 1 : min(uint32_t(ix), d.width - 1);
    const uint32_t y0 = iy < 0 ? d.height - 1 : min(uint32_t(iy), d.height - 1);
    const uint32_t x1 = x0 + 1 >= d.width ? 0 : x0 + 1;
    const uint32_t y1 = y0 + 1 >= d.height ? 0 : y0 + 1;
    const float4 t00 = texel(s, d, x0, y0), t10 = texel(s, d, x1, y0);
    const float4 t01 = texel(s, d, x0, y1), t11 = texel(s, d, x1, y1);
    const float w00 = (1.0f - ax) * (1.0f - ay), w10 = ax * (1.0f - ay), w01 = (1.0f - ax) * ay, w11 = ax * ay;
    return make_float4(w00 * t00.x + w10 * t10.x + w01 * t01.x + w11 * t11.x,
                       w00 * t00.y + w10 * t10.y + w01 * t01.y + w11 * t11.y,
                       w00 * t00.z + w10 * t10.z + w01 * t01.z + w11 * t11.z,
                       w00 * t00.w + w10 * t10.w + w01 * t01.w + w11 * t11.w);
}

// get_emission / get_base_colour (material.hlsli:23-41)
__device__ __forceinline__ float3 material_emission(const DeviceScene& s, const MaterialRec& m, float2 uv)
{
    if (map_bits(m) & MAP_EMISSION)
        return xyz(sample_texture(s, float_as_u32(m.emission.x), uv)) * m.emission.w;
    return emission_of(m);
}

__device__ __forceinline__ float4 material_base_colour(const DeviceScene& s, const MaterialRec& m, float2 uv)
{
    if (map_bits(m) & MAP_BASE_COLOUR)
        return sample_texture(s, float_as_u32(m.base_colour.x), uv);
    return m.base_colour;
}

// get_roughness_metalness_normal (material.hlsli:43-70); normal is the tangent-space map normal
// or (0, 0, 1).
__device__ __forceinline__ void material_rmn(const DeviceScene& s, const MaterialRec& m, float2 uv, float& roughness,
                                             float& metalness, float3& normal)
{
    const uint32_t b = map_bits(m);
    if (b & MAP_ROUGHNESS_METALNESS) {
        const float4 t = sample_texture(s, float_as_u32(m.rmnb.x), uv);
        roughness = t.y;
        metalness = t.z;
    } else {
        roughness = (b & MAP_ROUGHNESS) ? sample_texture(s, float_as_u32(m.rmnb.x), uv).x : m.rmnb.x;
        metalness = (b & MAP_METALNESS) ? sample_texture(s, float_as_u32(m.rmnb.y), uv).x : m.rmnb.y;
    }
    if (b & MAP_NORMAL) {
        const float4 t = sample_texture(s, float_as_u32(m.rmnb.z), uv);
        normal = normalize3(f3(t.x * 2.0f - 1.0f, t.y * 2.0f - 1.0f, t.z * 2.0f - 1.0f));
    } else {
        normal = f3(0.0f, 0.0f, 1.0f);
    }
}

__device__ __forceinline__ float3 material_specular(const DeviceScene& s, const MaterialRec& m, float2 uv)
{
    return (map_bits(m) & MAP_SPECULAR) ? xyz(sample_texture(s, float_as_u32(m.specular.x), uv)) : xyz(m.specular);
}

__device__ __forceinline__ float3 material_sheen(const DeviceScene& s, const MaterialRec& m, float2 uv)
{
    return (map_bits(m) & MAP_SHEEN) ? xyz(sample_texture(s, float_as_u32(m.sheen.x), uv)) : xyz(m.sheen);
}

// Orthonormal basis around a unit vector (Duff et al. 2017, util.hlsli:57-64).
__device__ __forceinline__ void onb(float3 n, float3& x, float3& y)
{
    const float sg = n.z < 0.0f ? -1.0f : 1.0f;
    const float a = -1.0f / (sg + n.z);
    const float b = n.x * n.y * a;
    x = f3(1.0f + sg * n.x * n.x * a, sg * b, -sg * n.x);
    y = f3(b, sg + n.y * n.y * a, -n.y);
}

// Texture coordinates of a triangle point (geometry.hlsli:39-48).
__device__ __forceinline__ float2 vertex_uv(const DeviceScene& s, uint32_t v)
{
    const float4 t = s.vertices[size_t(v) * 2 + 1];
    return make_float2(t.z, t.w);
}

// sample_light (lightsample.hlsli:7-143) split at its dependent loads, so that the shade
// stage keeps its own loads in flight between them: alias_pick / alias_follow are the two
// halves of sample_alias_table (the bin u falls in, then that bin or its alias, u rescaled
// each time), load_alias_bin reads one bin, light_sample_at evaluates the chosen light at the
// shading point from what the walk read.
template <int PAYLOAD>
struct AliasBin {
    float p, q;
    uint32_t alias;
    uint32_t w[PAYLOAD];  // the bin's payload words (light: type and 6 words; emitter: 3 vertices)
};

template <int PAYLOAD>
__device__ __forceinline__ AliasBin<PAYLOAD> load_alias_bin(const uint32_t* b)
{
    AliasBin<PAYLOAD> r;
    r.p = u32_as_float(b[0]);
    r.q = u32_as_float(b[1]);
    r.alias = b[2];
    for (int k = 0; k < PAYLOAD; k++)
        r.w[k] = b[3 + k];
    return r;
}

__device__ __forceinline__ uint32_t alias_pick(uint32_t bin_count, float& u)
{
    const uint32_t i = min(bin_count - 1, uint32_t(u * float(bin_count)));
    u = fminf((u * float(bin_count)) - float(i), just_below_one());
    return i;
}

// True when the alias is taken.
__device__ __forceinline__ bool alias_follow(float q, float& u)
{
    if (u < q) {
        u = fminf(u / q, just_below_one());
        return false;
    }
    u = fminf((u - q) / (1.0f - q), just_below_one());
    return true;
}

__device__ __forceinline__ float3 xform_point_rows(float4 r0, float4 r1, float4 r2, float3 p)
{
    return f3(r0.x * p.x + r0.y * p.y + r0.z * p.z + r0.w, r1.x * p.x + r1.y * p.y + r1.z * p.z + r1.w,
              r2.x * p.x + r2.y * p.y + r2.z * p.z + r2.w);
}

__device__ __forceinline__ float3 xform_rows(float4 r0, float4 r1, float4 r2, float3 v)
{
    return f3(r0.x * v.x + r0.y * v.y + r0.z * v.z, r1.x * v.x + r1.y * v.y + r1.z * v.z,
              r2.x * v.x + r2.y * v.y + r2.z * v.z);
}

// lb: the chosen light bin; for an emitting object (type 3) also eb: the chosen triangle bin,
// eh / m0-m2: the emitter instance's indices and object->world rows, v0-v2: the triangle's
// vertices, emission / rmnb: the emitter material's emission and map bits.
__device__ __forceinline__ float3 light_sample_at(const DeviceScene& s, const AliasBin<7>& lb, const AliasBin<3>& eb,
                                                  uint4 eh, float4 m0, float4 m1, float4 m2, float4 v0, float4 v1,
                                                  float4 v2, float4 emission, float4 rmnb, float2 u12, float3 pt,
                                                  float3& light_dir, float& light_dist, float& pdf, bool& is_delta)
{
    pdf = lb.p;
    const uint32_t type = lb.w[0];
    if (type == 3u) {
        pdf *= eb.p;
        const float3 pa = xform_point_rows(m0, m1, m2, xyz(v0));
        const float3 pb = xform_point_rows(m0, m1, m2, xyz(v1));
        const float3 pc = xform_point_rows(m0, m1, m2, xyz(v2));
        const float3 bary = random_barycentric(u12);
        const float3 on_tri = bary.x * pa + bary.y * pb + bary.z * pc;
        light_dir = on_tri - pt;
        light_dist = 1.0f;
        const float3 tn = cross3(pb - pa, pc - pa);
        const float area = sqrtf(dot3(tn, tn)) * 0.5f;
        const float cosine = fabsf(dot3(normalize3(light_dir), normalize3(tn)));
        const float dist2 = dot3(light_dir, light_dir);
        pdf *= dist2 / (area * cosine);
        is_delta = false;
        if (!isfinite(pdf))
            return f3(0, 0, 0);
        if (!(float_as_u32(rmnb.w) & MAP_EMISSION))
            return f3(emission.x, emission.y, emission.z) * emission.w;
        const float2 ua = vertex_uv(s, eh.x + eb.w[0]), ub = vertex_uv(s, eh.x + eb.w[1]),
                     uc = vertex_uv(s, eh.x + eb.w[2]);
        const float2 uv = make_float2(bary.x * ua.x + bary.y * ub.x + bary.z * uc.x,
                                      bary.x * ua.y + bary.y * ub.y + bary.z * uc.y);
        return material_emission(s, s.materials[eh.w], uv);
    } else if (type == 1u) {
        const float3 lp = f3(u32_as_float(lb.w[1]), u32_as_float(lb.w[2]), u32_as_float(lb.w[3]));
        light_dir = lp - pt;
        light_dist = 1.0f;
        is_delta = true;
        return (1.0f / dot3(light_dir, light_dir)) *
               f3(u32_as_float(lb.w[4]), u32_as_float(lb.w[5]), u32_as_float(lb.w[6]));
    } else if (type == 2u) {
        light_dir = f3(u32_as_float(lb.w[1]), u32_as_float(lb.w[2]), u32_as_float(lb.w[3]));
        light_dist = INFINITY;
        is_delta = true;
        return f3(u32_as_float(lb.w[4]), u32_as_float(lb.w[5]), u32_as_float(lb.w[6]));
    }
    light_dist = 0.0f;
    is_delta = true;
    return f3(0, 0, 0);
}

__device__ __forceinline__ float correct_component(float x) { return (x < 0.0f || !isfinite(x)) ? 0.0f : x; }

}  // namespace ptk
