// Device build of the light sampler's alias tables (src/aliastable.h:21-67,
// src/scene/lightsample.cpp:5-178): the table over the lights (emitting objects weighted by
// (r + g + b) * strength of their emission - for an emission map the Monte Carlo mean of
// calculate_object_intensity, lightsample.cpp:41-87 - point and directional lights by
// r + g + b) and one table per emissive mesh over its triangles (weighted by the object-space
// cross-product length). Weights are computed in parallel; each table is then built by one
// thread with the same arithmetic and the same LIFO under/overfull lists as the host build, so
// the words are bit-identical to the ones the host packs (the bins' payload words are kept).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "device.h"

namespace ptk {

namespace {

constexpr uint32_t LIGHT_TYPE_EMITTER = 3;

struct TableJob {
    uint32_t word_offset;  // first word of the table
    uint32_t n;            // bins
    uint32_t stride;       // words per bin
    uint32_t weight_offset;  // first weight in the weight buffer
};

// The random stream calculate_object_intensity draws from: Xshiro128 (src/rng.cpp:37-60)
// seeded by four Splitmix32(0xbd7a0d48) outputs, which the reference's g++ build passes in
// reverse order (its argument list is evaluated right to left), uniform floats result / 2^32.
struct IntensityRng {
    uint32_t st[4];
    __device__ IntensityRng()
    {
        uint32_t sm = 0xbd7a0d48u, d[4];
        for (int k = 0; k < 4; k++) {
            uint32_t z = (sm += 0x9e3779b9u);
            z = (z ^ (z >> 16)) * 0x85ebca6bu;
            z = (z ^ (z >> 13)) * 0xc2b2ae35u;
            d[k] = z ^ (z >> 16);
        }
        st[0] = d[3], st[1] = d[2], st[2] = d[1], st[3] = d[0];
    }
    __device__ static uint32_t rotl(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }
    __device__ float next()
    {
        const uint32_t result = rotl(st[0] + st[3], 7) + st[0];
        const uint32_t t = st[1] << 9;
        st[2] ^= st[0];
        st[3] ^= st[1];
        st[1] ^= st[2];
        st[0] ^= st[3];
        st[2] ^= t;
        st[3] = rotl(st[3], 11);
        return float(result) / 4294967296.0f;
    }
};

// calculate_object_intensity of a mesh with an emission map (lightsample.cpp:41-87): the mean
// of r + g + b of the texel nearest to the interpolated UV (raw 8-bit channels / 255, the
// fraction of the UV rounded to a texel, negative fractions clamped to the last texel) over 20
// uniform points per triangle, in triangle order, times the strength.
__device__ float textured_intensity(const float4* vertices, const uint32_t* indices, const pt_mesh_desc& m,
                                    const pt_texture_desc& tx, const uint8_t* texels, float strength)
{
#pragma clang fp contract(off)
    IntensityRng rng;
    const uint32_t spt = 20;
    double total = 0.0;
    for (uint32_t t = 0; t < m.tri_count; t++) {
        const uint32_t* ix = indices + m.index_start + 3 * size_t(t);
        const float4 a = vertices[size_t(m.vertex_start + ix[0]) * 2 + 1];
        const float4 b = vertices[size_t(m.vertex_start + ix[1]) * 2 + 1];
        const float4 c = vertices[size_t(m.vertex_start + ix[2]) * 2 + 1];
        double sum = 0.0;
        for (uint32_t j = 0; j < spt; j++) {
            const float beta = 1.0f - sqrtf(rng.next());
            const float gamma = (1.0f - beta) * rng.next();
            const float alpha = 1.0f - beta - gamma;
            const float u = alpha * a.z + beta * b.z + gamma * c.z;
            const float v = alpha * a.w + beta * b.w + gamma * c.w;
            float ip;
            const float fx = roundf(modff(u, &ip) * float(tx.width));
            const float fy = roundf(modff(v, &ip) * float(tx.height));
            const uint32_t x = fx < 0.0f ? tx.width - 1 : min(uint32_t(fx), tx.width - 1);
            const uint32_t y = fy < 0.0f ? tx.height - 1 : min(uint32_t(fy), tx.height - 1);
            const uint8_t* p = texels + tx.offset + (size_t(y) * tx.width + x) * tx.channels;
            const float c0 = p[0] / 255.0f;
            const float c1 = tx.channels >= 2 ? p[1] / 255.0f : 0.0f;
            const float c2 = tx.channels == 4 ? p[2] / 255.0f : 0.0f;
            sum += double(c0 + c1 + c2);
        }
        total += sum;
    }
    return float(total * double(strength) / double(size_t(m.tri_count) * spt));
}

__global__ void k_light_weights(const uint32_t* words, uint32_t light_count, const InstanceData* instances,
                                const MaterialRec* materials, const uint32_t* instance_mesh,
                                const pt_mesh_desc* meshes, const float4* vertices, const uint32_t* indices,
                                const pt_texture_desc* textures, uint32_t texture_count, const uint8_t* texels,
                                float* w)
{
#pragma clang fp contract(off)
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= light_count)
        return;
    const uint32_t* b = words + size_t(i) * PT_LIGHT_BIN_WORDS;
    if (b[3] == LIGHT_TYPE_EMITTER) {
        const MaterialRec& mat = materials[instances[b[6]].material_index];
        const float4 e = mat.emission;
        const uint32_t tex = __float_as_uint(e.x);
        if ((__float_as_uint(mat.rmnb.w) & MAP_EMISSION) && tex < texture_count)
            w[i] = textured_intensity(vertices, indices, meshes[instance_mesh[b[6]]], textures[tex], texels, e.w);
        else
            w[i] = (e.x + e.y + e.z) * e.w;
    } else {
        w[i] = __uint_as_float(b[7]) + __uint_as_float(b[8]) + __uint_as_float(b[9]);
    }
}

// Triangle weights and vertex-index payloads of every emitter table; one thread per triangle.
__global__ void k_tri_weights(const float4* vertices, const uint32_t* indices, const pt_mesh_desc* meshes,
                              const uint32_t* table_mesh, const TableJob* jobs, uint32_t ntables, uint32_t total,
                              uint32_t* words, float* w)
{
#pragma clang fp contract(off)
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total)
        return;
    // emitter tables are jobs 1..ntables, their weights contiguous from jobs[1].weight_offset
    uint32_t lo = 1, hi = ntables + 1;
    const uint32_t base = jobs[1].weight_offset;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].weight_offset - base <= g)
            lo = mid;
        else
            hi = mid;
    }
    const TableJob j = jobs[lo];
    const uint32_t t = g - (j.weight_offset - base);
    const pt_mesh_desc m = meshes[table_mesh[lo - 1]];
    const uint32_t i0 = indices[m.index_start + 3 * t], i1 = indices[m.index_start + 3 * t + 1],
                   i2 = indices[m.index_start + 3 * t + 2];
    const float4 a = vertices[size_t(m.vertex_start + i0) * 2], b = vertices[size_t(m.vertex_start + i1) * 2],
                 c = vertices[size_t(m.vertex_start + i2) * 2];
    const float abx = b.x - a.x, aby = b.y - a.y, abz = b.z - a.z;
    const float acx = c.x - a.x, acy = c.y - a.y, acz = c.z - a.z;
    const float cx = aby * acz - abz * acy, cy = abz * acx - abx * acz, cz = abx * acy - aby * acx;
    w[j.weight_offset + t] = sqrtf(cx * cx + cy * cy + cz * cz);
    uint32_t* bin = words + j.word_offset + size_t(t) * j.stride;
    bin[3] = i0;
    bin[4] = i1;
    bin[5] = i2;
}

// Vose's method, one thread per table: p = w / sum (the sum in double, in bin order),
// u = p * n, LIFO lists of under- and overfull bins, leftovers q = 1.
__global__ void k_vose(const TableJob* jobs, uint32_t njobs, const float* w, uint32_t* words, uint32_t* scratch)
{
#pragma clang fp contract(off)
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= njobs)
        return;
    const TableJob j = jobs[k];
    const float* wt = w + j.weight_offset;
    uint32_t* bins = words + j.word_offset;
    uint32_t* under = scratch + 2 * size_t(j.weight_offset);
    uint32_t* over = under + j.n;
    double sum = 0.0;
    for (uint32_t i = 0; i < j.n; i++)
        sum += double(wt[i]);
    const double inv_sum = 1.0 / sum;
    uint32_t nu = 0, no = 0;
    for (uint32_t i = 0; i < j.n; i++) {
        const float p = float(double(wt[i]) * inv_sum);
        const float u = p * float(j.n);
        uint32_t* b = bins + size_t(i) * j.stride;
        b[0] = __float_as_uint(p);
        b[1] = __float_as_uint(u);
        b[2] = i;
        if (u < 1.0f)
            under[nu++] = i;
        else
            over[no++] = i;
    }
    while (nu > 0 && no > 0) {
        const uint32_t s = under[nu - 1], l = over[no - 1];
        uint32_t* bs = bins + size_t(s) * j.stride;
        uint32_t* bl = bins + size_t(l) * j.stride;
        bs[2] = l;
        bl[1] = __float_as_uint(__uint_as_float(bl[1]) + (__uint_as_float(bs[1]) - 1.0f));
        nu--;
        if (__uint_as_float(bl[1]) < 1.0f) {
            under[nu++] = l;
            no--;
        }
    }
    for (uint32_t i = 0; i < nu; i++)
        bins[size_t(under[i]) * j.stride + 1] = __float_as_uint(1.0f);
    for (uint32_t i = 0; i < no; i++)
        bins[size_t(over[i]) * j.stride + 1] = __float_as_uint(1.0f);
}

// Clears p, q, alias (and the triangle payloads) so that the build provably writes them.
__global__ void k_clear_bins(const TableJob* jobs, uint32_t njobs, uint32_t* words)
{
    const uint32_t k = blockIdx.y;
    if (k >= njobs)
        return;
    const TableJob j = jobs[k];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < j.n; i += gridDim.x * blockDim.x) {
        uint32_t* b = words + j.word_offset + size_t(i) * j.stride;
        b[0] = b[1] = b[2] = 0u;
        if (k > 0)
            b[3] = b[4] = b[5] = 0u;
    }
}

}  // namespace

hipError_t build_light_tables(Ctx& c)
{
    if (c.light_count == 0)
        return hipSuccess;
    hipStream_t st = c.stream;
    // job 0: the light table; jobs 1..: the emitter tables in order
    std::vector<TableJob> jobs;
    std::vector<uint32_t> table_mesh;
    jobs.push_back(TableJob{0, c.light_count, PT_LIGHT_BIN_WORDS, 0});
    uint32_t wofs = c.light_count, tri_total = 0;
    for (const EmitterTable& t : c.emitter_tables) {
        jobs.push_back(TableJob{t.word_offset, t.tris, PT_EMITTER_BIN_WORDS, wofs});
        table_mesh.push_back(t.mesh);
        wofs += t.tris;
        tri_total += t.tris;
    }
    const uint32_t njobs = uint32_t(jobs.size());
    TableJob* d_jobs = nullptr;
    uint32_t* d_table_mesh = nullptr;
    float* d_w = nullptr;
    uint32_t* d_scratch = nullptr;
    hipError_t e = hipMalloc(&d_jobs, njobs * sizeof(TableJob));
    if (e == hipSuccess)
        e = hipMalloc(&d_table_mesh, std::max<size_t>(table_mesh.size(), 1) * 4);
    if (e == hipSuccess)
        e = hipMalloc(&d_w, size_t(wofs) * 4);
    if (e == hipSuccess)
        e = hipMalloc(&d_scratch, size_t(wofs) * 8);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_jobs, jobs.data(), njobs * sizeof(TableJob), hipMemcpyHostToDevice, st);
    if (e == hipSuccess && !table_mesh.empty())
        e = hipMemcpyAsync(d_table_mesh, table_mesh.data(), table_mesh.size() * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_clear_bins, dim3(64, njobs), dim3(256), 0, st, d_jobs, njobs, c.d_light);
        hipLaunchKernelGGL(k_light_weights, dim3((c.light_count + 63) / 64), dim3(64), 0, st, c.d_light,
                           c.light_count, c.d_instances, c.d_materials, c.d_instance_mesh, c.d_meshes,
                           c.d_vertices, c.d_indices, c.d_tex, c.scene.tex_count, c.d_texels, d_w);
        if (tri_total)
            hipLaunchKernelGGL(k_tri_weights, dim3((tri_total + 255) / 256), dim3(256), 0, st, c.d_vertices,
                               c.d_indices, c.d_meshes, d_table_mesh, d_jobs, njobs - 1, tri_total, c.d_light, d_w);
        hipLaunchKernelGGL(k_vose, dim3((njobs + 63) / 64), dim3(64), 0, st, d_jobs, njobs, d_w, c.d_light, d_scratch);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    void* ptrs[] = {d_jobs, d_table_mesh, d_w, d_scratch};
    for (void* p : ptrs)
        if (p)
            (void)hipFree(p);
    return e;
}

}  // namespace ptk
