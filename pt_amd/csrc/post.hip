// Output post-processing: the PNG path of the headless dispatcher. PathTraceDispatcher::
// put_result (src/pathtrace/ptdispatch.cpp:155-204) runs the ACES tone mapper over the
// accumulation (ToneMapper constructed with src_is_srgb = false, dst_is_srgb = true,
// ptdispatch.cpp:29; shader shaders/compute/tonemap.cs.hlsl) and write_png
// (src/io/imagewrite.cpp:39-62) converts each channel with (uchar)clamp(x * 255, 0, 255).
// Here both happen in one kernel that writes RGBA8; the accumulation itself is left as it is
// (the reference tone-maps it in place, after which it only writes the file).
#include <hip/hip_runtime.h>

#include "device.h"

namespace ptk {

namespace {

// tonemap.cs.hlsl:30-32, with the shader's threshold 0.00031308 (the sRGB standard's is
// 0.0031308; kept as the shader has it)
__device__ __forceinline__ float tm_to_srgb(float x)
{
    return x <= 0.00031308f ? 12.92f * x : 1.055f * powf(x, 1.0f / 2.4f) - 0.055f;
}

// tonemap.cs.hlsl:48-53
__device__ __forceinline__ float tm_fit(float v)
{
    const float a = v * (v + 0.0245786f) - 0.000090537f;
    const float b = v * (0.983729f * v + 0.4329510f) + 0.238081f;
    return a / b;
}

__device__ __forceinline__ float tm_saturate(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

__device__ __forceinline__ uint32_t tm_byte(float x)
{
    return uint32_t(fminf(fmaxf(x * 255.0f, 0.0f), 255.0f));  // float_to_byte, imagewrite.cpp:39-42
}

// aces() of tonemap.cs.hlsl:55-66 for src_is_srgb = false, dst_is_srgb = true, then the 8-bit
// conversion; alpha passes through the tone mapper unchanged.
__global__ void k_tonemap_rgba8(const float4* accum, uint32_t width, uint32_t height, uint32_t* out)
{
    const uint32_t n = width * height;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4 c = accum[i];
        const float s0 = tm_to_srgb(c.x), s1 = tm_to_srgb(c.y), s2 = tm_to_srgb(c.z);
        // srgb_to_rrt_mat (tonemap.cs.hlsl:17-22), mul(matrix, column)
        const float r0 = 0.59719f * s0 + 0.35458f * s1 + 0.04823f * s2;
        const float r1 = 0.07600f * s0 + 0.90834f * s1 + 0.01566f * s2;
        const float r2 = 0.02840f * s0 + 0.13383f * s1 + 0.83777f * s2;
        const float f0 = tm_fit(r0), f1 = tm_fit(r1), f2 = tm_fit(r2);
        // odt_to_srgb_mat (tonemap.cs.hlsl:24-29)
        const float o0 = 1.60475f * f0 - 0.53108f * f1 - 0.07367f * f2;
        const float o1 = -0.10208f * f0 + 1.10813f * f1 - 0.00605f * f2;
        const float o2 = -0.00327f * f0 - 0.07276f * f1 + 1.07602f * f2;
        out[i] = tm_byte(tm_saturate(o0)) | (tm_byte(tm_saturate(o1)) << 8) | (tm_byte(tm_saturate(o2)) << 16) |
                 (tm_byte(c.w) << 24);
    }
}

}  // namespace

hipError_t tonemap_rgba8(const float4* accum, uint32_t width, uint32_t height, uint32_t* d_out, hipStream_t st)
{
    const uint64_t n = uint64_t(width) * height;
    const unsigned blocks = unsigned(std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 65535u * 4)));
    hipLaunchKernelGGL(k_tonemap_rgba8, dim3(blocks), dim3(256), 0, st, accum, width, height, d_out);
    return hipGetLastError();
}

}  // namespace ptk
