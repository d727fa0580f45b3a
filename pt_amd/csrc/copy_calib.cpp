// Plain device copy of a known size, for calibrating rocprofv3 FETCH_SIZE / WRITE_SIZE
// counters against a kernel whose HBM traffic is exactly known (1 GiB read + 1 GiB write
// per launch). Built as pt_amd/bin/copy_calib; not part of the library.
#include <hip/hip_runtime.h>

#include <cstdio>

__global__ void copy_f4(const float4* __restrict__ src, float4* __restrict__ dst, size_t n)
{
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

__global__ void fill_f4(float4* dst, size_t n)
{
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        dst[i] = make_float4(1.0f, 2.0f, 3.0f, 4.0f);
}

int main()
{
    const size_t bytes = size_t(1) << 30;
    const size_t n = bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        fprintf(stderr, "copy_calib: allocation failed\n");
        return 1;
    }
    hipLaunchKernelGGL(fill_f4, dim3(4096), dim3(256), 0, 0, a, n);
    for (int k = 0; k < 3; k++)
        hipLaunchKernelGGL(copy_f4, dim3(8192), dim3(256), 0, 0, a, b, n);
    if (hipDeviceSynchronize() != hipSuccess) {
        fprintf(stderr, "copy_calib: kernel failed\n");
        return 1;
    }
    printf("copy_calib: 3 copies of %zu bytes\n", bytes);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
