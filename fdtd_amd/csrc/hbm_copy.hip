// hbm_copy.hip -- a plain streaming copy kernel of known size.
//
// Measurement aid only (not part of the engine's ABI): its HIP-event rate is
// the practical HBM ceiling for a read+write stream on this device, and the
// FETCH_SIZE / WRITE_SIZE counters it produces under rocprofv3 calibrate the
// engine's counter-derived HBM traffic (bytes the counter reports per byte
// actually moved).
#include <hip/hip_runtime.h>
#include <cstdint>

__global__ void __launch_bounds__(256) hbm_copy_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                       std::int64_t n)
{
    std::int64_t i = (std::int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const std::int64_t stride = (std::int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

extern "C" {

// Copy `bytes` (multiple of 16) `iters` times; returns the mean GB/s counted
// as read + write bytes, or a negative value on a HIP error.
double hbm_copy_gbs(std::uint64_t bytes, int iters)
{
    void *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess) return -1.0;
    if (hipMalloc(&b, bytes) != hipSuccess) { (void)hipFree(a); return -1.0; }
    (void)hipMemset(a, 1, bytes);
    (void)hipMemset(b, 0, bytes);
    const std::int64_t n = (std::int64_t)(bytes / 16);
    const int blocks = 256 * 16;
    hipEvent_t t0, t1;
    (void)hipEventCreate(&t0);
    (void)hipEventCreate(&t1);
    hipLaunchKernelGGL(hbm_copy_kernel, dim3(blocks), dim3(256), 0, 0, (const float4 *)a, (float4 *)b, n);
    (void)hipEventRecord(t0, 0);
    for (int q = 0; q < iters; ++q)
        hipLaunchKernelGGL(hbm_copy_kernel, dim3(blocks), dim3(256), 0, 0, (const float4 *)a, (float4 *)b, n);
    (void)hipEventRecord(t1, 0);
    (void)hipEventSynchronize(t1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, t0, t1);
    hipError_t e = hipGetLastError();
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return -2.0;
    return 2.0 * (double)bytes * iters / (ms * 1e-3) / 1e9;
}

}  // extern "C"
