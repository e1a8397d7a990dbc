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

// Variants: U float4 per thread in a fixed (non-grid-stride) layout, with
// ordinary or non-temporal loads/stores.
typedef float v4f __attribute__((ext_vector_type(4)));

template <int U, bool NT>
__global__ void __launch_bounds__(256) hbm_copy_wide(const v4f *__restrict__ src, v4f *__restrict__ dst,
                                                     std::int64_t n)
{
    const std::int64_t base = ((std::int64_t)blockIdx.x * 256 * U) + threadIdx.x;
    v4f v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const std::int64_t i = base + (std::int64_t)u * 256;
        if (i < n) v[u] = NT ? __builtin_nontemporal_load(src + i) : src[i];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const std::int64_t i = base + (std::int64_t)u * 256;
        if (i < n) {
            if (NT) __builtin_nontemporal_store(v[u], dst + i);
            else dst[i] = v[u];
        }
    }
}

extern "C" {

// variant: 0 grid-stride (hbm_copy_kernel), 1..: U = 1, 2, 4, 8 plain, then
// the same four with non-temporal loads/stores (5..8).
double hbm_copy_variant_gbs(std::uint64_t bytes, int iters, int variant)
{
    void *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess) return -1.0;
    if (hipMalloc(&b, bytes) != hipSuccess) { (void)hipFree(a); return -1.0; }
    (void)hipMemset(a, 1, bytes);
    (void)hipMemset(b, 0, bytes);
    const std::int64_t n = (std::int64_t)(bytes / 16);
    const int us[4] = {1, 2, 4, 8};
    const int U = variant >= 1 ? us[(variant - 1) % 4] : 1;
    const bool nt = variant >= 5;
    const unsigned blocks = (unsigned)((n + 256 * U - 1) / (256 * U));
    auto launch = [&]() {
        const v4f *s = (const v4f *)a;
        v4f *d = (v4f *)b;
        if (variant == 0) {
            hipLaunchKernelGGL(hbm_copy_kernel, dim3(4096), dim3(256), 0, 0, (const float4 *)a, (float4 *)b, n);
            return;
        }
        switch (U * (nt ? -1 : 1)) {
        case 1: hipLaunchKernelGGL((hbm_copy_wide<1, false>), dim3(blocks), dim3(256), 0, 0, s, d, n); break;
        case 2: hipLaunchKernelGGL((hbm_copy_wide<2, false>), dim3(blocks), dim3(256), 0, 0, s, d, n); break;
        case 4: hipLaunchKernelGGL((hbm_copy_wide<4, false>), dim3(blocks), dim3(256), 0, 0, s, d, n); break;
        case 8: hipLaunchKernelGGL((hbm_copy_wide<8, false>), dim3(blocks), dim3(256), 0, 0, s, d, n); break;
        case -1: hipLaunchKernelGGL((hbm_copy_wide<1, true>), dim3(blocks), dim3(256), 0, 0, s, d, n); break;
        case -2: hipLaunchKernelGGL((hbm_copy_wide<2, true>), dim3(blocks), dim3(256), 0, 0, s, d, n); break;
        case -4: hipLaunchKernelGGL((hbm_copy_wide<4, true>), dim3(blocks), dim3(256), 0, 0, s, d, n); break;
        default: hipLaunchKernelGGL((hbm_copy_wide<8, true>), dim3(blocks), dim3(256), 0, 0, s, d, n); break;
        }
    };
    hipEvent_t t0, t1;
    (void)hipEventCreate(&t0);
    (void)hipEventCreate(&t1);
    launch();
    (void)hipEventRecord(t0, 0);
    for (int q = 0; q < iters; ++q) launch();
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
