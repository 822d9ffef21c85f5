// Plain device-to-device copy of a known byte count: the calibration kernel for HBM
// byte counters (FETCH_SIZE / WRITE_SIZE read under a profiler for this kernel give
// the counter's bytes per algorithmic byte). Built as its own small library, outside
// the C ABI of physkit_amd.h.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace
{
__global__ void __launch_bounds__(256) k_copy_check(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n)
{
    size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

__global__ void __launch_bounds__(256) k_copy_fill(uint4 *__restrict__ dst, size_t n)
{
    size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = make_uint4(static_cast<uint32_t>(i), 1u, 2u, 3u);
}
} // namespace

extern "C" int pkc_copy_check(size_t bytes, int reps, double *ms_per_copy)
{
    if (bytes == 0 || bytes % 16 != 0 || reps < 1) return 1;
    size_t n = bytes / 16;
    uint4 *a = nullptr, *b = nullptr;
    if (hipMalloc(reinterpret_cast<void **>(&a), bytes) != hipSuccess) return 2;
    if (hipMalloc(reinterpret_cast<void **>(&b), bytes) != hipSuccess)
    {
        (void)hipFree(a);
        return 2;
    }
    const int grid = 256 * 8 * 4;
    k_copy_fill<<<grid, 256>>>(a, n);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps; ++r) k_copy_check<<<grid, 256>>>(a, b, n);
    (void)hipEventRecord(e1, nullptr);
    int rc = hipEventSynchronize(e1) == hipSuccess ? 0 : 3;
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (ms_per_copy) *ms_per_copy = ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}
