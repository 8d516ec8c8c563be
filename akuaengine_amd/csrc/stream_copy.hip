// stream_copy.hip -- a plain device-to-device copy kernel of known size: measures the achievable
// HBM bandwidth beside the 8 TB/s peak, and calibrates the rocprofv3 FETCH_SIZE / WRITE_SIZE
// counters (a copy of B bytes must read B and write B).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/akua.h"

namespace {

__global__ __launch_bounds__(256) void k_stream_copy(const float4* __restrict__ src,
                                                     float4* __restrict__ dst, size_t n) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * 256;
    for (; i < n; i += stride) dst[i] = src[i];
}

}  // namespace

extern "C" akua_status akua_stream_copy_bandwidth(int64_t bytes, int32_t reps, double* gbps) {
    if (bytes < 16 || reps < 1 || !gbps) return AKUA_ERR_INVALID_ARG;
    const size_t n = (size_t)bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    akua_status st = AKUA_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return AKUA_ERR_NO_DEVICE;
    if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess) {
        st = AKUA_ERR_OUT_OF_MEMORY;
    } else if (hipMemset(a, 0, n * 16) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
               hipEventCreate(&e1) != hipSuccess) {
        st = AKUA_ERR_HIP;
    } else {
        const unsigned grid = 256u * 8u;  // 8 workgroups per CU, grid-stride
        k_stream_copy<<<grid, 256>>>(a, b, n);  // warm-up
        (void)hipEventRecord(e0, nullptr);
        for (int r = 0; r < reps; ++r) k_stream_copy<<<grid, 256>>>(a, b, n);
        (void)hipEventRecord(e1, nullptr);
        float ms = 0.0f;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            st = AKUA_ERR_HIP;
        else
            *gbps = 2.0 * (double)n * 16.0 * reps / (ms * 1e-3) / 1e9;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return st;
}
