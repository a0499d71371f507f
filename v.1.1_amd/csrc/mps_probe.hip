// mps_probe.hip -- stream-copy probe: measured HBM ceiling of the device and the
// calibration kernel for rocprofv3 FETCH_SIZE / WRITE_SIZE counters (a copy of B bytes
// must read B and write B bytes).
#include <hip/hip_runtime.h>

#include "../../include/mps_gpu.h"

__global__ __launch_bounds__(256) void k_copy_probe(const double4* __restrict__ src, double4* __restrict__ dst,
                                                     long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

extern "C" int mps_bandwidth_probe(int32_t device, int64_t bytes, int32_t reps, double* copy_gbs) {
    if (bytes < 4096 || reps <= 0 || !copy_gbs) return MPS_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return MPS_ERR_HIP;
    const long long n = bytes / 32;
    double4 *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = MPS_OK;
    if (hipMalloc(&a, n * 32) != hipSuccess || hipMalloc(&b, n * 32) != hipSuccess ||
        hipStreamCreate(&st) != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
        hipMemsetAsync(a, 0, n * 32, st) != hipSuccess) {
        rc = MPS_ERR_HIP;
    } else {
        int cus = 256;
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, device) == hipSuccess) cus = p.multiProcessorCount;
        const int grid = cus * 8;
        k_copy_probe<<<grid, 256, 0, st>>>(a, b, n);  // warm-up
        (void)hipEventRecord(e0, st);
        for (int r = 0; r < reps; r++) k_copy_probe<<<grid, 256, 0, st>>>(a, b, n);
        (void)hipEventRecord(e1, st);
        if (hipEventSynchronize(e1) != hipSuccess) {
            rc = MPS_ERR_HIP;
        } else {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            *copy_gbs = 2.0 * (double)n * 32.0 * reps / (ms * 1e-3) / 1e9;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}
