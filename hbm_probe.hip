// hbm_probe.hip -- calibration kernel for HBM traffic counters: copies a
// buffer of known size so that rocprofv3's FETCH_SIZE / WRITE_SIZE readings
// of the search kernels can be scaled to bytes.
//   hbm_probe [MiB]     (default 1024)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

int main(int argc, char** argv) {
    const size_t mib = argc > 1 ? strtoull(argv[1], nullptr, 10) : 1024;
    const size_t bytes = mib << 20;
    const size_t n = bytes / sizeof(uint4);
    uint4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        fprintf(stderr, "hbm_probe: allocation failed\n");
        return 1;
    }
    hipMemset(a, 1, bytes);
    hipMemset(b, 0, bytes);
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    for (int rep = 0; rep < 3; ++rep) {
        hipEventRecord(e0);
        hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, 0, a, b, n);
        hipEventRecord(e1);
        hipEventSynchronize(e1);
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        printf("copy %zu MiB: %.3f ms, %.1f GB/s (read+write)\n", mib, ms, 2.0 * bytes / (ms * 1e-3) / 1e9);
    }
    hipFree(a);
    hipFree(b);
    return 0;
}
