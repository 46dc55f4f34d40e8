// hbm_copy.hip -- calibration program for the HBM traffic counters.
//
// Copies 1 GiB three times with a plain grid-stride kernel.  Profiling it with
// `rocprofv3 --pmc FETCH_SIZE` and `--pmc WRITE_SIZE` gives what the counters
// read for a known 1 GiB read and 1 GiB write per dispatch; the DFS kernels'
// counter readings are scaled by those factors.
#include <hip/hip_runtime.h>

#include <cstdio>

__global__ void hbm_copy_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

int main() {
    const size_t bytes = 1ull << 30;
    const size_t n = bytes / sizeof(uint4);
    uint4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
    if (hipMemset(a, 1, bytes) != hipSuccess) return 1;
    for (int r = 0; r < 3; ++r) {
        hipLaunchKernelGGL(hbm_copy_kernel, dim3(8192), dim3(256), 0, 0, a, b, n);
        if (hipDeviceSynchronize() != hipSuccess) return 1;
    }
    std::printf("copied %zu bytes, 3 dispatches\n", bytes);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
