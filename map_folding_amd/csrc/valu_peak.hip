// valu_peak.hip -- measures the int32 VALU throughput the roofline divides by.
//
// Every lane of a full grid (32 waves per CU) runs 16 independent chains of
// `a += b; b ^= a` (two full-rate 32-bit VALU ops per chain step, no closed
// form the compiler could fold), so the SIMDs issue back-to-back VALU work
// with enough independent instructions to cover the ALU latency.  The kernel
// time comes from HIP events (best of 5); ops/s = lanes x iterations x 16
// chains x 2 ops / time.  The chains' xor is written out so that no work is
// dead.  On one MI355X this reads 6.1-6.2e13 ops/s, 1.57x the CUs x 64 lanes
// x 2.4 GHz figure, which is therefore no ceiling for int32 VALU work; the
// bench divides by this measured rate.  The last line is `peak_ops_per_s V`
// for scripts.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

constexpr int kChains = 16;

__global__ __launch_bounds__(256) void valu_peak_kernel(uint32_t iters, uint32_t *out) {
    uint32_t a[kChains], b[kChains];
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
    for (int c = 0; c < kChains; ++c) {
        a[c] = t * 2654435761u + static_cast<uint32_t>(c);
        b[c] = t ^ (0x9e3779b9u * static_cast<uint32_t>(c + 1));
    }
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) {
            a[c] += b[c];
            b[c] ^= a[c];
        }
    }
    uint32_t x = 0;
#pragma unroll
    for (int c = 0; c < kChains; ++c) x ^= a[c] ^ b[c];
    out[t] = x;
}

int main() {
    int dev = 0, cus = 0, clock_khz = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 1;
    if (hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeClockRate, dev) != hipSuccess) return 1;
    const uint32_t block = 256, waves_per_cu = 32;
    const uint32_t grid = static_cast<uint32_t>(cus) * waves_per_cu / (block / 64);
    const uint32_t iters = 1u << 16;
    uint32_t *out = nullptr;
    if (hipMalloc(&out, static_cast<size_t>(grid) * block * sizeof(uint32_t)) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    hipLaunchKernelGGL(valu_peak_kernel, dim3(grid), dim3(block), 0, 0, iters, out);   // warm-up
    if (hipDeviceSynchronize() != hipSuccess) return 1;
    float best_ms = 1e30f;
    for (int r = 0; r < 5; ++r) {
        if (hipEventRecord(e0, 0) != hipSuccess) return 1;
        hipLaunchKernelGGL(valu_peak_kernel, dim3(grid), dim3(block), 0, 0, iters, out);
        if (hipEventRecord(e1, 0) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 1;
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1;
        if (ms < best_ms) best_ms = ms;
    }
    const double ops = static_cast<double>(grid) * block * iters * kChains * 2.0;
    const double measured = ops / (best_ms * 1e-3);
    const double spec = static_cast<double>(cus) * 64.0 * 2.4e9;
    std::printf("cus %d reported_clock_mhz %.0f grid %u block %u iters %u kernel_ms %.3f\n", cus, clock_khz / 1e3, grid,
                block, iters, best_ms);
    std::printf("int32 VALU ops/s measured %.4e, spec (CUs x 64 x 2.4 GHz) %.4e, ratio %.3f\n", measured, spec,
                measured / spec);
    std::printf("peak_ops_per_s %.6e\n", measured);
    (void)hipFree(out);
    return 0;
}
