// mf_wave.hpp -- wave64 helpers shared by the DFS kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "mf_device.hpp"

namespace mf {

// 128-bit exact accumulation into acc[lo], acc[hi]: the adder that wraps the
// low word carries into the high word.
__device__ __forceinline__ void add_u128(unsigned long long *acc, unsigned long long v) {
    if (v == 0) return;
    const unsigned long long old = atomicAdd(acc + ACC_COUNT_LO, v);
    if (old + v < old) atomicAdd(acc + ACC_COUNT_HI, 1ull);
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Wave-aggregated slot allocation: every lane with `want` gets a distinct
// index from *counter with one atomic per wave.  Must be called by the whole
// wave (converged).  Returns the lane's index (undefined when !want).
__device__ __forceinline__ uint32_t wave_alloc(uint32_t *counter, bool want, uint32_t lane, uint32_t *total_out) {
    const unsigned long long m = __ballot(want);
    if (m == 0) {
        *total_out = 0;
        return 0;
    }
    const uint32_t cnt = static_cast<uint32_t>(__popcll(m));
    const uint32_t first = static_cast<uint32_t>(__builtin_ctzll(m));
    uint32_t base = 0;
    if (lane == first) base = atomicAdd(counter, cnt);
    base = __shfl(base, first, 64);
    *total_out = cnt;
    return base + static_cast<uint32_t>(__popcll(m & ((1ull << lane) - 1ull)));
}

constexpr uint32_t kNoSlot = 0xffffffffu;

// Reserve the wave's spill slots in one step: every lane passes how many
// items it will write; lane 0 adds the wave's total to *counter with one
// fetch-and-add.  If the range would pass `cap`, the wave takes it back
// (fetch-and-sub) and gets kNoSlot in every lane.  A wave that adds while
// such an overshoot stands sees a start past the failed range and fails
// too, so the successful ranges are exactly [0, final counter) with no hole.
// (A compare-and-swap loop would avoid the overshoot, but thousands of waves
// spill at the same moment at a round's end and their retries made a round
// three times slower.)  Returns, in every lane, the first of the wave's
// `total` consecutive slots; a wave with nothing to write reserves nothing
// and gets 0.  Must be called by the whole wave (converged).
__device__ __forceinline__ uint32_t spill_reserve(uint32_t *counter, uint32_t cap, uint32_t need, uint32_t lane) {
    uint32_t total = need;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) total += __shfl_xor(total, off, 64);
    if (total == 0) return 0;
    uint32_t base = kNoSlot;
    if (lane == 0) {
        const uint32_t old = atomicAdd(counter, total);
        if (old <= cap && total <= cap - old) {
            base = old;
        } else {
            atomicSub(counter, total);
        }
    }
    return __shfl(base, 0, 64);
}

// The same reservation with per-lane ranges: returns the lane's first slot
// (lanes' ranges are consecutive in lane order), or kNoSlot in every lane if
// the wave's total does not fit (nothing reserved).
__device__ __forceinline__ uint32_t lane_reserve(uint32_t *counter, uint32_t cap, uint32_t need, uint32_t lane) {
    uint32_t incl = need;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t v = __shfl_up(incl, off, 64);
        if (lane >= static_cast<uint32_t>(off)) incl += v;
    }
    const uint32_t total = __shfl(incl, 63, 64);
    if (total == 0) return 0;
    uint32_t base = kNoSlot;
    if (lane == 0) {
        const uint32_t old = atomicAdd(counter, total);
        if (old <= cap && total <= cap - old) {
            base = old;
        } else {
            atomicSub(counter, total);
        }
    }
    base = __shfl(base, 0, 64);
    return base == kNoSlot ? kNoSlot : base + (incl - need);
}

// Copy 64 state bytes between an item and a lane's lane-interleaved LDS rows
// (dword w of the lane at lane_base + w*256).
__device__ __forceinline__ void state_to_lds(uint8_t *lds, uint32_t lane_base, const uint8_t *src) {
    const uint4 *s4 = reinterpret_cast<const uint4 *>(src);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint4 v = s4[q];
        *reinterpret_cast<uint32_t *>(lds + lane_base + (4 * q + 0) * 256u) = v.x;
        *reinterpret_cast<uint32_t *>(lds + lane_base + (4 * q + 1) * 256u) = v.y;
        *reinterpret_cast<uint32_t *>(lds + lane_base + (4 * q + 2) * 256u) = v.z;
        *reinterpret_cast<uint32_t *>(lds + lane_base + (4 * q + 3) * 256u) = v.w;
    }
}

__device__ __forceinline__ void lds_to_state(const uint8_t *lds, uint32_t lane_base, uint8_t *dst) {
    uint4 *d4 = reinterpret_cast<uint4 *>(dst);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        uint4 v;
        v.x = *reinterpret_cast<const uint32_t *>(lds + lane_base + (4 * q + 0) * 256u);
        v.y = *reinterpret_cast<const uint32_t *>(lds + lane_base + (4 * q + 1) * 256u);
        v.z = *reinterpret_cast<const uint32_t *>(lds + lane_base + (4 * q + 2) * 256u);
        v.w = *reinterpret_cast<const uint32_t *>(lds + lane_base + (4 * q + 3) * 256u);
        d4[q] = v;
    }
}

// Byte address of entry `idx` of a lane's 64-byte row (4 entries per dword,
// dword w of the lane at lane_base + w*256): lanes never share a bank.
__device__ __forceinline__ uint32_t row_addr(uint32_t lane_base, uint32_t idx) {
    return lane_base + ((idx & ~3u) << 6) + (idx & 3u);
}

}  // namespace mf
