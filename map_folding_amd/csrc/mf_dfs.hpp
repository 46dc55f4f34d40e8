// mf_dfs.hpp -- per-lane DFS state in LDS, shared by the sym and Lunnon kernels.
//
// Each lane owns one 64-entry byte row (sym: cycle successor per cell;
// Lunnon: "below" link per leaf) and one stack with an entry per DFS level
// holding the children of that level's node not yet entered.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "mf_wave.hpp"

namespace mf {

__device__ __forceinline__ uint64_t below_mask(uint32_t i) { return (i >= 64u) ? ~0ull : ((1ull << i) - 1ull); }

// Row layout, byte-interleaved: entry c of lane L at byte c*64 + L -- one VALU
// op per address (the four lanes of a quad share dwords, so lanes of a quad
// reading different entries can meet in a bank).  The dword-interleaved
// alternative (dword w of lane L at word w*64 + L: conflict-free, four VALU ops
// per address) measured 11% slower on the 8x8 depth-36 chunk.
// Entry c of a lane's row sits kStride bytes after entry c-1.  A stride that
// is not a power of two is one 24-bit multiply-add (c < 2^24); written out,
// because the compiler otherwise forms a 64-bit multiply-add for it.
template <uint32_t kStride = 64>
__device__ __forceinline__ uint32_t row_at(uint32_t lane_base, uint32_t c) {
    if constexpr ((kStride & (kStride - 1)) == 0) {
        return lane_base + c * kStride;
    } else {
        uint32_t r;
        asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(c), "s"(kStride), "v"(lane_base));
        return r;
    }
}
__device__ __forceinline__ uint32_t row_lane_base(uint32_t wbase, uint32_t lane) { return wbase + lane; }

// Byte of a lane row at `addr`: a power-of-two stride indexes the block (the
// row offset folds into the ds instruction); otherwise the absolute 32-bit
// LDS address is used (the kernels' only LDS is the dynamic block at offset 0),
// which keeps the address arithmetic 32-bit.
template <uint32_t kStride>
__device__ __forceinline__ uint32_t row_ld(const uint8_t *lds, uint32_t addr) {
    if constexpr ((kStride & (kStride - 1)) == 0) {
        return lds[addr];
    } else {
        return *reinterpret_cast<const __attribute__((address_space(3))) uint8_t *>(addr);
    }
}
template <uint32_t kStride>
__device__ __forceinline__ void row_st(uint8_t *lds, uint32_t addr, uint32_t v) {
    if constexpr ((kStride & (kStride - 1)) == 0) {
        lds[addr] = static_cast<uint8_t>(v);
    } else {
        *reinterpret_cast<__attribute__((address_space(3))) uint8_t *>(addr) = static_cast<uint8_t>(v);
    }
}

// Copy a work item's first kBytes state bytes (a multiple of 16) into the
// lane's row, and back.  The row holds kRow entries (kRow <= kBytes); state
// bytes past it are zero in every item (cells past n*n) and are not stored.
template <uint32_t kBytes = 64, uint32_t kRow = kBytes, uint32_t kStride = 64>
__device__ __forceinline__ void row_load(uint8_t *lds, uint32_t lb, const uint8_t *src) {
    const uint4 *s4 = reinterpret_cast<const uint4 *>(src);
#pragma unroll
    for (uint32_t q = 0; q < kBytes / 16; ++q) {
        const uint4 v = s4[q];
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k)
            if (16u * q + k < kRow) lds[lb + (16u * q + k) * kStride] = static_cast<uint8_t>(w[k >> 2] >> (8 * (k & 3)));
    }
}

template <uint32_t kBytes = 64, uint32_t kRow = kBytes, uint32_t kStride = 64>
__device__ __forceinline__ void row_store(const uint8_t *lds, uint32_t lb, uint8_t *dst) {
    uint4 *d4 = reinterpret_cast<uint4 *>(dst);
#pragma unroll
    for (uint32_t q = 0; q < kBytes / 16; ++q) {
        uint32_t w[4] = {0, 0, 0, 0};
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k)
            if (16u * q + k < kRow)
                w[k >> 2] |= static_cast<uint32_t>(lds[lb + (16u * q + k) * kStride]) << (8 * (k & 3));
        d4[q] = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

// The same for a row of `rows` entries known only at run time (rows <= 64):
// entries past it are zero in every item and are not stored.
template <uint32_t kStride = 64>
__device__ __forceinline__ void row_load_n(uint8_t *lds, uint32_t lb, const uint8_t *src, uint32_t rows) {
    const uint4 *s4 = reinterpret_cast<const uint4 *>(src);
#pragma unroll
    for (uint32_t q = 0; q < 4; ++q) {
        const uint4 v = s4[q];
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k)
            if (16u * q + k < rows) lds[lb + (16u * q + k) * kStride] = static_cast<uint8_t>(w[k >> 2] >> (8 * (k & 3)));
    }
}

template <uint32_t kStride = 64>
__device__ __forceinline__ void row_store_n(const uint8_t *lds, uint32_t lb, uint8_t *dst, uint32_t rows) {
    uint4 *d4 = reinterpret_cast<uint4 *>(dst);
#pragma unroll
    for (uint32_t q = 0; q < 4; ++q) {
        uint32_t w[4] = {0, 0, 0, 0};
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k)
            if (16u * q + k < rows) w[k >> 2] |= static_cast<uint32_t>(lds[lb + (16u * q + k) * kStride]) << (8 * (k & 3));
        d4[q] = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

// Lowest set bit of a u64 or u128 mask (nonzero).
__device__ __forceinline__ uint32_t mask_ctz(uint64_t v) { return static_cast<uint32_t>(__builtin_ctzll(v)); }
// The same for any mask: an empty one gives the top bit's index (for values
// computed by lanes that do not use them).
__device__ __forceinline__ uint32_t mask_ctz_any(uint64_t v) { return mask_ctz(v | (1ull << 63)); }
// u128: both halves' ctz (each defined for an empty half) and a select, no
// branch.
__device__ __forceinline__ uint32_t mask_ctz(unsigned __int128 v) {
    const uint64_t lo = static_cast<uint64_t>(v), hi = static_cast<uint64_t>(v >> 64);
    const uint32_t a = static_cast<uint32_t>(__builtin_ctzll(lo | (1ull << 63)));
    const uint32_t b = 64u + static_cast<uint32_t>(__builtin_ctzll(hi | (1ull << 63)));
    return lo != 0 ? a : b;
}
__device__ __forceinline__ uint32_t mask_ctz_any(unsigned __int128 v) {
    return mask_ctz(v | (static_cast<unsigned __int128>(1) << 127));
}
__device__ __forceinline__ uint32_t mask_popcount(uint64_t v) { return static_cast<uint32_t>(__popcll(v)); }
__device__ __forceinline__ uint32_t mask_popcount(unsigned __int128 v) {
    return static_cast<uint32_t>(__popcll(static_cast<uint64_t>(v)) + __popcll(static_cast<uint64_t>(v >> 64)));
}

struct Lane {
    uint32_t cl;      // LDS byte address of this lane's row entry 0
    uint32_t sl;      // LDS byte address of this lane's stack slot 0
    uint32_t base;    // level held by stack slot 0
    uint32_t s0;      // sl - base * slot bytes (mod 2^32): level l's slot is s0 + l * slot bytes
};

// ---------------------------------------------------------- stack encodings
// An entry names the current child j of its level's node (ctz of what is
// left) and the siblings after it.  Children are entered in ascending order,
// so "the siblings after j" is always every child of the node above j.

// One u64 mask per level.
struct WideStack {
    static constexpr uint32_t kSlotBytes = 512;   // u64 x 64 lanes
    static constexpr uint32_t kLaneBytes = 8;
    using Entry = uint64_t;
    using MaskT = uint64_t;
    static __device__ __forceinline__ uint64_t get(const uint8_t *lds, const Lane &ln, uint32_t level) {
        return *reinterpret_cast<const uint64_t *>(lds + (ln.s0 + level * kSlotBytes));
    }
    static __device__ __forceinline__ void put(uint8_t *lds, const Lane &ln, uint32_t level, uint64_t e) {
        *reinterpret_cast<uint64_t *>(lds + (ln.s0 + level * kSlotBytes)) = e;
    }
    // The same at a slot's LDS byte address.
    static __device__ __forceinline__ uint64_t get_at(const uint8_t *lds, uint32_t addr) {
        return *reinterpret_cast<const uint64_t *>(lds + addr);
    }
    static __device__ __forceinline__ void put_at(uint8_t *lds, uint32_t addr, uint64_t e) {
        *reinterpret_cast<uint64_t *>(lds + addr) = e;
    }
    static __device__ __forceinline__ uint64_t encode(uint64_t g) { return g; }
    static __device__ __forceinline__ uint64_t encode(uint64_t g, uint32_t) { return g; }
    static __device__ __forceinline__ uint64_t encode_at(uint32_t, uint64_t g, uint32_t) { return g; }
    static __device__ __forceinline__ uint32_t child(uint64_t e) { return static_cast<uint32_t>(__builtin_ctzll(e)); }
    // 0 = no sibling after the current child, 1 = advance() gives it.
    static __device__ __forceinline__ uint32_t more(uint64_t e) { return (e & (e - 1)) != 0 ? 1u : 0u; }
    static __device__ __forceinline__ bool has_next(uint64_t e) { return (e & (e - 1)) != 0; }
    static __device__ __forceinline__ bool recompute(uint64_t) { return false; }
    static __device__ __forceinline__ bool any_more(uint64_t e) { return (e & (e - 1)) != 0; }
    static __device__ __forceinline__ uint64_t advance(uint64_t e) { return e & (e - 1); }
    // The siblings after the current child as a mask (for splitting); never
    // unknown for this encoding.
    static __device__ __forceinline__ uint64_t rest(uint64_t e, bool &unknown) {
        unknown = false;
        return e & (e - 1);
    }
};

// 16 bits per level: bits [0, B) current child j, [B, 2B) next sibling j2,
// bit 2B "more": children follow that are not recorded (after j2 when j2 is
// recorded, else after j) and the node is recomputed to find them; bit 2B+1
// "next": j2 is recorded.  So an entry below kMore has no sibling left, one in
// [kMore, kNext) recomputes, and one at or above kNext advances to j2 -- two
// compares decide a backtracking step.  B = 6 for up to 64 cells (u64 masks),
// 7 for up to 128 (u128 masks).  A level's slot holds one entry per stack of
// the wave: 64, or 32 when two lanes share a subtree.
template <typename Mask, uint32_t B, uint32_t kStacks = 64>
struct CompactStackT {
    static_assert(2 * B + 2 <= 16, "entry must fit 16 bits");
    static constexpr uint32_t kSlotBytes = 2 * kStacks;   // u16 x stacks
    static constexpr uint32_t kLaneBytes = 2;
    static constexpr uint32_t kField = (1u << B) - 1u;
    static constexpr uint32_t kMore = 1u << (2 * B);
    static constexpr uint32_t kNext = 2u << (2 * B);
    using Entry = uint32_t;
    using MaskT = Mask;
    static __device__ __forceinline__ uint32_t get(const uint8_t *lds, const Lane &ln, uint32_t level) {
        uint32_t e = *reinterpret_cast<const uint16_t *>(lds + (ln.s0 + level * kSlotBytes));
        // The load zero-extends: tell the compiler nothing needs masking.
        asm("" : "+v"(e));
        return e;
    }
    static __device__ __forceinline__ void put(uint8_t *lds, const Lane &ln, uint32_t level, uint32_t e) {
        *reinterpret_cast<uint16_t *>(lds + (ln.s0 + level * kSlotBytes)) = static_cast<uint16_t>(e);
    }
    // The same at a slot's LDS byte address.
    static __device__ __forceinline__ uint32_t get_at(const uint8_t *lds, uint32_t addr) {
        uint32_t e = *reinterpret_cast<const uint16_t *>(lds + addr);
        asm("" : "+v"(e));
        return e;
    }
    static __device__ __forceinline__ void put_at(uint8_t *lds, uint32_t addr, uint32_t e) {
        *reinterpret_cast<uint16_t *>(lds + addr) = static_cast<uint16_t>(e);
    }
    static __device__ __forceinline__ uint32_t encode(Mask g) {
        const uint32_t j = mask_ctz(g);
        const Mask r = g & (g - 1);
        if (r == 0) return j;
        const uint32_t j2 = mask_ctz(r);
        return j | (j2 << B) | kNext | ((r & (r - 1)) != 0 ? kMore : 0u);
    }
    // The same with the number of children (popcount of g) already known.
    static __device__ __forceinline__ uint32_t encode(Mask g, uint32_t count) {
        const uint32_t j = mask_ctz(g);
        if (count == 1) return j;
        const uint32_t j2 = mask_ctz(g & (g - 1));
        return j | (j2 << B) | kNext | (count > 2 ? kMore : 0u);
    }
    // The same with the lowest child j (ctz of g) known too, without
    // branches (a wave runs both sides anyway): with one child the second
    // ctz reads an empty mask and is not used.
    static __device__ __forceinline__ uint32_t encode_at(uint32_t j, Mask g, uint32_t count) {
        const uint32_t j2 = mask_ctz_any(g & (g - 1)) & kField;
        const uint32_t two = j | (j2 << B) | kNext | (count > 2 ? kMore : 0u);
        return count == 1 ? j : two;
    }
    static __device__ __forceinline__ uint32_t child(uint32_t e) { return e & kField; }
    // A recorded next sibling: advance() gives it.
    static __device__ __forceinline__ bool has_next(uint32_t e) { return e >= kNext; }
    // Children after j, recorded or not.
    static __device__ __forceinline__ bool any_more(uint32_t e) { return e >= kMore; }
    // No recorded sibling but more children: recompute the node above j.
    static __device__ __forceinline__ bool recompute(uint32_t e) { return e >= kMore && e < kNext; }
    // 0 = none, 1 = a recorded next sibling (advance()), 2 = recompute.
    static __device__ __forceinline__ uint32_t more(uint32_t e) { return e >= kNext ? 1u : (e >= kMore ? 2u : 0u); }
    static __device__ __forceinline__ uint32_t advance(uint32_t e) { return ((e >> B) & kField) | (e & kMore); }
    static __device__ __forceinline__ Mask rest(uint32_t e, bool &unknown) {
        unknown = (e & kMore) != 0;
        return (e >= kNext && !unknown) ? (static_cast<Mask>(1) << ((e >> B) & kField)) : static_cast<Mask>(0);
    }
};

using CompactStack = CompactStackT<uint64_t, 6>;
using CompactStack128 = CompactStackT<unsigned __int128, 7>;

}  // namespace mf
