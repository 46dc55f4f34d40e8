// sym_kernels.hip -- symmetry-reduced cycle DFS on gfx950.
//
// Restates recurse_count / get_good / iter_good (sym_fold.cpp:179-233) as a
// per-lane iterative DFS:
//   * one lane owns one subtree (a WorkItem);
//   * its cycle (cell -> successor, bytes) lives in LDS, byte-interleaved
//     across the wave's lanes (entry c of lane L at byte c*S + L, S = 68 for
//     the u64 instances so that quad lanes reading different cells fall in
//     different banks, 64 for 9x9);
//   * the partner of every cell on each of the 4 pages sits in LDS tables
//     (byte p of word c, and page-major bytes for the walks), with one page
//     word per level naming the one or two pages its get_good walks;
//   * node = one get_good call (sym_fold9.cpp:342-344); a node at the last
//     level adds popcount(good) leaves instead of recursing.
//
// The DFS stack holds, per level, the children of that level's node not yet
// entered.  Two encodings:
//   compact (default)  16 bits per level: the current child j, the next
//                      sibling, and whether more siblings follow.  When a
//                      level's known siblings run out but more follow, the
//                      node's good mask is recomputed (uncounted) and masked
//                      to the children above j.  Nodes with three or more
//                      children are ~10% of the tree at 8x8 depth 36, so the
//                      recomputes cost little, while the stack shrinks from
//                      8 to 2 bytes per level -- 118 instead of 280 bytes of
//                      LDS per lane, 20 instead of 8 resident waves per CU.
//   wide               one u64 mask per level (kept for comparison).
// Walk semantics: for each page with an earlier neighbour (the anchor), get_good
// follows cycle successors from the anchor; a cell whose page partner is
// already placed is replaced by that partner; every cell reached is a legal
// insertion point; the page masks are intersected.  A level's two walks run
// in one loop (walk_mode below), which runs until every lane of the wave has
// closed both of its walks.
//
// Two schedules share the core:
//   static      one lane per input item, one launch (the plain static-frontier
//               schedule);
//   persistent  resident lanes claim input items a wave at a time; once the
//               input is drained, a lane still busy after `tail` iterations
//               splits its remaining work -- the node it is about to expand
//               plus the untried siblings of every level on its stack -- into
//               the spill queue, which is the next round's input.
// Counts never depend on encoding or schedule: every subtree is counted once.
#include <hip/hip_runtime.h>

#include "mf_counters.hpp"
#include "mf_device.hpp"
#include "mf_dfs.hpp"
#include "mf_wave.hpp"

namespace mf {
namespace {

// Cell geometry of a launch: up to 64 cells (u64 masks, 80-byte items) or up
// to 96 (9x9: 81 cells, u128 masks, 128-byte items).  LDS: a partner word per
// cell (bits 8p..8p+6 = partner on page p, n2 = none; bit 8p+7 set when that
// partner is below the cell, i.e. get_good walks page p for it), the
// partners page-major (byte kPageOff + kCells*p + c), a level word per cell
// (the pages its get_good walks, level_word below), then per wave a
// kCells-entry byte row per lane and the stack.
// Row stride of the u64 geometries: 68 bytes (17 dwords, odd) instead of 64
// puts the 4 lanes of a quad in different banks when they read different
// cells; 2.3% less time on the 8x8 chunk and 3.5% on 7x7, for 4 more bytes
// per entry (5 blocks per CU at 8x8 depth 36 and 6 at 7x7 depth 24, as with
// 64).  9x9 keeps 64: its 81-entry rows would lose a block per CU.
#ifndef MF_SYM_ROW_STRIDE
#define MF_SYM_ROW_STRIDE 68
#endif
// Cells below i / at or above f as masks.  For u128 masks (i, f < 128) the
// generic 128-bit shifts cost about a dozen instructions; here one
// v_lshlrev_b64 of all ones (it shifts by the amount's low 6 bits), bit 6 of
// the amount as an all-ones / zero mask m, and one v_bitop3_b32 per dword
// (truth table of (S0, S1) = (0xF0, 0xCC), S2 unused): with s = ~0 << (i & 63),
// below(i) = (~s | m, ~s & m) -> 0xCF, 0x0C; from(f) = (s & ~m, s | ~m) ->
// 0x30, 0xF3 (low half, high half).
__device__ __forceinline__ uint64_t mask_below(uint64_t, uint32_t i) { return (1ull << i) - 1ull; }
__device__ __forceinline__ uint64_t mask_from(uint64_t, uint32_t f) { return ~0ull << f; }
template <uint32_t kLo, uint32_t kHi>
__device__ __forceinline__ unsigned __int128 halves128(uint32_t i) {
    uint64_t sh;
    asm("v_lshlrev_b64 %0, %1, -1" : "=v"(sh) : "v"(i));
    const uint32_t m = static_cast<uint32_t>(static_cast<int32_t>(i << 25) >> 31);   // i >= 64
    const uint32_t s0 = static_cast<uint32_t>(sh), s1 = static_cast<uint32_t>(sh >> 32);
    uint32_t d0, d1, d2, d3;
    asm("v_bitop3_b32 %0, %1, %2, %2 bitop3:%3" : "=v"(d0) : "v"(s0), "v"(m), "i"(kLo));
    asm("v_bitop3_b32 %0, %1, %2, %2 bitop3:%3" : "=v"(d1) : "v"(s1), "v"(m), "i"(kLo));
    asm("v_bitop3_b32 %0, %1, %2, %2 bitop3:%3" : "=v"(d2) : "v"(s0), "v"(m), "i"(kHi));
    asm("v_bitop3_b32 %0, %1, %2, %2 bitop3:%3" : "=v"(d3) : "v"(s1), "v"(m), "i"(kHi));
    return (static_cast<unsigned __int128>((static_cast<uint64_t>(d3) << 32) | d2) << 64) |
           ((static_cast<uint64_t>(d1) << 32) | d0);
}
[[maybe_unused]] __device__ __forceinline__ unsigned __int128 mask_below(unsigned __int128, uint32_t i) {
    return halves128<0xCF, 0x0C>(i);
}
[[maybe_unused]] __device__ __forceinline__ unsigned __int128 mask_from(unsigned __int128, uint32_t f) {
    return halves128<0x30, 0xF3>(f);
}

template <uint32_t kCells_, typename Mask_, typename Item_, typename Stack_, uint32_t kRowCells_ = kCells_,
          uint32_t kRowStride_ = MF_SYM_ROW_STRIDE, uint32_t kLanes_ = 1, int kWalk_ = -1>
struct Geometry {
    static constexpr uint32_t kCells = kCells_;
    // get_good's walks: 1 one page at a time, 2 a level's one or two pages
    // in one paired loop, -1 by mask width (walk_mode below).
    static constexpr int kWalk = kWalk_;
    // Lanes per subtree: 1, or 2 (a lane pair shares one row and one stack,
    // keeps identical DFS state, and splits each node's page walks).
    static constexpr uint32_t kLanes = kLanes_;
    static constexpr uint32_t kRowCells = kRowCells_;   // entries of a lane's LDS row (cells of the largest map)
    static constexpr uint32_t kRowStride = kRowStride_;   // bytes from entry c to entry c+1 of a lane
    using Mask = Mask_;
    using Item = Item_;
    using Stack = Stack_;
    static constexpr uint32_t kPartnerOff = 0;
    static constexpr uint32_t kPageOff = 4 * kCells;
    static constexpr uint32_t kLevelOff = 8 * kCells;
    static constexpr uint32_t kTableBytes = 12 * kCells;
    // Every lane's row, rounded so that the stack after it stays 16-byte aligned.
    static constexpr uint32_t kCycleBytes = (kRowCells * kRowStride + 15u) & ~15u;
    static constexpr uint32_t kStateBytes = kCells;
    // Cells below i.  i is a level (at most n2 - 1) or a floor j + 1 with
    // j < i, so it stays under the mask width (64 up to 8x8, 128 for 9x9).
    static __device__ __forceinline__ Mask below(uint32_t i) { return mask_below(static_cast<Mask>(0), i); }
    // Cells at or above f (a recompute's floor, f < the mask width).
    static __device__ __forceinline__ Mask from(uint32_t f) { return mask_from(static_cast<Mask>(0), f); }
    static __device__ __forceinline__ Mask item_mask(const WorkItem *it) { return it->mask; }
    static __device__ __forceinline__ Mask item_mask(const WorkItem128 *it) {
        return (static_cast<Mask>(it->mask_hi) << 64) | it->mask_lo;
    }
    static __device__ __forceinline__ void set_item_mask(WorkItem *it, Mask m) { it->mask = static_cast<uint64_t>(m); }
    static __device__ __forceinline__ void set_item_mask(WorkItem128 *it, Mask m) {
        it->mask_lo = static_cast<uint64_t>(m);
        it->mask_hi = static_cast<uint64_t>(static_cast<unsigned __int128>(m) >> 64);
    }
};

using G64 = Geometry<64, uint64_t, WorkItem, CompactStack>;
using G64Wide = Geometry<64, uint64_t, WorkItem, WideStack>;
// 9x9: 81 cells in LDS rows (5,184 B per wave instead of 6,144), so four
// 4-wave blocks fit a CU at depth 44 instead of three.
// 9x9 rows at a 65-byte stride: lanes of a quad reading different cells land
// in different banks, and 81 rows of 65 B (5,280 B per wave) still fit four
// blocks per CU (68 would not); 5.2% less time on the 9x9 depth-44 window
// than 64 (the u64 instances measured 0.2% slower at 65 than at 68).
#ifndef MF_SYM96_ROW_STRIDE
#define MF_SYM96_ROW_STRIDE 65
#endif
using G96 = Geometry<96, unsigned __int128, WorkItem128, CompactStack128, 81, MF_SYM96_ROW_STRIDE>;
// Up to 7x7: 49-entry rows (3,136 B per wave): six 4-wave blocks per CU at
// 7x7 depth 24 instead of five.
// The 49-cell rows have room for a wider stride within the 7 blocks per CU
// of the 7x7 chunk: 73 bytes runs it 0.7% faster than 68 (71 and 75 the
// same, 65 0.5% slower).
#ifndef MF_SYM49_ROW_STRIDE
#define MF_SYM49_ROW_STRIDE 73
#endif
using G49 = Geometry<64, uint64_t, WorkItem, CompactStack, 49, MF_SYM49_ROW_STRIDE>;
// Two lanes per subtree, up to 8x8: 32 rows per wave at a 36-byte stride (9
// dwords, odd) and 32 stacks, half the LDS of G64 per wave.
using G64Pair = Geometry<64, uint64_t, WorkItem, CompactStackT<uint64_t, 6, 32>, 64, 36, 2>;
// Up to 8x8 with one page walk at a time: for cell numberings in which some
// level walks no page or more than two (ring orders, other centers), which
// the paired walk does not cover.
using G64Any = Geometry<64, uint64_t, WorkItem, CompactStack, 64, MF_SYM_ROW_STRIDE, 1, 1>;
using G96Any = Geometry<96, unsigned __int128, WorkItem128, CompactStack128, 81, MF_SYM96_ROW_STRIDE, 1, 1>;

// The geometry of a launch: `cells` = n*n of the map; core 3 = two lanes per
// subtree (re-split schedule, up to 8x8); cores 4 (up to 8x8) and 5 (9x9) =
// one page walk at a time (sym_walk_core picks them).
enum class Geo { k49, k64, k64Wide, k96, k64Pair, k64Any, k96Any };
inline Geo geometry(int core, uint32_t cells) {
    if (cells > 64) return core == 5 ? Geo::k96Any : Geo::k96;
    if (core == 1) return Geo::k64Wide;
    if (core == 3) return Geo::k64Pair;
    if (core == 4) return Geo::k64Any;
    return cells <= 49 ? Geo::k49 : Geo::k64;
}

template <typename G>
__device__ __forceinline__ uint32_t cyc_addr(uint32_t lane_base, uint32_t cell) {
    return row_at<G::kRowStride>(lane_base, cell);
}
// Cycle entry of `cell` in the lane's row.  With a stride that is not a power
// of two the entry is read by absolute 32-bit LDS address (the kernel's only
// LDS is the dynamic block at offset 0, checked when the tables are staged),
// which keeps the address a single 24-bit multiply-add; a power-of-two stride
// indexes the block, which lets the row offset fold into the ds instruction.
template <typename G>
__device__ __forceinline__ uint32_t cyc_ld(const uint8_t *lds, uint32_t lane_base, uint32_t cell) {
    if constexpr ((G::kRowStride & (G::kRowStride - 1)) == 0) {
        return lds[cyc_addr<G>(lane_base, cell)];
    } else {
        return *reinterpret_cast<const __attribute__((address_space(3))) uint8_t *>(cyc_addr<G>(lane_base, cell));
    }
}
template <typename G>
__device__ __forceinline__ void cyc_st(uint8_t *lds, uint32_t lane_base, uint32_t cell, uint32_t v) {
    if constexpr ((G::kRowStride & (G::kRowStride - 1)) == 0) {
        lds[cyc_addr<G>(lane_base, cell)] = static_cast<uint8_t>(v);
    } else {
        *reinterpret_cast<__attribute__((address_space(3))) uint8_t *>(cyc_addr<G>(lane_base, cell)) =
            static_cast<uint8_t>(v);
    }
}

// Insert cell i after cell j (iter_good, sym_fold.cpp:210-213).
template <typename G>
__device__ __forceinline__ void place(uint8_t *lds, const Lane &ln, uint32_t j, uint32_t i) {
    const uint32_t nx = cyc_ld<G>(lds, ln.cl, j);
    cyc_st<G>(lds, ln.cl, i, nx);
    cyc_st<G>(lds, ln.cl, j, i);
}
// Cycle entry at an LDS byte address (from cyc_addr).
template <typename G>
__device__ __forceinline__ uint32_t cyc_ld_at(const uint8_t *lds, uint32_t addr) {
    if constexpr ((G::kRowStride & (G::kRowStride - 1)) == 0) {
        return lds[addr];
    } else {
        return *reinterpret_cast<const __attribute__((address_space(3))) uint8_t *>(addr);
    }
}
template <typename G>
__device__ __forceinline__ void cyc_st_at(uint8_t *lds, uint32_t addr, uint32_t v) {
    if constexpr ((G::kRowStride & (G::kRowStride - 1)) == 0) {
        lds[addr] = static_cast<uint8_t>(v);
    } else {
        *reinterpret_cast<__attribute__((address_space(3))) uint8_t *>(addr) = static_cast<uint8_t>(v);
    }
}
// place() with the address of cell i's entry already known.
template <typename G>
__device__ __forceinline__ void place_at(uint8_t *lds, const Lane &ln, uint32_t j, uint32_t i, uint32_t addr_i) {
    const uint32_t nx = cyc_ld<G>(lds, ln.cl, j);
    cyc_st_at<G>(lds, addr_i, nx);
    cyc_st<G>(lds, ln.cl, j, i);
}
// Remove cell i again (sym_fold.cpp:215): its successor goes back to j.
template <typename G>
__device__ __forceinline__ void unplace(uint8_t *lds, const Lane &ln, uint32_t j, uint32_t i) {
    cyc_st<G>(lds, ln.cl, j, cyc_ld<G>(lds, ln.cl, i));
}

struct Work {
    uint32_t i;        // level of the node to expand next
    uint32_t root;     // item root level: finished when backtracking reaches it
    uint32_t origin;
    uint32_t mult;
    uint32_t floor;    // 0, or j+1: the next expansion recomputes node i and
                       // keeps only children above j (not counted as a node)
    uint64_t leaves;   // leaves of the current item so far
    uint32_t inodes;   // nodes, page walks and walk steps of the current item
    uint32_t iwalks;   // (32-bit: folded into the lane totals before they can
    uint32_t isteps;   // wrap, fold_item below)
    bool have;
    bool popping;      // backtracking continues this step (no node to expand)
    bool primary;      // the lane that counts and spills for its subtree (lanes per subtree > 1)
};

struct Totals {
    unsigned long long sum = 0;     // sum of mult * leaves
    unsigned long long nodes = 0, walks = 0, steps = 0, leaves = 0, items = 0;
};

// get_good (sym_fold.cpp:179-204; u128 masks as sym_fold9.cpp:286-311) for
// cell i against this lane's cycle.  Lanes of a wave sit at different levels,
// so they walk different subsets of the 4 pages (about 2 each).  Walk modes:
//   0  one loop per page index: the wave runs every page's loop;
//   1  rounds: round k walks each lane's k-th page with an earlier neighbour
//      (ascending, so the walks and their early exit are unchanged), and the
//      wave runs max(#pages) rounds; the page offset becomes a per-lane
//      value (12% less time than mode 0 on the 8x8 depth-36 chunk);
//   2  pairs: the level's two pages (every cell past the spiral's first few
//      walks one or two; the level table names them) walk in one loop, both
//      chains stepping every iteration, so that their LDS loads are in
//      flight together: a walk step is two dependent LDS loads.  A closed
//      chain walks its orbit again under a scalar open mask, which adds
//      nothing to its reach: 7 VALU per chain step.  13% less time than
//      mode 1 on the 8x8 chunk.
// The default is pairs for every one-lane geometry whose levels all walk one
// or two pages (every spiral numbering; G64Any and G96Any, for the other
// numberings, run rounds).  For u128 masks (9x9) pairs pay since each reach
// bit is 6 VALU (or_bit): 10% less time than rounds on the 9x9 depth-44
// window (with the generic 128-bit shifts they had been 5.6% slower).
// MF_SYM_WALK_MODE overrides it.
#ifndef MF_SYM_WALK_MODE
#define MF_SYM_WALK_MODE -1
#endif
// Wave priority inside the walk loops (1: 1.4% less time on the 8x8 chunk
// than 0; 2 and 3 the same as 1).
#ifndef MF_SYM_WALK_PRIO
#define MF_SYM_WALK_PRIO 1
#endif
constexpr int kWalkPrio = MF_SYM_WALK_PRIO;
// The paired loop runs two steps per trip: a closed chain's extra steps
// change nothing, so the exit test runs every second step (3% less time on
// the 8x8 and 7x7 chunks than every step; three and four per trip are 1.9%
// and 2.6% slower on the 8x8 chunk).
#ifndef MF_SYM_WALK_UNROLL
#define MF_SYM_WALK_UNROLL 2
#endif
constexpr int kWalkUnroll = MF_SYM_WALK_UNROLL;
template <typename G>
constexpr int walk_mode() {
    // (Two lanes per subtree split a level's pages between the pair.)
    return G::kLanes == 2         ? 1
           : MF_SYM_WALK_MODE >= 0 ? MF_SYM_WALK_MODE
           : G::kWalk >= 0         ? G::kWalk
                                   : 2;
}

// Where a page-p walk moves from cell j at level i: the page partner if it is
// already placed, else j itself.  `prow` is the page's row of partner bytes.
template <typename G>
__device__ __forceinline__ uint32_t page_row(uint32_t p) { return G::kPageOff + G::kCells * p; }
template <typename G>
__device__ __forceinline__ uint32_t hop(uint32_t prow, uint32_t i, uint32_t j) {
    // Absolute LDS address (the kernel's only LDS is the dynamic block at
    // offset 0): a constant row offset folds into the ds instruction.
    const uint32_t v = *reinterpret_cast<const __attribute__((address_space(3))) uint8_t *>(prow + j);
    return v < i ? v : j;
}

// reach |= bit j.  For u128 masks (cells < 128) the generic 128-bit shift
// costs about 15 instructions; here: the bit in a 64-bit word
// (v_lshlrev_b64 shifts by the amount's low 6 bits), bit 6 of j as an
// all-ones / zero mask, and one v_bitop3_b32 per dword that ORs in the word
// under its half's mask (truth table of (S0, S1, S2) = (0xF0, 0xCC, 0xAA):
// S0 | (S1 & ~S2) = 0xF4, S0 | (S1 & S2) = 0xF8).
__device__ __forceinline__ void or_bit(uint64_t &r, uint32_t j) { r |= 1ull << j; }
[[maybe_unused]] __device__ __forceinline__ void or_bit(unsigned __int128 &r, uint32_t j) {
    uint64_t b;
    asm("v_lshlrev_b64 %0, %1, 1" : "=v"(b) : "v"(j));
    const uint32_t hi = static_cast<uint32_t>(static_cast<int32_t>(j << 25) >> 31);   // j >= 64
    const uint32_t bl = static_cast<uint32_t>(b), bh = static_cast<uint32_t>(b >> 32);
    uint32_t d0 = static_cast<uint32_t>(r), d1 = static_cast<uint32_t>(r >> 32), d2 = static_cast<uint32_t>(r >> 64),
             d3 = static_cast<uint32_t>(r >> 96);
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0xf4" : "=v"(d0) : "v"(d0), "v"(bl), "v"(hi));
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0xf4" : "=v"(d1) : "v"(d1), "v"(bh), "v"(hi));
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0xf8" : "=v"(d2) : "v"(d2), "v"(bl), "v"(hi));
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0xf8" : "=v"(d3) : "v"(d3), "v"(bh), "v"(hi));
    r = (static_cast<unsigned __int128>((static_cast<uint64_t>(d3) << 32) | d2) << 64) |
        ((static_cast<uint64_t>(d1) << 32) | d0);
}

// One page walk from `anchor`: every cell reached is a legal insertion point.
template <typename G>
__device__ __forceinline__ typename G::Mask page_walk(const uint8_t *__restrict__ lds, uint32_t cl, uint32_t i,
                                                      uint32_t prow, uint32_t anchor, uint32_t &steps) {
    using Mask = typename G::Mask;
    Mask reach = 0;
    or_bit(reach, anchor);
    uint32_t j = cyc_ld<G>(lds, cl, anchor);
    __builtin_amdgcn_s_setprio(kWalkPrio);   // as in the paired walk (good_mask)
    while (j != anchor) {
        ++steps;
        j = hop<G>(prow, i, j);
        or_bit(reach, j);
        j = cyc_ld<G>(lds, cl, j);
    }
    __builtin_amdgcn_s_setprio(0);
    return reach;
}

// a & b & ~m, one v_bitop3_b32 per dword (truth table of (S0, S1, S2) =
// (0xF0, 0xCC, 0xAA): 0xF0 & 0xCC & ~0xAA = 0x40).
__device__ __forceinline__ uint64_t and_and_not(uint64_t a, uint64_t b, uint64_t m) {
    uint32_t lo, hi;
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x40"
        : "=v"(lo)
        : "v"(static_cast<uint32_t>(a)), "v"(static_cast<uint32_t>(b)), "v"(static_cast<uint32_t>(m)));
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x40"
        : "=v"(hi)
        : "v"(static_cast<uint32_t>(a >> 32)), "v"(static_cast<uint32_t>(b >> 32)),
          "v"(static_cast<uint32_t>(m >> 32)));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}
// a | b | c, one v_bitop3_b32 per dword (0xF0 | 0xCC | 0xAA = 0xFE): the
// three-input OR of the bitwise class that issues at the full rate, where
// v_or3_b32 issues at about two thirds of it on gfx950.
__device__ __forceinline__ uint64_t or3(uint64_t a, uint64_t b, uint64_t c) {
    uint32_t lo, hi;
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0xfe"
        : "=v"(lo)
        : "v"(static_cast<uint32_t>(a)), "v"(static_cast<uint32_t>(b)), "v"(static_cast<uint32_t>(c)));
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0xfe"
        : "=v"(hi)
        : "v"(static_cast<uint32_t>(a >> 32)), "v"(static_cast<uint32_t>(b >> 32)),
          "v"(static_cast<uint32_t>(c >> 32)));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}
[[maybe_unused]] __device__ __forceinline__ unsigned __int128 or3(unsigned __int128 a, unsigned __int128 b,
                                                                 unsigned __int128 c) {
    return a | b | c;
}
[[maybe_unused]] __device__ __forceinline__ unsigned __int128 and_and_not(unsigned __int128 a, unsigned __int128 b,
                                                         unsigned __int128 m) {
    const uint64_t lo = and_and_not(static_cast<uint64_t>(a), static_cast<uint64_t>(b), static_cast<uint64_t>(m));
    const uint64_t hi = and_and_not(static_cast<uint64_t>(a >> 64), static_cast<uint64_t>(b >> 64),
                                    static_cast<uint64_t>(m >> 64));
    return (static_cast<unsigned __int128>(hi) << 64) | lo;
}

// The pages get_good walks at level i, as one word: bits 0-6 the first
// walked page's anchor (its partner of cell i), bit 7 set when a second page
// is walked, bits 8-14 the second anchor, bits 16-23 / 24-31 the two pages'
// partner-row offsets from kPageOff, shifted right by row_shift (96-cell
// tables: the offsets 96p are multiples of 32).  With one page the second
// repeats the first.  kLevelGeneric: no page or more than two; a map with
// such a level runs one page walk at a time (sym_walk_core).
constexpr uint32_t kLevelGeneric = 0x8000u;
template <typename G>
constexpr uint32_t row_shift() {
    return 3u * G::kCells <= 255u ? 0u : 5u;
}
template <typename G>
__host__ __device__ __forceinline__ uint32_t level_word(uint32_t partners, uint32_t i) {
    static_assert(((3u * G::kCells) >> row_shift<G>()) <= 255u &&
                      ((G::kCells >> row_shift<G>()) << row_shift<G>()) == G::kCells,
                  "page rows must fit the level word");
    uint32_t n = 0, a0 = 0, a1 = 0, r0 = 0, r1 = 0;
#pragma unroll
    for (uint32_t p = 0; p < 4; ++p) {
        const uint32_t o = (partners >> (8u * p)) & 0xffu;
        if (o < i) {
            if (n == 0) {
                a0 = o;
                r0 = (G::kCells * p) >> row_shift<G>();
            } else if (n == 1) {
                a1 = o;
                r1 = (G::kCells * p) >> row_shift<G>();
            }
            ++n;
        }
    }
    if (n == 0 || n > 2) return kLevelGeneric;
    if (n == 1) {
        a1 = a0;
        r1 = r0;
    }
    return a0 | (n == 2 ? 0x80u : 0u) | (a1 << 8) | (r0 << 16) | (r1 << 24);
}

template <typename G, bool EVC>
__device__ __forceinline__ typename G::Mask good_mask(const uint8_t *__restrict__ lds, uint32_t cl, uint32_t i,
                                                      uint32_t &walks, uint32_t &steps) {
    using Mask = typename G::Mask;
    Mask good = G::below(i);
    // One load brings the 4 partners and the flags of the pages to walk: a
    // flag at bit 8p+7 gives the page (p = bit >> 3) and the partner's shift
    // (bit - 7).
    const uint32_t ew = reinterpret_cast<const uint32_t *>(lds + G::kPartnerOff)[i];
    uint32_t act = ew & 0x80808080u;
    if constexpr (G::kLanes == 2) {
        // Two lanes per subtree: per round the even lane walks the first of
        // the next two pages, the odd lane the second (or the first again
        // when there is no second), each to its own chain's end; the pair
        // swaps reaches and step counts.  Both lanes hold the same state, so
        // both run every round and the swap always finds its partner.
        const bool odd = (threadIdx.x & 1u) != 0u;
        while (act != 0 && good != 0) {
            const uint32_t b1 = static_cast<uint32_t>(__builtin_ctz(act));
            act &= act - 1u;
            uint32_t b2 = b1;
            const bool two = act != 0;
            if (two) {
                b2 = static_cast<uint32_t>(__builtin_ctz(act));
                act &= act - 1u;
            }
            const uint32_t b = odd ? b2 : b1;
            uint32_t st = 0;
            const Mask mine = page_walk<G>(lds, cl, i, page_row<G>(b >> 3), (ew >> (b - 7u)) & 0x7fu, st);
            const Mask other = __shfl_xor(mine, 1, 64);
            const uint32_t ost = static_cast<uint32_t>(__shfl_xor(static_cast<int>(st), 1, 64));
            // Event counts as the one-page-at-a-time order.
            good &= odd ? other : mine;
            ++walks;
            steps += odd ? ost : st;
            if (two && good != 0) {
                ++walks;
                steps += odd ? st : ost;
            }
            good &= odd ? mine : other;
        }
        return good;
    }
    if constexpr (walk_mode<G>() == 2) {
        // Pairs: both pages of the level in one loop.  Every cell i >= 1 of
        // a spiral map walks one or two pages (a DFS level is at least 1);
        // other maps run one page at a time, and a kernel staged with any
        // other level stops before its first walk.
        const uint32_t lw = reinterpret_cast<const uint32_t *>(lds + G::kLevelOff)[i];
        {
            const uint32_t a1 = lw & 0x7fu, a2 = (lw >> 8) & 0x7fu;
            const uint32_t r1 = G::kPageOff + (((lw >> 16) & 0xffu) << row_shift<G>()),
                           r2 = G::kPageOff + ((lw >> 24) << row_shift<G>());
            uint32_t j1 = cyc_ld<G>(lds, cl, a1), j2 = cyc_ld<G>(lds, cl, a2);
            Mask reach1 = 0, reach2 = 0;
            or_bit(reach1, a1);
            or_bit(reach2, a2);
            // A chain that closes walks its orbit again (back through the
            // anchor: the anchor's partner on this page is cell i, not yet
            // placed, so the walk re-enters the orbit), which only re-adds
            // cells already in reach.  Whether a chain is still open is a
            // lane mask that only ever loses lanes (scalar ANDs), so the loop
            // carries no per-step select.  A walk visits each cell of its
            // orbit once, so its step count is |reach| - 1.  With one page
            // the second chain repeats the first.
            bool o1 = j1 != a1, o2 = j2 != a2;
            // Waves inside a walk loop issue first: their next instruction is
            // usually waiting on an LDS load, so handing them the issue slot
            // as soon as it lands keeps more loads in flight.
            __builtin_amdgcn_s_setprio(kWalkPrio);
            do {
                // kWalkUnroll paired steps; each chain's new cells join its
                // reach two at a time in one three-input OR per dword.
                [[maybe_unused]] Mask bits1[kWalkUnroll], bits2[kWalkUnroll];
#pragma unroll
                for (int u = 0; u < kWalkUnroll; ++u) {
                    const uint32_t h1 = hop<G>(r1, i, j1);
                    const uint32_t h2 = hop<G>(r2, i, j2);
                    j1 = cyc_ld<G>(lds, cl, h1);
                    j2 = cyc_ld<G>(lds, cl, h2);
                    // The loaded bytes are already zero-extended: keep the
                    // compiler from narrowing the loop-carried cells to 8
                    // bits (which costs a mask per step).
                    asm("" : "+v"(j1));
                    asm("" : "+v"(j2));
                    if constexpr (sizeof(Mask) > 8) {
                        or_bit(reach1, h1);
                        or_bit(reach2, h2);
                    } else {
                        bits1[u] = static_cast<Mask>(1) << h1;
                        bits2[u] = static_cast<Mask>(1) << h2;
                    }
                    o1 = o1 && j1 != a1;
                    o2 = o2 && j2 != a2;
                }
                if constexpr (sizeof(Mask) == 8) {
#pragma unroll
                    for (int u = 0; u + 1 < kWalkUnroll; u += 2) {
                        reach1 = or3(reach1, bits1[u], bits1[u + 1]);
                        reach2 = or3(reach2, bits2[u], bits2[u + 1]);
                    }
                    if constexpr ((kWalkUnroll & 1) != 0) {
                        reach1 |= bits1[kWalkUnroll - 1];
                        reach2 |= bits2[kWalkUnroll - 1];
                    }
                }
            } while (o1 || o2);
            __builtin_amdgcn_s_setprio(0);
            // cells below i reached by both walks
            const Mask above = G::from(i);
            if (EVC) {
                // Event counts as the one-page-at-a-time order: the second
                // walk happens only if the first left good cells.
                ++walks;
                steps += mask_popcount(reach1) - 1u;
                if ((lw & 0x80u) != 0u && (reach1 & ~above) != 0) {
                    ++walks;
                    steps += mask_popcount(reach2) - 1u;
                }
            }
            return and_and_not(reach1, reach2, above);
        }
    }
    if constexpr (walk_mode<G>() == 1) {
        while (act != 0 && good != 0) {
            const uint32_t b = static_cast<uint32_t>(__builtin_ctz(act));
            act &= act - 1u;
            ++walks;
            good &= page_walk<G>(lds, cl, i, page_row<G>(b >> 3), (ew >> (b - 7u)) & 0x7fu, steps);
        }
        return good;
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const uint32_t anchor = (ew >> (8 * p)) & 0x7fu;
        if (anchor < i && good != 0) {
            ++walks;
            good &= page_walk<G>(lds, cl, i, page_row<G>(p), anchor, steps);
        }
    }
    return good;
}

// One node: expand level w.i (or recompute it above w.floor), then move to
// the next node of the subtree.  Returns true when the item is finished.
// Backtracking pops at most this many levels per step; a lane with more to
// pop skips the next step's walks and continues.  Bounding the pops keeps one
// lane's long unwind from holding the whole wave (a lane pops 1.0 levels per
// node on average, the deepest lane of a wave ~9).  Each popped level is one
// guarded region with two LDS reads and a write.  With the current pop passes
// a bound of 4 runs the 8x8 depth-36 chunk 1.3% faster than 3 (2: 5.4%
// slower, 5 and 6: 1.2% and 2.6% slower than 4), and 9x9 is fastest with 4
// too; the 7x7 chunk (49-cell rows) is 0.4% faster with 3.  MF_SYM_MAX_POPS
// > 0 overrides the per-geometry bound.
#ifndef MF_SYM_MAX_POPS
#define MF_SYM_MAX_POPS 0
#endif
template <typename G>
constexpr uint32_t max_pops() {
    return MF_SYM_MAX_POPS > 0 ? MF_SYM_MAX_POPS : (G::kRowCells <= 49 ? 3u : 4u);
}

template <typename G, bool EV>
__device__ __forceinline__ bool dfs_step(uint8_t *lds, const Lane &ln, uint32_t last, Work &w) {
    using Stack = typename G::Stack;
    constexpr uint32_t P = max_pops<G>();
    // The cycle entry and stack slot of level i - P: those of levels i - P .. i
    // sit at constant offsets from them, which fold into the LDS instructions.
    // (The bases may wrap below the lane's region; only the offsets of levels
    // the lane holds are dereferenced.)
    const uint32_t i0 = w.i;
    const uint32_t cb = cyc_addr<G>(ln.cl - P * G::kRowStride, i0);
    const uint32_t sb = (ln.s0 - P * Stack::kSlotBytes) + i0 * Stack::kSlotBytes;
    // The child this step enters -- the node's first child, or after a
    // backtrack the next sibling of the level it stops at -- goes in by one
    // shared block at the end: its level's new stack entry and its cell.
    bool enter = false, pop = w.popping;
    typename Stack::Entry ent = 0;
    uint32_t kid = 0;
    if (!w.popping) {
        uint32_t nw = 0, ns = 0;
        // A recompute (floor j + 1) keeps the children above j.
        const typename G::Mask g =
            good_mask<G, EV>(lds, ln.cl, i0, nw, ns) & G::from(w.floor);
        const uint32_t cnt = mask_popcount(g);   // children of this node
        if (w.floor == 0) {
            ++w.inodes;
            if (EV) {
                w.iwalks += nw;
                w.isteps += ns;
            }
        }
        w.floor = 0;
        // The last level counts its children as leaves (it is never
        // recomputed: its nodes enter no children).
        const bool leaf = i0 == last;
        if (leaf) w.leaves += cnt;
        enter = g != 0 && !leaf;
        pop = !enter;
        // (Computed by every expanding lane: with no child they are unused.)
        kid = mask_ctz_any(g);
        ent = Stack::encode_at(kid, g, cnt);
    }
    // Backtracking, one level per pass while the lane pops: remove the level's
    // child.  A level with more children (recorded or not) stops it, and so
    // does the item's root.  Each pass is one region with selects inside: a
    // lane at its root rewrites its own cycle entry unchanged instead of
    // branching.
    // Only `pop` and VGPR values cross the passes: whether the pops stopped
    // at a level with more children (ae != 0: such an entry is nonzero) and
    // whether they reached the root are read off afterwards.
    typename Stack::Entry ae = 0;
#pragma unroll
    for (uint32_t k = 0; k < P; ++k) {
        if (pop) {
            // level i0 - 1 - k
            const uint32_t slot = sb + (P - 1 - k) * Stack::kSlotBytes, cell = cb + (P - 1 - k) * G::kRowStride;
            const typename Stack::Entry e = Stack::get_at(lds, slot);
            const uint32_t c = cyc_ld_at<G>(lds, cell);
            const bool root = w.i == w.root;
            cyc_st<G>(lds, ln.cl, root ? w.i - 1 : Stack::child(e), c);   // unplace (at the root: unchanged)
            const bool more = !root && Stack::any_more(e);
            ae = more ? e : ae;
            w.i = root ? w.i : w.i - 1;
            pop = !(root || more);
        }
    }
    // A lane that popped and stopped without a level with more children
    // passed its root: the item is finished.
    const bool fin = !enter && !pop && ae == 0;
    // Stopped at a level with more children: its recorded next sibling, or a
    // recompute of the node above the current child.
    const bool nx = Stack::has_next(ae);
    if (Stack::recompute(ae)) w.floor = Stack::child(ae) + 1;
    if (nx) {
        ent = Stack::advance(ae);
        kid = Stack::child(ent);
    }
    if (enter || nx) {
        Stack::put(lds, ln, w.i, ent);
        place<G>(lds, ln, kid, w.i);
        ++w.i;
    }
    w.popping = pop;
    return fin;
}

// ------------------------------------------------------------ item plumbing

// Item kinds: mask != 0: enter the listed children of node `level`;
// mask == 0 and map == 0: expand node `level` (counted); mask == 0 and
// map == f > 0: recompute node `level` and enter its children above f-1.
template <typename G>
__device__ __forceinline__ void load_item_common(uint8_t *lds, const Lane &ln,
                                                 const typename G::Item *__restrict__ it, Work &w) {
    row_load<G::kStateBytes, G::kRowCells, G::kRowStride>(lds, ln.cl, it->state);
    w.origin = it->origin;
    w.mult = it->mult;
    w.root = it->level;
    w.leaves = 0;
    w.inodes = 0;
    w.iwalks = 0;
    w.isteps = 0;
    w.floor = 0;
    w.have = true;
    w.popping = false;
}

template <typename G>
__device__ __forceinline__ void load_item(uint8_t *lds, const Lane &ln, const typename G::Item *__restrict__ it,
                                          Work &w) {
    using Stack = typename G::Stack;
    load_item_common<G>(lds, ln, it, w);
    const typename G::Mask mask = G::item_mask(it);
    const uint32_t level = it->level;
    if (mask == 0) {
        w.i = level;
        w.floor = it->map;
    } else {
        // A sibling mask is always every child of the node above some child
        // (children are entered in ascending order), so an encoding that
        // records only part of it and recomputes the rest is exact.
        const typename Stack::Entry e = Stack::encode(mask);
        Stack::put(lds, ln, level, e);
        place<G>(lds, ln, Stack::child(e), level);
        w.i = level + 1;
    }
}

// Fold the current item's counters into the lane totals and the per-origin
// results (when an item finishes and when it is split).
__device__ __forceinline__ void retire(const SymArgs &a, const Work &w, Totals &t) {
    if (!w.primary) return;
    t.sum += static_cast<unsigned long long>(w.mult) * w.leaves;
    t.nodes += w.inodes;
    t.walks += w.iwalks;
    t.steps += w.isteps;
    t.leaves += w.leaves;
    t.items += 1;
    if (a.per_count != nullptr && w.leaves != 0) atomicAdd(a.per_count + w.origin, static_cast<unsigned long long>(w.leaves));
    if (a.per_nodes != nullptr && w.inodes != 0) atomicAdd(a.per_nodes + w.origin, static_cast<unsigned long long>(w.inodes));
}

// Fold the item's 32-bit event counters into the lane totals once one of
// them reaches kFoldAt (mf_counters.hpp): called after every step of the
// static schedule and every `budget` steps of the persistent one.
__device__ __forceinline__ void fold_item(const SymArgs &a, Work &w, Totals &t) {
    if (((w.inodes | w.iwalks | w.isteps) & kFoldAt) == 0u) return;
    FoldTotals f;
    uint32_t no_cands = 0;
    const uint32_t moved = fold_item_counters(w.inodes, w.iwalks, w.isteps, no_cands, f);
    if (!w.primary) return;   // the partner lane folded the same counts the same way
    t.nodes += f.nodes;
    t.walks += f.walks;
    t.steps += f.steps;
    if (a.per_nodes != nullptr && moved != 0) atomicAdd(a.per_nodes + w.origin, static_cast<unsigned long long>(moved));
}

__device__ __forceinline__ void flush_wave(const SymArgs &a, uint32_t lane, const Totals &t) {
    add_u128(a.acc, t.sum);
    const unsigned long long n = wave_sum(t.nodes), wk = wave_sum(t.walks), st = wave_sum(t.steps),
                             lv = wave_sum(t.leaves), it = wave_sum(t.items);
    if (lane == 0) {
        if (n) atomicAdd(a.acc + ACC_NODES, n);
        if (wk) atomicAdd(a.acc + ACC_WALKS, wk);
        if (st) atomicAdd(a.acc + ACC_STEPS, st);
        if (lv) atomicAdd(a.acc + ACC_LEAVES, lv);
        if (it) atomicAdd(a.acc + ACC_ITEMS, it);
    }
}

template <typename G>
__device__ __forceinline__ Lane make_lane(uint32_t levels, uint32_t base_level) {
    const uint32_t tid = threadIdx.x, slot = (tid & 63u) / G::kLanes, wave = tid >> 6;
    const uint32_t wave_bytes = G::kCycleBytes + levels * G::Stack::kSlotBytes;
    const uint32_t wbase = G::kTableBytes + wave * wave_bytes;
    Lane ln;
    ln.cl = row_lane_base(wbase, slot);
    ln.sl = wbase + G::kCycleBytes + slot * G::Stack::kLaneBytes;
    ln.base = base_level;
    ln.s0 = ln.sl - base_level * G::Stack::kSlotBytes;
    return ln;
}

// Returns false (for the whole block, with ACC_ERROR bit 4 raised) when the
// tables do not fit this geometry: the kernel then returns without walking.
template <typename G>
__device__ __forceinline__ bool stage_tables(const SymArgs &a, uint8_t *lds) {
    const uint32_t tid = threadIdx.x;
    if (tid < G::kCells) {
        // Partners are at most n2 <= 81 < 128: bit 7 of each byte is free.
        const uint32_t w = a.edge_word[tid];
        uint32_t flagged = w;
#pragma unroll
        for (uint32_t p = 0; p < 4; ++p) {
            lds[G::kPageOff + G::kCells * p + tid] = static_cast<uint8_t>(w >> (8u * p));
            flagged |= (((w >> (8u * p)) & 0xffu) < tid ? 0x80u : 0u) << (8u * p);
        }
        reinterpret_cast<uint32_t *>(lds + G::kPartnerOff)[tid] = flagged;
        reinterpret_cast<uint32_t *>(lds + G::kLevelOff)[tid] = level_word<G>(w, tid);
    }
    __syncthreads();
    // Checked by every wave on its own (a block-wide reduction would take
    // static LDS and move the dynamic block).  The walk addresses the page
    // table by absolute LDS offset.
    const auto *local = (const __attribute__((address_space(3))) uint8_t *)lds;
    bool bad = static_cast<uint32_t>(reinterpret_cast<size_t>(local)) != 0u;
    if constexpr (walk_mode<G>() == 2) {
        // The paired walk takes every level of the map from its level word:
        // a level that walks no page or more than two would send it round an
        // orbit that never returns to its anchor.  (Levels 1..last, one or
        // two per lane.)
        const uint32_t c = tid & 63u;
        const uint32_t *levels = reinterpret_cast<const uint32_t *>(lds + G::kLevelOff);
        bool generic = c >= 1u && c <= a.last && (levels[c] & kLevelGeneric) != 0u;
        if constexpr (G::kCells > 64) {
            if (c + 64u < G::kCells) generic = generic || (c + 64u <= a.last && (levels[c + 64u] & kLevelGeneric) != 0u);
        }
        bad = bad || __any(generic);
    }
    if (bad && tid == 0) atomicOr(a.acc + ACC_ERROR, 4ull);
    return !bad;
}

// Split everything the lane holds into the spill queue: the node about to be
// expanded, then for each stack level, deepest first, the siblings not yet
// entered (a recorded sibling as a mask, unrecorded ones as a recompute piece
// with a floor).  Called by the whole wave.  The wave first reserves all its
// slots with one compare-and-swap (spill_reserve); if the queue cannot take
// them, nothing is written and the wave keeps its work (returns false), so
// the queue never overflows whatever its capacity.
template <typename G>
__device__ __forceinline__ bool spill_all(const SymArgs &a, uint8_t *lds, const Lane &ln, uint32_t lane, Work &w,
                                          Totals &tot) {
    using Stack = typename G::Stack;
    using Item = typename G::Item;
    // Items this lane emits: the pending node (unless it is mid-backtrack,
    // when levels >= i are finished), then every level with siblings left.
    uint32_t need = 0;
    if (w.have && w.primary) {
        need = w.popping ? 0u : 1u;
        for (uint32_t t = w.root; t < w.i; ++t) need += Stack::more(Stack::get(lds, ln, t)) != 0u ? 1u : 0u;
    }
    // `off` is wave-uniform: the next free slot of the wave's reserved range.
    // Each pass writes the emitting lanes' items at consecutive slots in lane
    // order; the passes together write exactly `total` items.
    uint32_t off = spill_reserve(a.spill_count, a.spill_cap, need, lane);
    if (off == kNoSlot) return false;
    bool pend = w.have;
    uint32_t t = w.i;
    typename G::Mask mask = 0;
    uint32_t floor = w.floor;   // the pending expansion may itself be a recompute
    bool emit = pend && !w.popping && w.primary;
    const unsigned long long below_lane = (1ull << lane) - 1ull;
    while (__any(pend)) {
        const unsigned long long m = __ballot(emit);
        if (emit) {
            Item *dst = static_cast<Item *>(a.spill) + off + static_cast<uint32_t>(__popcll(m & below_lane));
            row_store<G::kStateBytes, G::kRowCells, G::kRowStride>(lds, ln.cl, dst->state);
            G::set_item_mask(dst, mask);
            dst->origin = w.origin;
            dst->mult = static_cast<uint16_t>(w.mult);
            dst->level = static_cast<uint8_t>(t);
            dst->map = static_cast<uint8_t>(floor);
        }
        off += static_cast<uint32_t>(__popcll(m));
        emit = false;
        if (pend) {
            if (t == w.root) {
                pend = false;
            } else {
                --t;
                const typename Stack::Entry e = Stack::get(lds, ln, t);
                const uint32_t j = Stack::child(e);
                unplace<G>(lds, ln, j, t);
                bool unknown = false;
                mask = Stack::rest(e, unknown);
                floor = 0;
                if (unknown) {
                    // Everything above j: recompute node t with floor j.
                    mask = 0;
                    floor = j + 1;
                }
                emit = (mask != 0 || floor != 0) && w.primary;
            }
        }
    }
    if (w.have) {
        retire(a, w, tot);
        w.have = false;
    }
    return true;
}

}  // namespace

// ---------------------------------------------------------------- static
template <typename G, bool EV>
__global__ __launch_bounds__(256) void sym_static_kernel(SymArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    if (!stage_tables<G>(a, lds)) return;
    const Lane ln = make_lane<G>(a.levels, a.base_level);
    const auto *items = static_cast<const typename G::Item *>(a.items);
    static_assert(G::kLanes == 1, "the static schedule runs one lane per item");
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    Totals tot;
    Work w;
    w.have = false;
    w.popping = false;
    w.primary = true;
    if (idx < a.num_items) load_item<G>(lds, ln, items + idx, w);
    while (w.have) {
        if (dfs_step<G, EV>(lds, ln, a.last, w)) {
            retire(a, w, tot);
            w.have = false;
        } else {
            fold_item(a, w, tot);
        }
    }
    flush_wave(a, threadIdx.x & 63u, tot);
}

// ------------------------------------------------------------- persistent
template <typename G, bool EV>
__device__ __forceinline__ void persistent_body(const SymArgs &a) {
    extern __shared__ __align__(16) uint8_t lds[];
    // Graph rounds read their item count from the previous round's spill
    // counter; an empty round returns at once (uniformly for the grid).
    const uint32_t num_items = a.num_items_dev ? *a.num_items_dev : a.num_items;
    if (num_items == 0) return;
    // A graph round's input is the previous round's spill queue, which has the
    // same capacity as this round's: a count past it means that round
    // overflowed (already flagged) -- read nothing.
    if (a.num_items_dev != nullptr && num_items > a.spill_cap) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(a.acc + ACC_ERROR, 1ull);
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(a.acc + ACC_ROUNDS, 1ull);
    if (!stage_tables<G>(a, lds)) return;
    const uint32_t lane = threadIdx.x & 63u;
    const Lane ln = make_lane<G>(a.levels, a.base_level);
    const auto *items = static_cast<const typename G::Item *>(a.items);
    const unsigned long long below = (1ull << lane) - 1ull;

    Totals tot;
    Work w;
    w.have = false;
    w.popping = false;
    w.primary = (lane % G::kLanes) == 0u;
    bool drained = false;   // wave-uniform: the input array has been handed out
    uint32_t tail_left = a.tail;

    for (;;) {
        if (!drained) {
            // Idle lanes claim the next input items, one atomic per wave.
            // (The lanes of a subtree are idle together: one claim per subtree.)
            const unsigned long long idle = __ballot(!w.have);
            if (idle != 0) {
                const uint32_t want = static_cast<uint32_t>(__popcll(idle)) / G::kLanes;
                const uint32_t first = static_cast<uint32_t>(__builtin_ctzll(idle));
                uint32_t got = 0;
                if (lane == first) got = atomicAdd(a.claim, want);
                got = __shfl(got, first, 64);
                if (got + want >= num_items) drained = true;
                if (!w.have) {
                    const uint32_t idx = got + static_cast<uint32_t>(__popcll(idle & below)) / G::kLanes;
                    if (idx < num_items) load_item<G>(lds, ln, items + idx, w);
                }
            } else {
                // Every lane is busy: look at the shared cursor anyway, so that
                // a wave of long items still learns that the input is gone.
                drained = __hip_atomic_load(a.claim, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= num_items;
            }
        }
        if (!__any(w.have)) break;

        for (uint32_t s = 0; s < a.budget && w.have; ++s) {
            if (dfs_step<G, EV>(lds, ln, a.last, w)) {
                retire(a, w, tot);
                w.have = false;
            }
        }
        if (w.have) fold_item(a, w, tot);

        if (drained) {
            tail_left = tail_left > a.budget ? tail_left - a.budget : 0;
            if (tail_left == 0) {
                // Progress guarantee: a split piece restarts the pending
                // expansion, so every lane first completes one counted node.
                while (__any(w.have && w.inodes == 0)) {
                    if (w.have && w.inodes == 0 && dfs_step<G, EV>(lds, ln, a.last, w)) {
                        retire(a, w, tot);
                        w.have = false;
                    }
                }
                if (spill_all<G>(a, lds, ln, lane, w, tot)) break;
                // The spill queue is full: the wave runs its work to the end
                // in this round, testing again every `budget` steps.
            }
        }
    }
    flush_wave(a, lane, tot);
}

template <typename G, bool EV>
__global__ __launch_bounds__(256) void sym_persistent_kernel(SymArgs a) {
    persistent_body<G, EV>(a);
}

// Up to 7x7 the rows are short enough that LDS holds 7 blocks per CU once
// the cases are deepened; the register budget is capped to match (at most
// 72 VGPRs, 7 waves per SIMD).
template <bool EV>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void sym_persistent_kernel49(SymArgs a) {
    persistent_body<G49, EV>(a);
}

// ------------------------------------------------------------- deepening
// One level of the frontier deepening that runs before the persistent rounds
// (u64 geometries): one lane per item.  An expand item below its target level
// is expanded -- one get_good, counted as a node exactly as the DFS counts it
// -- and each child (ascending insertion point, so siblings stay together)
// goes to the final array if it reaches the target, else to the next pass.
// Items at or past their target, and non-expand items, are copied to the
// final array unchanged (a non-expand item below the target cannot occur:
// the frontiers hold expand items only; it raises kErrDeepenFull).  Deeper
// items leave a shorter stack per lane in the persistent kernel, and more
// workgroups fit the LDS of a CU.
template <typename G, bool EV>
__global__ __launch_bounds__(256) void sym_deepen_kernel(SymArgs a) {
    using Mask = typename G::Mask;
    using Item = typename G::Item;
    extern __shared__ __align__(16) uint8_t lds[];
    if (!stage_tables<G>(a, lds)) return;
    const uint32_t lane = threadIdx.x & 63u;
    const Lane ln = make_lane<G>(1, 0);
    const auto *items = static_cast<const Item *>(a.items);
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    const bool have = idx < a.num_items;
    uint32_t level = 0, target = 0, origin = 0, mult = 0;
    bool expand = false;
    Mask good = 0;
    uint32_t nw = 0, ns = 0;
    if (have) {
        const Item *it = items + idx;
        level = it->level;
        origin = it->origin;
        mult = it->mult;
        const Mask m = G::item_mask(it);
        if (a.deep_pass == 0) {
            target = a.targets ? a.targets[idx] : a.target;
            expand = m == 0 && it->map == 0 && level < target;
            if ((m != 0 || it->map != 0) && level < target) atomicOr(a.acc + ACC_ERROR, kErrDeepenFull);
        } else {
            target = static_cast<uint32_t>(m);
            expand = level < target;
        }
        if (expand) {
            row_load<G::kStateBytes, G::kRowCells, G::kRowStride>(lds, ln.cl, it->state);
            good = good_mask<G, EV>(lds, ln.cl, level, nw, ns);
        }
    }
    // Slots: the children of an expanded item, or the item itself.
    const uint32_t cnt = expand ? mask_popcount(good) : 0u;
    const bool to_final = !expand || level + 1 >= target;
    const uint32_t nf = have ? (expand ? (to_final ? cnt : 0u) : 1u) : 0u;
    const uint32_t nc = (have && expand && !to_final) ? cnt : 0u;
    const uint32_t fslot = lane_reserve(a.final_count, a.final_cap, nf, lane);
    const uint32_t cslot = lane_reserve(a.spill_count, a.spill_cap, nc, lane);
    if (fslot == kNoSlot || cslot == kNoSlot) {
        if (lane == 0) atomicOr(a.acc + ACC_ERROR, kErrDeepenFull);
        return;
    }
    if (have && !expand) {
        static_cast<Item *>(a.final_out)[fslot] = items[idx];
        if (a.deep_pass != 0) G::set_item_mask(static_cast<Item *>(a.final_out) + fslot, 0);
    }
    unsigned long long nodes = 0, walks = 0, steps = 0;
    if (expand) {
        nodes = 1;
        if (EV) {
            walks = nw;
            steps = ns;
        }
        if (a.per_nodes != nullptr) atomicAdd(a.per_nodes + origin, 1ull);
        Item *dst = to_final ? static_cast<Item *>(a.final_out) + fslot : static_cast<Item *>(a.spill) + cslot;
        const Mask carry = to_final ? static_cast<Mask>(0) : static_cast<Mask>(target);
        for (Mask gm = good; gm != 0; gm &= gm - 1, ++dst) {
            const uint32_t j = mask_ctz(gm);
            place<G>(lds, ln, j, level);
            row_store<G::kStateBytes, G::kRowCells, G::kRowStride>(lds, ln.cl, dst->state);
            unplace<G>(lds, ln, j, level);
            G::set_item_mask(dst, carry);
            dst->origin = origin;
            dst->mult = static_cast<uint16_t>(mult);
            dst->level = static_cast<uint8_t>(level + 1);
            dst->map = 0;
        }
    }
    const unsigned long long n = wave_sum(nodes), wk = wave_sum(walks), st = wave_sum(steps);
    if (lane == 0) {
        if (n) atomicAdd(a.acc + ACC_NODES, n);
        if (wk) atomicAdd(a.acc + ACC_WALKS, wk);
        if (st) atomicAdd(a.acc + ACC_STEPS, st);
    }
}

// ------------------------------------------------------------- cost features
// The learned cost model's feature row of every case of a resident frontier,
// computed on the device from the case's own cycle -- the same 80 values as
// the host's sym_row (cost.cpp), which the model is trained on:
//   [0, 64)   placed cell c: (rank of c along the cycle from cell 0 + 1) / 64;
//             unplaced cell c < n*n: popcount of the intersection of the
//             page walks from its placed partners (cells below the level) /
//             level; 0 past n*n;
//   64..71    level / 64, level / n*n, children / 16, n*n / 64, 1 (sym),
//             constrained pages / 4, levels left / 64, n / 8;
//   [72, 80)  log2(1 + descendants at level + k) / 24, k = 1..8: a DFS to
//             depth 8 (estimate_work of sym_fold.cpp:235-250 at each depth).
// One lane per case.  LDS: the partner tables, each lane's cycle row, and per
// lane the DFS's remaining-children masks and current children (8 levels).
constexpr uint32_t kFeatureDepth = 8;
constexpr uint32_t kFeatureStackLevels = 36;   // 36 x 128 B per wave: 8 u64 masks + 8 child bytes per lane

template <typename G>
__global__ __launch_bounds__(256) void sym_features_kernel(SymArgs a, float *out) {
    using Mask = typename G::Mask;
    using Item = typename G::Item;
    extern __shared__ __align__(16) uint8_t lds[];
    if (!stage_tables<G>(a, lds)) return;
    const uint32_t lane = threadIdx.x & 63u;
    const Lane ln = make_lane<G>(kFeatureStackLevels, 0);
    const uint32_t sbase = ln.sl - lane * G::Stack::kLaneBytes;   // this wave's stack region
    unsigned long long *rem = reinterpret_cast<unsigned long long *>(lds + sbase) + lane;   // [k * 64]
    uint8_t *cur = lds + sbase + kFeatureDepth * 512u + lane;                            // [k * 64]
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= a.num_items) return;
    const Item *it = static_cast<const Item *>(a.items) + idx;
    const uint32_t level = it->level;
    const uint32_t n2 = a.last + 1;
    row_load<G::kStateBytes, G::kRowCells, G::kRowStride>(lds, ln.cl, it->state);
    float *f = out + static_cast<size_t>(idx) * 80u;
    // Placed cells: rank along the cycle from cell 0.
    uint32_t c = 0;
    for (uint32_t r = 0; r < level; ++r) {
        f[c] = static_cast<float>(r + 1) / 64.0f;
        c = cyc_ld<G>(lds, ln.cl, c);
        if (c == 0) break;
    }
    // Unplaced cells: their options against the current stack.
    uint32_t dummy = 0;
    for (uint32_t cell = level; cell < G::kCells; ++cell) {
        float v = 0.0f;
        if (cell < n2) {
            Mask opts = G::below(level);
            for (uint32_t p = 0; p < 4 && opts != 0; ++p) {
                const uint32_t anchor = lds[G::kPageOff + G::kCells * p + cell];
                if (anchor < level) opts &= page_walk<G>(lds, ln.cl, level, page_row<G>(p), anchor, dummy);
            }
            v = static_cast<float>(mask_popcount(opts)) / static_cast<float>(level);
        }
        f[cell] = v;
    }
    // Lookahead profile: descendants at depths 1..8 (one iterative DFS).
    unsigned long long prof[kFeatureDepth] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t k = 0, i = level;
    uint32_t nw = 0, ns = 0;
    for (;;) {
        bool descend = false;
        if (i < n2) {
            const Mask g = good_mask<G, false>(lds, ln.cl, i, nw, ns);
            const uint32_t cnt = mask_popcount(g);
#pragma unroll
            for (uint32_t q = 0; q < kFeatureDepth; ++q) prof[q] += (q == k) ? cnt : 0u;
            if (k + 1 < kFeatureDepth && i + 1 < n2 && g != 0) {
                const uint32_t j = mask_ctz(g);
                rem[k * 64u] = static_cast<unsigned long long>(g & (g - 1));
                cur[k * 64u] = static_cast<uint8_t>(j);
                place<G>(lds, ln, j, i);
                ++k;
                ++i;
                descend = true;
            }
        }
        if (descend) continue;
        bool resumed = false;
        while (k > 0) {
            --k;
            --i;
            unplace<G>(lds, ln, cur[k * 64u], i);
            unsigned long long r = rem[k * 64u];
            if (r != 0) {
                const uint32_t j = static_cast<uint32_t>(__builtin_ctzll(r));
                rem[k * 64u] = r & (r - 1);
                cur[k * 64u] = static_cast<uint8_t>(j);
                place<G>(lds, ln, j, i);
                ++k;
                ++i;
                resumed = true;
                break;
            }
        }
        if (!resumed) break;
    }
    uint32_t pages = 0;
    for (uint32_t p = 0; p < 4; ++p)
        pages += (level < n2 && lds[G::kPageOff + G::kCells * p + level] < level) ? 1u : 0u;
    f[64] = static_cast<float>(level) / 64.0f;
    f[65] = static_cast<float>(level) / static_cast<float>(n2);
    f[66] = static_cast<float>(prof[0]) / 16.0f;
    f[67] = static_cast<float>(n2) / 64.0f;
    f[68] = 1.0f;
    f[69] = static_cast<float>(pages) / 4.0f;
    f[70] = static_cast<float>(n2 - level) / 64.0f;
    const uint32_t n = static_cast<uint32_t>(sqrtf(static_cast<float>(n2)) + 0.5f);   // side of the square
    f[71] = static_cast<float>(n) / 8.0f;
#pragma unroll
    for (uint32_t q = 0; q < kFeatureDepth; ++q)
        f[72 + q] = static_cast<float>(log2(1.0 + static_cast<double>(prof[q]))) / 24.0f;
}

namespace {
template <typename G>
hipError_t launch_features(const SymArgs &a, float *out, hipStream_t s) {
    const uint32_t lds = G::kTableBytes + 4u * (G::kCycleBytes + kFeatureStackLevels * G::Stack::kSlotBytes);
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&sym_features_kernel<G>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(sym_features_kernel<G>, dim3((a.num_items + 255u) / 256u), dim3(256), lds, s, a, out);
    return hipGetLastError();
}
}  // namespace

// The core that runs `core`'s request for the map in `a` (`cells` = n*n):
// the paired walk (cores 0-2, and 9x9) needs every level 1..last to walk one
// or two pages, which every spiral numbering gives; a map with any other
// level, or a request for core 4, runs one page walk at a time: core 4 up to
// 8x8 (compact stack), core 5 for 9x9.  Core 3's two-lane walk already walks
// page by page.
int sym_walk_core(int core, const SymArgs &a, uint32_t cells) {
    bool paired = core != 4;
    for (uint32_t i = 1; i <= a.last && i < 96u; ++i)
        if ((level_word<G96>(a.edge_word[i], i) & kLevelGeneric) != 0u) paired = false;
    if (cells > 64) return paired ? 2 : 5;
    if (core == 3) return core;
    return paired ? core : 4;
}

hipError_t launch_sym_features(const SymArgs &a, float *out, hipStream_t s) {
    // Rows of 64 cells for every square up to 8x8.
    return sym_walk_core(0, a, 64) == 4 ? launch_features<G64Any>(a, out, s) : launch_features<G64>(a, out, s);
}

// Host-visible entry points (C++ linkage inside the library).
namespace {
template <typename G>
const void *pick(bool persistent, bool events) {
    if (persistent) return events ? reinterpret_cast<const void *>(&sym_persistent_kernel<G, true>)
                                  : reinterpret_cast<const void *>(&sym_persistent_kernel<G, false>);
    return events ? reinterpret_cast<const void *>(&sym_static_kernel<G, true>)
                  : reinterpret_cast<const void *>(&sym_static_kernel<G, false>);
}
}  // namespace

// core: 1 = wide u64 stack (up to 8x8), 3 = two lanes per subtree, 4 / 5 =
// one page walk at a time (sym_walk_core), otherwise the compact 16-bit stack;
// cells = n*n picks the row size (49, 64 or, for 9x9, 81 entries).
const void *sym_kernel_ptr(int core, bool persistent, bool events, uint32_t cells) {
    switch (geometry(core, cells)) {
        case Geo::k64Pair:   // re-split schedule only
            if (!persistent) return nullptr;
            return events ? reinterpret_cast<const void *>(&sym_persistent_kernel<G64Pair, true>)
                          : reinterpret_cast<const void *>(&sym_persistent_kernel<G64Pair, false>);
        case Geo::k96: return pick<G96>(persistent, events);
        case Geo::k96Any: return pick<G96Any>(persistent, events);
        case Geo::k64Wide: return pick<G64Wide>(persistent, events);
        case Geo::k64Any: return pick<G64Any>(persistent, events);
        case Geo::k49:
            if (persistent)
                return events ? reinterpret_cast<const void *>(&sym_persistent_kernel49<true>)
                              : reinterpret_cast<const void *>(&sym_persistent_kernel49<false>);
            return pick<G49>(false, events);
        default: return pick<G64>(persistent, events);
    }
}

// Between graph rounds: rewind the claim cursor and empty the next spill
// queue (a kernel, not a memset node, so that replays stay exact).
__global__ void round_reset_kernel(uint32_t *claim, uint32_t *out_count) {
    if (threadIdx.x == 0) {
        *claim = 0u;
        *out_count = 0u;
    }
}

hipError_t launch_round_reset(uint32_t *claim, uint32_t *out_count, hipStream_t s) {
    hipLaunchKernelGGL(round_reset_kernel, dim3(1), dim3(64), 0, s, claim, out_count);
    return hipGetLastError();
}

hipError_t launch_sym(const SymArgs &a, int core, bool persistent, uint32_t grid, uint32_t block,
                      uint32_t lds_bytes, bool events, uint32_t cells, hipStream_t s) {
    void *args[] = {const_cast<SymArgs *>(&a)};
    return hipLaunchKernel(sym_kernel_ptr(core, persistent, events, cells), dim3(grid), dim3(block), args,
                           lds_bytes, s);
}

namespace {
template <typename G>
uint32_t lds_of(uint32_t levels, uint32_t block) {
    return G::kTableBytes + (block / 64u) * (G::kCycleBytes + levels * G::Stack::kSlotBytes);
}
}  // namespace

// Deepening pass for the u64 geometries (nullptr for 9x9 and the wide stack).
const void *sym_deepen_ptr(int core, bool events, uint32_t cells) {
    switch (geometry(core, cells)) {
        case Geo::k49: return events ? reinterpret_cast<const void *>(&sym_deepen_kernel<G49, true>)
                                     : reinterpret_cast<const void *>(&sym_deepen_kernel<G49, false>);
        case Geo::k64: return events ? reinterpret_cast<const void *>(&sym_deepen_kernel<G64, true>)
                                     : reinterpret_cast<const void *>(&sym_deepen_kernel<G64, false>);
        case Geo::k64Any: return events ? reinterpret_cast<const void *>(&sym_deepen_kernel<G64Any, true>)
                                        : reinterpret_cast<const void *>(&sym_deepen_kernel<G64Any, false>);
        default: return nullptr;
    }
}

uint32_t sym_lds_bytes(int core, uint32_t levels, uint32_t block, uint32_t cells) {
    switch (geometry(core, cells)) {
        case Geo::k64Pair: return lds_of<G64Pair>(levels, block);
        case Geo::k96: return lds_of<G96>(levels, block);
        case Geo::k96Any: return lds_of<G96Any>(levels, block);
        case Geo::k64Wide: return lds_of<G64Wide>(levels, block);
        case Geo::k64Any: return lds_of<G64Any>(levels, block);
        case Geo::k49: return lds_of<G49>(levels, block);
        default: return lds_of<G64>(levels, block);
    }
}

}  // namespace mf
