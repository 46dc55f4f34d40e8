// lunnon_kernels.hip -- Lunnon map-folding DFS on gfx950.
//
// Restates the resume loop of mf.c:78-143 / solve_state_device
// (gpu_folding_v2.cu:381-433) per lane:
//   * state = the "below" links b[0..l-1] of the placed leaves (a circular
//     list rooted at leaf 0), one byte row per lane in LDS; a[] is never
//     needed because the leaf a new leaf was linked below is the stack's
//     current child;
//   * the gap computation of leaf l (mf.c:88-124) walks each dimension's
//     chain m -> d[i][l][b[m]] and keeps the positions reached by every
//     dimension that constrains l -- as u64 masks instead of a gap stack and
//     a per-leaf tally (positions are < l <= 64); the two dimensions of a
//     rectangle walk in one loop (gap_mask);
//   * every legal position is one insert (node).  An insert at position 0 of
//     a leaf l >= 2 breaks the normal-form filter b[0] == 1 (mf.c:83) and is
//     pruned at once, so it is counted but never entered; an insert of leaf n
//     is a completion worth +n (mf.c:84-85, process() :9-11);
//   * the stack uses the encodings of mf_dfs.hpp (compact by default: a
//     level whose unrecorded positions are needed recomputes its gaps,
//     uncounted, above the current one);
//   * the d tables of every map in the launch sit in LDS as bytes, lower
//     triangles only (d[i][l][m] for m <= l), one copy per distinct shape.
// Child order differs from the reference's gap-stack order; subtree counts,
// inserts and completions do not depend on it.
#include <hip/hip_runtime.h>

#include "mf_counters.hpp"
#include "mf_device.hpp"
#include "mf_dfs.hpp"
#include "mf_wave.hpp"

namespace mf {
namespace {

constexpr uint32_t kMapBytes = kMaxMaps * sizeof(LunnonMap);

struct LWork {
    uint32_t l;       // next leaf to insert (the node being expanded)
    uint32_t root;
    uint32_t n;
    uint32_t ndim;
    uint32_t tab;     // LDS byte address of the map's d table
    uint32_t tri;     // bytes per dimension of the triangular table: (n+1)(n+2)/2
    uint32_t origin;
    uint32_t map;
    uint32_t floor;   // 0, or j+1: recompute the gaps of leaf l, keep those above j
    uint64_t comps;   // completions of the current item
    uint32_t inodes;  // inserts, gap computations, walk steps, candidates of the item
    uint32_t iwalks;  // (32-bit: folded into the lane totals before they can wrap, lfold_item)
    uint32_t isteps;
    uint32_t icands;
    bool have;
    bool popping;     // backtracking continues this step (no gap computation)
    bool progressed;  // a counted gap computation of this item has completed
};

struct LTotals {
    unsigned long long sum = 0;   // sum of n * completions
    unsigned long long nodes = 0, walks = 0, steps = 0, cands = 0, comps = 0, items = 0;
};

// Byte stride between entries of a lane's link row: 73, so that lanes of a
// quad reading different leaves land in different banks.  Against 64 it runs
// the 6x6 depth-18 frontier 2.9% and the 64 rectangles 1.8% faster (65: 0.6%
// and 1.0%; 77 and 81 the same as 73, 85 slower), and the blocks per CU stay
// what the registers allow.  The rows are then addressed by absolute LDS
// address, as the kernels' only LDS is the dynamic block at offset 0.
#ifndef MF_LUNNON_ROW_STRIDE
#define MF_LUNNON_ROW_STRIDE 73
#endif
constexpr uint32_t kLStride = MF_LUNNON_ROW_STRIDE;

// Link leaf l directly below leaf m (mf.c:137-140 without a[]).
__device__ __forceinline__ void link(uint8_t *lds, const Lane &ln, uint32_t m, uint32_t l) {
    row_st<kLStride>(lds, row_at<kLStride>(ln.cl, l), row_ld<kLStride>(lds, row_at<kLStride>(ln.cl, m)));
    row_st<kLStride>(lds, row_at<kLStride>(ln.cl, m), l);
}
// Unlink leaf l that sits below leaf m (mf.c:131-132).
__device__ __forceinline__ void unlink(uint8_t *lds, const Lane &ln, uint32_t m, uint32_t l) {
    row_st<kLStride>(lds, row_at<kLStride>(ln.cl, m), row_ld<kLStride>(lds, row_at<kLStride>(ln.cl, l)));
}

// Legal positions for leaf l: mask of m < l reached by the walk of every
// dimension with d[i][l][l] != l (all m < l if none constrains l).  Every state
// here is one the folding DFS reaches (inputs are checked on the host by
// replaying their insertions, lunnon_reachable; split pieces are states of
// this DFS), and on those each walk returns to l within l steps, as in mf.c.
// The loop carries no step guard.  A guard's second loop exit made the
// compiler track each lane's exit reason: 22 scalar and branch instructions
// per step against 5 now (8 and 6 VALU), and the 6x6 depth-18 frontier ran in
// 217 ms against 142 ms without it -- the walk was bound by scalar issue.
// Wave priority inside the paired dimension walk, as in the symmetry DFS:
// a wave whose next instruction waits on a walk's LDS load gets the issue
// slot as soon as it lands (1: 1.7% less time on the 6x6 depth-18 frontier
// and 1.6% on the 64 rectangles than 0).
#ifndef MF_LUNNON_WALK_PRIO
#define MF_LUNNON_WALK_PRIO 1
#endif
constexpr int kLWalkPrio = MF_LUNNON_WALK_PRIO;

// How a gap computation walks the dimensions: kDimsAny, each lane's map
// gives them, walked one after the other; kDims2, every map of the launch is
// two-dimensional (lunnon_dims), the two walked one after the other
// (unrolled); kDims2Paired, the same maps with at most 63 leaves, both
// dimensions walked in one loop.
enum : uint32_t { kDimsAny = 0, kDims2 = 1, kDims2Paired = 2 };
template <uint32_t D>
__device__ __forceinline__ uint64_t gap_mask(uint8_t *__restrict__ lds, const Lane &ln, LWork &w,
                                             uint32_t &steps, uint32_t &cands) {
    const uint32_t l = w.l;
    const uint64_t all = below_mask(l);
    if constexpr (D == kDims2Paired) {
        const uint32_t r0 = w.tab + ((l * (l + 1u)) >> 1), r1 = r0 + w.tri;
        uint32_t m0 = lds[r0 + l], m1 = lds[r1 + l];
        const bool c0 = m0 != l, c1 = m1 != l;
        // Leaf l is not linked yet, so its own link entry (row entry l <= 63)
        // is free: while its gaps are computed it points at l.  A walk back
        // at l then steps to d[i][l][l], its first leaf, and walks its orbit
        // again, which adds nothing to its reach but bit l (outside `all`);
        // an unconstrained dimension stays at l.  So a closed walk needs no
        // mask, and the two walks' LDS loads are in flight together.
        row_st<kLStride>(lds, row_at<kLStride>(ln.cl, l), l);
        uint64_t reach0 = 0, reach1 = 0;
        bool o0 = c0, o1 = c1;
        // Two steps per trip: a closed walk's extra step changes nothing,
        // so the exit is tested every second step.
        if constexpr (kLWalkPrio > 0) __builtin_amdgcn_s_setprio(kLWalkPrio);
        do {
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                reach0 |= 1ull << m0;
                reach1 |= 1ull << m1;
                m0 = lds[r0 + row_ld<kLStride>(lds, row_at<kLStride>(ln.cl, m0))];
                m1 = lds[r1 + row_ld<kLStride>(lds, row_at<kLStride>(ln.cl, m1))];
                o0 = o0 && m0 != l;
                o1 = o1 && m1 != l;
            }
        } while (o0 || o1);
        if constexpr (kLWalkPrio > 0) __builtin_amdgcn_s_setprio(0);
        // A walk visits each leaf of its orbit once: its step count is the
        // orbit's size (leaves below l).
        reach0 = c0 ? reach0 & all : 0;
        reach1 = c1 ? reach1 & all : 0;
        steps += mask_popcount(reach0) + mask_popcount(reach1);
        const bool constrained = c0 || c1;
        cands += constrained ? mask_popcount(reach0 | reach1) : l;
        return (c0 ? reach0 : all) & (c1 ? reach1 : all);
    }
    uint64_t meet = all, reach_any = 0;
    bool constrained = false;
    // Row l of dimension i in the triangular table: entries m <= l (the walk
    // only looks up l itself and links of placed leaves, all < l).
    const uint32_t row0 = w.tab + ((l * (l + 1u)) >> 1);
    auto dimension = [&](uint32_t i) {
        const uint32_t row = row0 + i * w.tri;
        uint32_t m = lds[row + l];
        if (m == l) return;
        constrained = true;
        uint64_t reach = 0;
        do {
            ++steps;
            reach |= 1ull << m;
            m = lds[row + row_ld<kLStride>(lds, row_at<kLStride>(ln.cl, m))];
        } while (m != l);
        meet &= reach;
        reach_any |= reach;
    };
    if constexpr (D == kDims2) {
        dimension(0);
        dimension(1);
    } else {
        for (uint32_t i = 0; i < w.ndim; ++i) dimension(i);
    }
    cands += constrained ? static_cast<uint32_t>(__popcll(reach_any)) : l;
    return constrained ? meet : all;
}

// Backtracking pops at most this many levels per step (as in the symmetry
// DFS): a lane with more to pop skips the next step's gap computation.  With
// the paired dimension walk, 3 runs the 6x6 depth-18 frontier 0.8% and the
// 64 rectangles 1.6% faster than 4 (5: 0.4% / 0.9% slower than 4).
#ifndef MF_LUNNON_MAX_POPS
#define MF_LUNNON_MAX_POPS 3
#endif
constexpr uint32_t kLMaxPops = MF_LUNNON_MAX_POPS;

template <typename Stack, bool EV, uint32_t D>
__device__ __forceinline__ bool ldfs_step(uint8_t *lds, const Lane &ln, LWork &w) {
    constexpr uint32_t P = kLMaxPops;
    // The link entry and stack slot of leaf l - P: those of leaves l - P .. l
    // sit at constant offsets from them, which fold into the LDS instructions.
    // (The bases may point below the lane's region; a lane at its root reads
    // the entries of leaf root - 1, inside the wave's rows, and rewrites its
    // link entry unchanged.)
    const uint32_t l0 = w.l;
    uint32_t rb = row_at<kLStride>(ln.cl - P * kLStride, l0);
    uint32_t sb = (ln.s0 - P * Stack::kSlotBytes) + l0 * Stack::kSlotBytes;
    // Opaque to the compiler, which would otherwise re-derive every pass's
    // addresses from l0 instead of folding the offsets.
    asm("" : "+v"(rb));
    asm("" : "+v"(sb));
    // The child this step enters -- the node's first position, or after a
    // backtrack the next recorded sibling -- goes in by one shared block at
    // the end: its level's stack entry and its link.
    bool enter = false, pop = w.popping;
    typename Stack::Entry ent = 0;
    uint32_t kid = 0;
    if (!w.popping) {
        uint32_t ns = 0, nc = 0;
        uint64_t g = gap_mask<D>(lds, ln, w, ns, nc);
        if (w.floor == 0) {
            w.progressed = true;
            w.inodes += static_cast<uint32_t>(__popcll(g));
            if (EV) {
                ++w.iwalks;
                w.isteps += ns;
                w.icands += nc;
            }
        }
        // Position 0 is pruned by the normal-form filter (l >= 2); a
        // recompute keeps the positions above its floor j + 1 >= 1.
        g &= ~0ull << (w.floor == 0 ? 1u : w.floor);
        w.floor = 0;
        // Leaf n's positions are completions (a recompute never reaches n:
        // its inserts enter no children).
        const bool leaf = l0 == w.n;
        if (leaf) w.comps += static_cast<uint32_t>(__popcll(g));
        enter = g != 0 && !leaf;
        pop = !enter;
        if (enter) {
            kid = mask_ctz(g);
            ent = Stack::encode_at(kid, g, mask_popcount(g));
        }
    }
    // Backtracking, one leaf per pass while the lane pops: unlink it from
    // below its position.  A level with more positions (recorded or not)
    // stops it, and so does the item's root.  Each pass is one region with
    // selects inside.
    // Only `pop` and VGPR values cross the passes (as in the symmetry DFS).
    typename Stack::Entry ae = 0;
#pragma unroll
    for (uint32_t k = 0; k < P; ++k) {
        if (pop) {
            // leaf l0 - 1 - k
            const uint32_t slot = sb + (P - 1 - k) * Stack::kSlotBytes, at = rb + (P - 1 - k) * kLStride;
            const typename Stack::Entry e = Stack::get_at(lds, slot);
            const uint32_t c = row_ld<kLStride>(lds, at);
            const bool root = w.l == w.root;
            // unlink (mf.c:131-132): b[j] = b[l]; at the root b[l] unchanged
            row_st<kLStride>(lds, root ? at : row_at<kLStride>(ln.cl, Stack::child(e)), c);
            const bool more = !root && Stack::any_more(e);
            ae = more ? e : ae;
            w.l = root ? w.l : w.l - 1;
            pop = !(root || more);
        }
    }
    // Popped and stopped without a level with more positions: past the root.
    const bool fin = !enter && !pop && ae == 0;
    // Stopped at a level with more positions: its recorded next sibling, or
    // a recompute of the leaf's gaps above the current position.
    const bool nx = Stack::has_next(ae);
    if (Stack::recompute(ae)) w.floor = Stack::child(ae) + 1;
    if (nx) {
        ent = Stack::advance(ae);
        kid = Stack::child(ent);
    }
    if (enter || nx) {
        Stack::put(lds, ln, w.l, ent);
        link(lds, ln, kid, w.l);
        ++w.l;
    }
    w.popping = pop;
    return fin;
}

// Item kinds as for the symmetry DFS; the floor of a recompute piece travels
// in the upper byte of `mult` (the low byte is n).
template <typename Stack>
__device__ __forceinline__ void lload_item(uint8_t *lds, Lane &ln, const WorkItem *__restrict__ it,
                                           const LunnonMap *maps, uint32_t rows, LWork &w) {
    row_load_n<kLStride>(lds, ln.cl, it->state, rows);
    const LunnonMap mp = maps[it->map];
    w.n = mp.n;
    w.ndim = mp.ndim;
    w.tab = mp.table_off;
    w.tri = ((mp.n + 1u) * (mp.n + 2u)) >> 1;
    ln.base = mp.base;
    ln.s0 = ln.sl - mp.base * Stack::kSlotBytes;
    w.map = it->map;
    w.origin = it->origin;
    const uint32_t level = it->level;
    const uint64_t mask = it->mask;
    w.root = level;
    w.comps = 0;
    w.inodes = 0;
    w.iwalks = 0;
    w.isteps = 0;
    w.icands = 0;
    w.floor = 0;
    w.have = true;
    w.popping = false;
    w.progressed = false;
    if (mask == 0) {
        w.l = level;
        w.floor = it->mult >> 8;
    } else {
        const typename Stack::Entry e = Stack::encode(mask);
        Stack::put(lds, ln, level, e);
        link(lds, ln, Stack::child(e), level);
        w.l = level + 1;
    }
}

__device__ __forceinline__ void lretire(const LunnonArgs &a, const LWork &w, LTotals &t) {
    const unsigned long long v = static_cast<unsigned long long>(w.n) * w.comps;
    t.sum += v;
    t.nodes += w.inodes;
    t.walks += w.iwalks;
    t.steps += w.isteps;
    t.cands += w.icands;
    t.comps += w.comps;
    t.items += 1;
    if (a.per_count != nullptr && w.comps != 0) atomicAdd(a.per_count + w.origin, static_cast<unsigned long long>(w.comps));
    if (a.per_nodes != nullptr && w.inodes != 0) atomicAdd(a.per_nodes + w.origin, static_cast<unsigned long long>(w.inodes));
    if (a.per_map != nullptr && v != 0) atomicAdd(a.per_map + w.map, v);
}

// Fold the item's 32-bit event counters into the lane totals once one of
// them reaches kFoldAt (mf_counters.hpp).
__device__ __forceinline__ void lfold_item(const LunnonArgs &a, LWork &w, LTotals &t) {
    if (((w.inodes | w.iwalks | w.isteps | w.icands) & kFoldAt) == 0u) return;
    FoldTotals f;
    const uint32_t moved = fold_item_counters(w.inodes, w.iwalks, w.isteps, w.icands, f);
    t.nodes += f.nodes;
    t.walks += f.walks;
    t.steps += f.steps;
    t.cands += f.cands;
    if (a.per_nodes != nullptr && moved != 0) atomicAdd(a.per_nodes + w.origin, static_cast<unsigned long long>(moved));
}

__device__ __forceinline__ void lflush(const LunnonArgs &a, uint32_t lane, const LTotals &t) {
    add_u128(a.acc, t.sum);
    const unsigned long long n = wave_sum(t.nodes), wk = wave_sum(t.walks), st = wave_sum(t.steps),
                             cd = wave_sum(t.cands), cp = wave_sum(t.comps), it = wave_sum(t.items);
    if (lane == 0) {
        if (n) atomicAdd(a.acc + ACC_NODES, n);
        if (wk) atomicAdd(a.acc + ACC_WALKS, wk);
        if (st) atomicAdd(a.acc + ACC_STEPS, st);
        if (cd) atomicAdd(a.acc + ACC_CANDS, cd);
        if (cp) atomicAdd(a.acc + ACC_LEAVES, cp);
        if (it) atomicAdd(a.acc + ACC_ITEMS, it);
    }
}

// LDS: [tables][maps][per wave: b rows of row_entries entries | stack].
__device__ __forceinline__ uint32_t tables_span(uint32_t bytes) { return (bytes + 15u) & ~15u; }

template <typename Stack>
__device__ __forceinline__ Lane lmake_lane(uint32_t tab_span, uint32_t levels, uint32_t rows) {
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t row_bytes = (rows * kLStride + 15u) & ~15u;   // the stack after it stays 16-byte aligned
    const uint32_t wbase = tab_span + kMapBytes + wave * (row_bytes + levels * Stack::kSlotBytes);
    Lane ln;
    ln.cl = row_lane_base(wbase, lane);
    ln.sl = wbase + row_bytes + lane * Stack::kLaneBytes;
    ln.base = 0;
    ln.s0 = ln.sl;
    return ln;
}

__device__ __forceinline__ void lstage(const LunnonArgs &a, uint8_t *lds, uint32_t tab_span) {
    const uint32_t words = (a.tables_bytes + 3u) / 4u;
    for (uint32_t k = threadIdx.x; k < words; k += blockDim.x) {
        uint32_t v = 0;
        for (uint32_t b = 0; b < 4u; ++b) {
            const uint32_t at = 4u * k + b;
            if (at < a.tables_bytes) v |= static_cast<uint32_t>(a.tables[at]) << (8u * b);
        }
        reinterpret_cast<uint32_t *>(lds)[k] = v;
    }
    LunnonMap *maps = reinterpret_cast<LunnonMap *>(lds + tab_span);
    for (uint32_t k = threadIdx.x; k < a.num_maps; k += blockDim.x) maps[k] = a.maps[k];
    __syncthreads();
}

// Split the lane's remaining work into the spill queue (see sym_kernels.hip).
template <typename Stack>
__device__ __forceinline__ void lspill_all(const LunnonArgs &a, uint8_t *lds, const Lane &ln, uint32_t lane,
                                           LWork &w, LTotals &tot) {
    bool pend = w.have;
    uint32_t t = w.l;
    uint64_t mask = 0;
    uint32_t floor = w.floor;
    // A lane in the middle of backtracking has no pending gap computation.
    bool emit = pend && !w.popping;
    while (__any(pend)) {
        uint32_t cnt = 0;
        const uint32_t slot = wave_alloc(a.spill_count, emit, lane, &cnt);
        if (emit) {
            if (slot < a.spill_cap) {
                WorkItem *dst = a.spill + slot;
                row_store_n<kLStride>(lds, ln.cl, dst->state, a.row_entries);
                dst->mask = mask;
                dst->origin = w.origin;
                dst->mult = static_cast<uint16_t>(w.n | (floor << 8));
                dst->level = static_cast<uint8_t>(t);
                dst->map = static_cast<uint8_t>(w.map);
            } else {
                atomicOr(a.acc + ACC_ERROR, 1ull);
            }
        }
        if (cnt != 0 && lane == 0) atomicAdd(a.acc + ACC_SPILLED, static_cast<unsigned long long>(cnt));
        emit = false;
        if (pend) {
            if (t == w.root) {
                pend = false;
            } else {
                --t;
                const typename Stack::Entry e = Stack::get(lds, ln, t);
                const uint32_t j = Stack::child(e);
                unlink(lds, ln, j, t);
                bool unknown = false;
                mask = Stack::rest(e, unknown);
                floor = 0;
                if (unknown) {
                    mask = 0;
                    floor = j + 1;
                }
                emit = mask != 0 || floor != 0;
            }
        }
    }
    if (w.have) {
        lretire(a, w, tot);
        w.have = false;
    }
}

}  // namespace

template <typename Stack, bool EV, uint32_t D>
__global__ __launch_bounds__(256) void lunnon_static_kernel(LunnonArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    const uint32_t span = tables_span(a.tables_bytes);
    lstage(a, lds, span);
    const LunnonMap *maps = reinterpret_cast<const LunnonMap *>(lds + span);
    Lane ln = lmake_lane<Stack>(span, a.levels, a.row_entries);
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    LTotals tot;
    LWork w;
    w.have = false;
    w.popping = false;
    if (idx < a.num_items) lload_item<Stack>(lds, ln, a.items + idx, maps, a.row_entries, w);
    while (w.have) {
        if (ldfs_step<Stack, EV, D>(lds, ln, w)) {
            lretire(a, w, tot);
            w.have = false;
        } else {
            lfold_item(a, w, tot);
        }
    }
    lflush(a, threadIdx.x & 63u, tot);
}

template <typename Stack, bool EV, uint32_t D>
__global__ __launch_bounds__(256) void lunnon_persistent_kernel(LunnonArgs a) {
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
    const uint32_t span = tables_span(a.tables_bytes);
    lstage(a, lds, span);
    const LunnonMap *maps = reinterpret_cast<const LunnonMap *>(lds + span);
    Lane ln = lmake_lane<Stack>(span, a.levels, a.row_entries);
    const uint32_t lane = threadIdx.x & 63u;
    const unsigned long long below = (1ull << lane) - 1ull;

    LTotals tot;
    LWork w;
    w.have = false;
    w.popping = false;
    bool drained = false;
    uint32_t tail_left = a.tail;

    for (;;) {
        if (!drained) {
            const unsigned long long idle = __ballot(!w.have);
            if (idle != 0) {
                const uint32_t want = static_cast<uint32_t>(__popcll(idle));
                const uint32_t first = static_cast<uint32_t>(__builtin_ctzll(idle));
                uint32_t got = 0;
                if (lane == first) got = atomicAdd(a.claim, want);
                got = __shfl(got, first, 64);
                if (got + want >= num_items) drained = true;
                if (!w.have) {
                    const uint32_t idx = got + static_cast<uint32_t>(__popcll(idle & below));
                    if (idx < num_items) lload_item<Stack>(lds, ln, a.items + idx, maps, a.row_entries, w);
                }
            } else {
                drained = __hip_atomic_load(a.claim, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= num_items;
            }
        }
        if (!__any(w.have)) break;

        for (uint32_t s = 0; s < a.budget && w.have; ++s) {
            if (ldfs_step<Stack, EV, D>(lds, ln, w)) {
                lretire(a, w, tot);
                w.have = false;
            }
        }
        if (w.have) lfold_item(a, w, tot);

        if (drained) {
            tail_left = tail_left > a.budget ? tail_left - a.budget : 0;
            if (tail_left == 0) {
                // Progress guarantee: every lane completes one counted gap
                // computation of its item before its remaining work is split.
                while (__any(w.have && !w.progressed)) {
                    if (w.have && !w.progressed && ldfs_step<Stack, EV, D>(lds, ln, w)) {
                        lretire(a, w, tot);
                        w.have = false;
                    }
                }
                lspill_all<Stack>(a, lds, ln, lane, w, tot);
                break;
            }
        }
    }
    lflush(a, lane, tot);
}

namespace {
template <typename Stack, uint32_t D>
const void *lpick(bool persistent, bool events) {
    if (persistent) return events ? reinterpret_cast<const void *>(&lunnon_persistent_kernel<Stack, true, D>)
                                  : reinterpret_cast<const void *>(&lunnon_persistent_kernel<Stack, false, D>);
    return events ? reinterpret_cast<const void *>(&lunnon_static_kernel<Stack, true, D>)
                  : reinterpret_cast<const void *>(&lunnon_static_kernel<Stack, false, D>);
}
}  // namespace

uint32_t lunnon_lds_bytes(int core, uint32_t tables_bytes, uint32_t levels, uint32_t block, uint32_t rows) {
    const uint32_t slot = core == 1 ? WideStack::kSlotBytes : CompactStack::kSlotBytes;
    return ((tables_bytes + 15u) & ~15u) + kMapBytes + (block / 64u) * (((rows * kLStride + 15u) & ~15u) + levels * slot);
}

// 2 when every map of the launch is two-dimensional (every rectangle) with
// at most 63 leaves, 1 when they are two-dimensional with more, else 0.
uint32_t lunnon_dims(const LunnonArgs &a) {
    uint32_t d = 2;
    for (uint32_t k = 0; k < a.num_maps; ++k) {
        if (a.maps[k].ndim != 2) return 0;
        if (a.maps[k].n > 63) d = 1;
    }
    return d;
}

// Whether `core` runs the launch with both dimensions walked in one loop
// (its rows then need entry n too: lunnon_args).
bool lunnon_paired(int core, const LunnonArgs &a) { return core != 1 && core != 4 && lunnon_dims(a) == 2; }

// core: 1 = wide u64 stack; 4 = compact stack, dimensions walked one after
// the other; otherwise compact, two-dimensional launches walking both
// dimensions in one loop.  dims as lunnon_dims gives it.
const void *lunnon_kernel_ptr(int core, bool persistent, bool events, uint32_t dims) {
    if (core == 1) return lpick<WideStack, kDimsAny>(persistent, events);
    if (dims == 0) return lpick<CompactStack, kDimsAny>(persistent, events);
    if (dims == 1 || core == 4) return lpick<CompactStack, kDims2>(persistent, events);
    return lpick<CompactStack, kDims2Paired>(persistent, events);
}

hipError_t launch_lunnon(const LunnonArgs &a, int core, bool persistent, uint32_t grid, uint32_t block,
                         uint32_t lds_bytes, bool events, hipStream_t s) {
    void *args[] = {const_cast<LunnonArgs *>(&a)};
    return hipLaunchKernel(lunnon_kernel_ptr(core, persistent, events, lunnon_dims(a)), dim3(grid), dim3(block), args,
                           lds_bytes, s);
}

}  // namespace mf
