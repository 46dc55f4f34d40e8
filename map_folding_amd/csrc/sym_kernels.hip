// sym_kernels.hip -- symmetry-reduced cycle DFS on gfx950.
//
// Restates recurse_count / get_good / iter_good (sym_fold.cpp:179-233) as a
// per-lane iterative DFS:
//   * one lane owns one subtree (a WorkItem);
//   * its cycle (cell -> successor, bytes) and a stack of remaining-children
//     masks (one u64 per level) live in LDS, lane-interleaved so that lanes
//     never share a bank: cycle dword w of lane L is at word (w*64 + L), stack
//     slot s of lane L at u64 (s*64 + L);
//   * the four edge pages are packed one word per cell in LDS (page p in byte p);
//   * node = one get_good call (sym_fold9.cpp:342-344); a node at the last
//     level adds popcount(good) leaves instead of recursing.
// Two schedules share this core:
//   sym_static_kernel      one lane per input item, one launch (the plain
//                          static-frontier schedule);
//   sym_persistent_kernel  resident lanes claim items from the input array a
//                          wave at a time; once the input is drained, a lane
//                          that is still busy after `tail` more steps splits
//                          its remaining work (the node it is about to expand,
//                          plus the untried siblings of every level on its
//                          stack) into the spill queue, which the host feeds
//                          back as the next round's input.
// Counts never depend on the schedule: every subtree is counted exactly once.
#include <hip/hip_runtime.h>

#include "mf_device.hpp"
#include "mf_wave.hpp"

namespace mf {
namespace {

constexpr uint32_t kEdgeBytes = 256;
constexpr uint32_t kCycleBytes = 4096;   // 16 words x 64 lanes

__device__ __forceinline__ uint32_t cyc_addr(uint32_t lane_base, uint32_t cell) { return row_addr(lane_base, cell); }

// get_good for cell i against this lane's cycle.  Walk semantics of
// sym_fold.cpp:188-195: from the anchor (cell i's neighbour on page p), follow
// cycle successors; a cell whose page-p partner is already placed is replaced
// by that partner; every cell reached is a legal insertion point.
template <bool EV>
__device__ __forceinline__ uint64_t good_mask(const uint8_t *__restrict__ lds, uint32_t cl, uint32_t i,
                                              uint32_t &walks, uint32_t &steps) {
    uint64_t good = (i >= 64u) ? ~0ull : ((1ull << i) - 1ull);
    const uint32_t ew = reinterpret_cast<const uint32_t *>(lds)[i];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const uint32_t anchor = (ew >> (8 * p)) & 0xffu;
        if (anchor < i && good != 0) {
            if (EV) ++walks;
            uint64_t reach = 1ull << anchor;
            uint32_t j = lds[cyc_addr(cl, anchor)];
            while (j != anchor) {
                if (EV) ++steps;
                const uint32_t partner = lds[j * 4u + p];
                j = (partner < i) ? partner : j;
                reach |= 1ull << j;
                j = lds[cyc_addr(cl, j)];
            }
            good &= reach;
        }
    }
    return good;
}

struct Lane {
    uint32_t cl;      // LDS byte address of this lane's cycle (dword 0)
    uint32_t sl;      // LDS byte address of this lane's stack slot 0
    uint32_t base;    // level held by stack slot 0
};

__device__ __forceinline__ uint64_t stk_get(const uint8_t *lds, const Lane &ln, uint32_t level) {
    return *reinterpret_cast<const uint64_t *>(lds + ln.sl + (level - ln.base) * 512u);
}
__device__ __forceinline__ void stk_put(uint8_t *lds, const Lane &ln, uint32_t level, uint64_t m) {
    *reinterpret_cast<uint64_t *>(lds + ln.sl + (level - ln.base) * 512u) = m;
}

// Insert cell i after cell j (iter_good, sym_fold.cpp:210-213).
__device__ __forceinline__ void place(uint8_t *lds, const Lane &ln, uint32_t j, uint32_t i) {
    const uint8_t nx = lds[cyc_addr(ln.cl, j)];
    lds[cyc_addr(ln.cl, i)] = nx;
    lds[cyc_addr(ln.cl, j)] = static_cast<uint8_t>(i);
}
// Remove cell i again (sym_fold.cpp:215): its successor goes back to j.
__device__ __forceinline__ void unplace(uint8_t *lds, const Lane &ln, uint32_t j, uint32_t i) {
    lds[cyc_addr(ln.cl, j)] = lds[cyc_addr(ln.cl, i)];
}

struct Work {
    uint32_t i;        // next level to expand
    uint32_t root;     // item root level: finished when backtracking reaches it
    uint32_t origin;
    uint32_t mult;
    uint64_t leaves;   // leaves of the current item so far
    uint32_t inodes;   // nodes, page walks and walk steps of the current item
    uint32_t iwalks;
    uint32_t isteps;
    bool have;
};

// Lane totals over all items (64-bit; item counters are folded in when an
// item is finished or split).
struct Totals {
    unsigned long long sum = 0;     // sum of mult * leaves
    unsigned long long nodes = 0, walks = 0, steps = 0, leaves = 0, items = 0;
};

__device__ __forceinline__ void load_item(uint8_t *lds, const Lane &ln, const WorkItem *__restrict__ it, Work &w) {
    state_to_lds(lds, ln.cl, it->state);
    const uint64_t mask = it->mask;
    const uint32_t level = it->level;
    w.origin = it->origin;
    w.mult = it->mult;
    w.root = level;
    w.leaves = 0;
    w.inodes = 0;
    w.iwalks = 0;
    w.isteps = 0;
    w.have = true;
    if (mask == 0) {
        w.i = level;
    } else {
        stk_put(lds, ln, level, mask);
        place(lds, ln, static_cast<uint32_t>(__builtin_ctzll(mask)), level);
        w.i = level + 1;
    }
}

// One node: expand level w.i, then move to the next node of the subtree (or
// finish).  Returns true when the item is finished.
template <bool EV>
__device__ __forceinline__ bool dfs_step(uint8_t *lds, const Lane &ln, uint32_t last, Work &w) {
    uint64_t g = good_mask<EV>(lds, ln.cl, w.i, w.iwalks, w.isteps);
    ++w.inodes;
    if (w.i == last) {
        w.leaves += static_cast<uint32_t>(__builtin_popcountll(g));
        g = 0;
    }
    if (g != 0) {
        stk_put(lds, ln, w.i, g);
        place(lds, ln, static_cast<uint32_t>(__builtin_ctzll(g)), w.i);
        ++w.i;
        return false;
    }
    for (;;) {
        if (w.i == w.root) return true;
        --w.i;
        uint64_t m = stk_get(lds, ln, w.i);
        unplace(lds, ln, static_cast<uint32_t>(__builtin_ctzll(m)), w.i);
        m &= m - 1;
        if (m != 0) {
            stk_put(lds, ln, w.i, m);
            place(lds, ln, static_cast<uint32_t>(__builtin_ctzll(m)), w.i);
            ++w.i;
            return false;
        }
    }
}

// Fold the current item's counters into the lane totals and the per-origin
// results (used when an item finishes and when it is split).
__device__ __forceinline__ void retire(const SymArgs &a, const Work &w, Totals &t) {
    t.sum += static_cast<unsigned long long>(w.mult) * w.leaves;
    t.nodes += w.inodes;
    t.walks += w.iwalks;
    t.steps += w.isteps;
    t.leaves += w.leaves;
    t.items += 1;
    if (a.per_count != nullptr && w.leaves != 0) atomicAdd(a.per_count + w.origin, static_cast<unsigned long long>(w.leaves));
    if (a.per_nodes != nullptr && w.inodes != 0) atomicAdd(a.per_nodes + w.origin, static_cast<unsigned long long>(w.inodes));
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

__device__ __forceinline__ Lane make_lane(uint32_t levels, uint32_t base_level) {
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t wave_bytes = kCycleBytes + levels * 512u;
    const uint32_t wbase = kEdgeBytes + wave * wave_bytes;
    Lane ln;
    ln.cl = wbase + lane * 4u;
    ln.sl = wbase + kCycleBytes + lane * 8u;
    ln.base = base_level;
    return ln;
}

}  // namespace

// ---------------------------------------------------------------- static
template <bool EV>
__global__ __launch_bounds__(256) void sym_static_kernel(SymArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    const uint32_t tid = threadIdx.x;
    if (tid < 64) reinterpret_cast<uint32_t *>(lds)[tid] = a.edge_word[tid];
    __syncthreads();
    const Lane ln = make_lane(a.levels, a.base_level);
    const uint32_t idx = blockIdx.x * blockDim.x + tid;
    Totals tot;
    Work w;
    w.have = false;
    if (idx < a.num_items) load_item(lds, ln, a.items + idx, w);
    while (w.have) {
        if (dfs_step<EV>(lds, ln, a.last, w)) {
            retire(a, w, tot);
            w.have = false;
        }
    }
    flush_wave(a, tid & 63u, tot);
}

// ------------------------------------------------------------- persistent
template <bool EV>
__global__ __launch_bounds__(256) void sym_persistent_kernel(SymArgs a) {
    extern __shared__ __align__(16) uint8_t lds[];
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    if (tid < 64) reinterpret_cast<uint32_t *>(lds)[tid] = a.edge_word[tid];
    __syncthreads();
    const Lane ln = make_lane(a.levels, a.base_level);
    const unsigned long long below = (1ull << lane) - 1ull;

    Totals tot;
    Work w;
    w.have = false;
    bool drained = false;   // wave-uniform: the input array has been handed out
    uint32_t tail_left = a.tail;

    for (;;) {
        // Idle lanes claim the next input items, one atomic per wave.
        if (!drained) {
            const unsigned long long idle = __ballot(!w.have);
            if (idle != 0) {
                const uint32_t want = static_cast<uint32_t>(__popcll(idle));
                const uint32_t first = static_cast<uint32_t>(__builtin_ctzll(idle));
                uint32_t got = 0;
                if (lane == first) got = atomicAdd(a.claim, want);
                got = __shfl(got, first, 64);
                if (got + want >= a.num_items) drained = true;
                if (!w.have) {
                    const uint32_t idx = got + static_cast<uint32_t>(__popcll(idle & below));
                    if (idx < a.num_items) load_item(lds, ln, a.items + idx, w);
                }
            } else {
                // Every lane is busy: look at the shared cursor anyway, so that
                // a wave of long items still learns that the input is gone.
                drained = __hip_atomic_load(a.claim, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= a.num_items;
            }
        }
        if (!__any(w.have)) break;

        for (uint32_t s = 0; s < a.budget && w.have; ++s) {
            if (dfs_step<EV>(lds, ln, a.last, w)) {
                retire(a, w, tot);
                w.have = false;
            }
        }

        if (drained) {
            tail_left = tail_left > a.budget ? tail_left - a.budget : 0;
            if (tail_left == 0) {
                // Split everything this lane still holds into the spill queue:
                // first the node about to be expanded, then for each stack level
                // (deepest first) the siblings not yet entered.
                bool pend = w.have;
                uint32_t t = w.i;          // level of the next piece to emit
                uint64_t mask = 0;         // 0 = expand node t
                bool emit = pend;
                while (__any(pend)) {
                    const unsigned long long em = __ballot(emit);
                    if (em != 0) {
                        const uint32_t cnt = static_cast<uint32_t>(__popcll(em));
                        const uint32_t first = static_cast<uint32_t>(__builtin_ctzll(em));
                        uint32_t slot0 = 0;
                        if (lane == first) slot0 = atomicAdd(a.spill_count, cnt);
                        slot0 = __shfl(slot0, first, 64);
                        if (emit) {
                            const uint32_t slot = slot0 + static_cast<uint32_t>(__popcll(em & below));
                            if (slot < a.spill_cap) {
                                WorkItem *dst = a.spill + slot;
                                lds_to_state(lds, ln.cl, dst->state);
                                dst->mask = mask;
                                dst->origin = w.origin;
                                dst->mult = static_cast<uint16_t>(w.mult);
                                dst->level = static_cast<uint8_t>(t);
                                dst->map = 0;
                            } else {
                                atomicOr(a.acc + ACC_ERROR, 1ull);
                            }
                        }
                        if (lane == first) atomicAdd(a.acc + ACC_SPILLED, static_cast<unsigned long long>(cnt));
                    }
                    // Step up one level: undo that level's current child; the
                    // remaining siblings become the next piece.
                    emit = false;
                    if (pend) {
                        if (t == w.root) {
                            pend = false;
                        } else {
                            --t;
                            const uint64_t m = stk_get(lds, ln, t);
                            unplace(lds, ln, static_cast<uint32_t>(__builtin_ctzll(m)), t);
                            mask = m & (m - 1);
                            emit = mask != 0;
                        }
                    }
                }
                if (w.have) {
                    // Leaves and nodes found so far stay with this lane.
                    retire(a, w, tot);
                    w.have = false;
                }
                break;
            }
        }
    }
    flush_wave(a, lane, tot);
}

// Host-visible launchers (C++ linkage inside the library).
void launch_sym_static(const SymArgs &a, uint32_t grid, uint32_t block, uint32_t lds_bytes, bool events,
                       hipStream_t s) {
    if (events)
        hipLaunchKernelGGL(sym_static_kernel<true>, dim3(grid), dim3(block), lds_bytes, s, a);
    else
        hipLaunchKernelGGL(sym_static_kernel<false>, dim3(grid), dim3(block), lds_bytes, s, a);
}

void launch_sym_persistent(const SymArgs &a, uint32_t grid, uint32_t block, uint32_t lds_bytes, bool events,
                           hipStream_t s) {
    if (events)
        hipLaunchKernelGGL(sym_persistent_kernel<true>, dim3(grid), dim3(block), lds_bytes, s, a);
    else
        hipLaunchKernelGGL(sym_persistent_kernel<false>, dim3(grid), dim3(block), lds_bytes, s, a);
}

const void *sym_kernel_ptr(bool persistent, bool events) {
    if (persistent) return events ? reinterpret_cast<const void *>(&sym_persistent_kernel<true>)
                                  : reinterpret_cast<const void *>(&sym_persistent_kernel<false>);
    return events ? reinterpret_cast<const void *>(&sym_static_kernel<true>)
                  : reinterpret_cast<const void *>(&sym_static_kernel<false>);
}

uint32_t sym_lds_bytes(uint32_t levels, uint32_t block) {
    return kEdgeBytes + (block / 64u) * (kCycleBytes + levels * 512u);
}

}  // namespace mf
