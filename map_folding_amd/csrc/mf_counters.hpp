// mf_counters.hpp -- per-item event counters that cannot wrap.
//
// A lane counts the events of its current work item (nodes, page walks /
// gap computations, walk steps, candidates) in 32-bit registers: 64-bit
// increments cost two VALU instructions each on the hot path.  An item of the
// static schedule can run for far more than 2^32 events, so the kernels fold
// the 32-bit counters into the lane's 64-bit totals once any of them reaches
// kFoldAt.  The kernels test after every DFS step (static schedule) or every
// `budget` steps (persistent schedule; the host caps the budget at
// kMaxFoldBudget), and one step adds at most kMaxStepEvents to any counter,
// so no counter passes 2^32 between two tests.
//
// This header has no HIP dependency so that the host test suite compiles the
// fold logic as it is.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define MF_HD __host__ __device__
#else
#define MF_HD
#endif

namespace mf {

constexpr uint32_t kFoldAt = 0x80000000u;
// Events one DFS step can add to one counter: sym walks 4 pages of at most
// 128 cells; Lunnon walks at most 4 dimensions of at most 64 leaves, and an
// insert step adds at most 64 inserts / candidates.
constexpr uint32_t kMaxStepEvents = 4u * 128u;
// The largest persistent budget (steps between two tests) the host allows.
constexpr uint32_t kMaxFoldBudget = 1u << 20;
static_assert(static_cast<uint64_t>(kMaxFoldBudget) * kMaxStepEvents < (1ull << 32) - kFoldAt,
              "a counter below kFoldAt must stay below 2^32 for kMaxFoldBudget steps");

struct FoldTotals {
    unsigned long long nodes = 0, walks = 0, steps = 0, cands = 0;
};

// When any counter has reached kFoldAt, move the counters into `t`.  The
// node counter keeps 1 (an item that has counted a node still shows it: the
// persistent kernels' progress guarantee reads it).  Returns the nodes moved,
// which the caller also adds to the item's per-origin node log; 0 when
// nothing was folded.
MF_HD inline uint32_t fold_item_counters(uint32_t &nodes, uint32_t &walks, uint32_t &steps, uint32_t &cands,
                                         FoldTotals &t) {
    if (((nodes | walks | steps | cands) & kFoldAt) == 0u) return 0u;
    const uint32_t moved = nodes > 0u ? nodes - 1u : 0u;
    t.nodes += moved;
    t.walks += walks;
    t.steps += steps;
    t.cands += cands;
    nodes -= moved;
    walks = 0u;
    steps = 0u;
    cands = 0u;
    return moved;
}

}  // namespace mf
