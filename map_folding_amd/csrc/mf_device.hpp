// mf_device.hpp -- work-item formats and kernel arguments shared by the HIP
// kernels and the host library.
//
// A work item is one subtree root of either DFS:
//   * mask == 0: "expand" -- the node at `level` itself (get_good / gap
//     computation not yet done for it);
//   * mask != 0: "siblings" -- the children of the node at `level` listed in
//     mask (a subset of its legal insertion points) that nobody has entered.
// Every subtree of the search is covered by exactly one item at any time, so
// splitting an item into pieces never changes any count.
#pragma once

#include <cstdint>

namespace mf {

// sym: cycle[c] = successor of cell c in the cyclic stack order (cells < level).
// Lunnon: state[m] = "below" link b[m] of placed leaf m (leaves < level).
struct alignas(16) WorkItem {
    uint8_t state[64];
    uint64_t mask;
    uint32_t origin;   // index of the input item this subtree came from
    uint16_t mult;     // sym: orbit multiplier (<= 16)
    uint8_t level;     // sym: next cell to place; Lunnon: next leaf to insert
    uint8_t map;       // Lunnon batch: map index (table selector)
};
static_assert(sizeof(WorkItem) == 80, "WorkItem layout");

// 9x9 symmetry cases: 81-cell cycles and u128 sibling masks.
struct alignas(16) WorkItem128 {
    uint8_t state[96];
    uint64_t mask_lo;
    uint64_t mask_hi;
    uint32_t origin;
    uint16_t mult;
    uint8_t level;
    uint8_t map;       // floor of a recompute piece (0 = none)
    uint8_t pad[8];
};
static_assert(sizeof(WorkItem128) == 128, "WorkItem128 layout");

// Global accumulators (device memory, zeroed before a solve).
enum AccSlot {
    ACC_COUNT_LO = 0,  // sum of mult * count, 128 bits
    ACC_COUNT_HI,
    ACC_NODES,
    ACC_WALKS,
    ACC_STEPS,
    ACC_LEAVES,
    ACC_CANDS,
    ACC_ITEMS,         // work items finished or spilled
    ACC_SPILLED,       // items written to the spill queue
    ACC_ERROR,         // nonzero: capacity overflow or consistency failure
    ACC_ROUNDS,        // persistent launches that had input
    ACC_SLOTS
};

// Per-map table descriptor for the Lunnon kernels.
struct LunnonMap {
    uint32_t table_off;   // byte offset of this map's d table in the table blob
    uint16_t n;           // leaves
    uint8_t ndim;
    uint8_t base;         // frontier leaf D+1 of this map: stack slot 0
};

constexpr int kMaxMaps = 64;

struct SymArgs {
    uint32_t edge_word[96];     // 4 pages of cell c packed into one word (partner, n2 = none)
    const void *items;          // WorkItem (up to 64 cells) or WorkItem128 (9x9)
    uint32_t num_items;
    const uint32_t *num_items_dev;  // persistent: when set, the item count is read here (graph rounds)
    uint32_t last;              // n2 - 1: the level whose good mask is counted
    uint32_t base_level;        // lowest level any item starts at (stack slot 0)
    uint32_t levels;            // stack slots per lane
    uint32_t budget;            // persistent: core steps between claim/spill checks
    uint32_t tail;              // persistent: steps a lane may run after the input is drained
    uint32_t *claim;            // persistent: next input index
    void *spill;                // persistent: output queue of split work (item type as `items`)
    uint32_t *spill_count;
    uint32_t spill_cap;
    unsigned long long *acc;    // AccSlot
    unsigned long long *per_count;  // per origin (nullable)
    unsigned long long *per_nodes;  // per origin (nullable)
    // Deepening pass (sym_deepen_kernel): items below their target level are
    // expanded by one level; children that reach the target, and input items
    // already at or past it, go to `final_out`, the others to `spill` with
    // their target in the mask field (expand items carry no sibling mask).
    void *final_out;
    uint32_t *final_count;
    uint32_t final_cap;
    uint32_t target;            // first pass: the target of every item without a per-item one
    const uint8_t *targets;     // first pass: per-item target levels (nullable)
    uint32_t deep_pass;         // 0: first pass (items are frontier items), 1: later (target in mask)
};

// ACC_ERROR bits: 1 spill-queue overflow, 2/4 consistency checks, 16 a
// deepening pass ran out of room (the host then solves without deepening).
constexpr unsigned long long kErrDeepenFull = 16ull;

struct LunnonArgs {
    const uint8_t *tables;      // concatenated d tables, [ndim][n+1][n+1] bytes each
    uint32_t tables_bytes;
    uint32_t row_entries;       // entries of a lane's LDS row: the largest map's leaf count n, plus one for the paired two-dimensional walk (at most 64)
    uint32_t num_maps;
    LunnonMap maps[kMaxMaps];
    const WorkItem *items;
    uint32_t num_items;
    const uint32_t *num_items_dev;  // persistent: when set, the item count is read here (graph rounds)
    uint32_t levels;
    uint32_t budget;
    uint32_t tail;
    uint32_t *claim;
    WorkItem *spill;
    uint32_t *spill_count;
    uint32_t spill_cap;
    unsigned long long *acc;
    unsigned long long *per_count;
    unsigned long long *per_nodes;
    unsigned long long *per_map;    // per map exact count (nullable)
};

}  // namespace mf
