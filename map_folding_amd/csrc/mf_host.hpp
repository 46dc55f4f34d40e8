// mf_host.hpp -- host-side problem tables and frontier generators (C++17).
//
// These are the product's own generators (the CPU half of the hot path): the
// Lunnon frontier collector and the symmetry-reduced seed / case stream.  They
// define the work-item order that every count and per-item result refers to,
// so their orders follow the reference exactly (cited per function).
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

// HIP's stream and event handles (hipStream_t, hipEvent_t) without the HIP
// headers, which the g++-built host sources do not include.
struct ihipStream_t;
struct ihipEvent_t;

namespace mf {

constexpr int kMaxLeaves = 64;   // n = prod(dims) <= 64 (one u64 mask per level)
constexpr int kMaxDims = 4;      // GPU Lunnon kernel: up to 4 dimensions
constexpr int kCells = 64;       // sym: case row bytes up to 8x8 (one u64 mask)
constexpr int kMaxCells = 128;   // sym: case row bytes for 9x9 (81 cells, u128 masks)

// ------------------------------------------------------------------ Lunnon

struct LunnonProblem {
    int ndim = 0;
    int n = 0;
    int dims[kMaxDims] = {0, 0, 0, 0};
    // d[i][l][m] for l, m in [0, n]; m = 0 is 0.  Stored as bytes: leaves <= 64.
    std::vector<uint8_t> d;  // ndim * (n+1) * (n+1)
    int stride() const { return n + 1; }
    uint8_t at(int i, int l, int m) const { return d[(static_cast<size_t>(i) * (n + 1) + l) * (n + 1) + m]; }
};

// Returns an empty optional-like flag through `ok`; error text in `err`.
bool lunnon_problem(const int32_t *dims, int ndim, LunnonProblem &pr, std::string &err);

// True if the folding DFS (mf.c:82-142, with the normal-form filter) reaches
// this state: leaves depth, depth-1, ..., 2 each sit directly below a position
// that the gap walks of the smaller state allow, never below the root.  The
// DFS kernels rely on it: every gap walk of a reachable state and of its
// descendants returns to its leaf, so they carry no step guard.
bool lunnon_reachable(const LunnonProblem &pr, const uint8_t *state, int depth);

// Canonical frontier (collect_states order).  Appends b[0..D] rows to `out`
// until `max_states` rows were written; returns the total frontier size.
int64_t lunnon_frontier(const LunnonProblem &pr, int depth, uint8_t *out, int64_t max_states);

// ------------------------------------------------------------------ sym

struct SymProblem {
    int n = 0;
    int n2 = 0;
    int center = 0;                  // c: side of the seed center (c*c cells)
    int order = 0;                   // cell numbering: 0 spiral, else MF_ORDER's ring order (odd n)
    int lock_radius = 0;             // rings up to this radius keep spiral order (MF_LOCK_CENTER_N)
    uint8_t edge[4][kMaxCells];      // edge[p][c] = neighbour of cell c on page p, n2 if none
    uint32_t edge_word[kMaxCells];   // the 4 pages of cell c packed, page p in bits 8p..8p+7
    int map[kMaxCells];              // spiral number of (x, y) at map[x*n + y]
    // Bytes per case row: 64 up to 8x8, 128 for 9x9.
    int case_bytes() const { return n2 <= 64 ? kCells : kMaxCells; }
};

// Seed center and cell numbering of the symmetry DFS.  The defaults are the
// reference's: center c = 5 for n = 9, else 4 (n even) or 3 (n odd), and the
// center-out spiral.  sym_fold9 lets MF_CENTER_N / MF_ORDER /
// MF_LOCK_CENTER_N change them (sym_fold9.cpp:160-229,968-980); its CLI reads
// those variables, the library reads no environment.
struct SymNumbering {
    int center = 0;          // 0 = the default for n; else 3..5, same parity as n (n >= 6)
    int order = 0;           // 0 spiral, 1 reverse, 2 opposite, 3 axes_first, 4 corners_first
    int lock_center_n = 3;   // rings of radius <= (lock_center_n - 1) / 2 keep spiral order
};

// The process-wide numbering (mf_sym_set_numbering); every entry point that
// takes only n uses it.
SymNumbering sym_numbering();
bool set_sym_numbering(const SymNumbering &num, std::string &err);

// n in 1..9 under the given numbering (an odd n only ever renumbers; the
// center override applies to n >= 6).
bool sym_problem(int n, const SymNumbering &num, SymProblem &pr, std::string &err);
// The same under the process-wide numbering.
bool sym_problem(int n, SymProblem &pr, std::string &err);

// Host get_good (generators, features).  sym_good: n <= 8; sym_good128: any n.
uint64_t sym_good(const SymProblem &pr, const uint8_t *cycle, int i);
unsigned __int128 sym_good128(const SymProblem &pr, const uint8_t *cycle, int i);

struct SymSeed {
    uint8_t cycle[kMaxCells];
    uint64_t mult;
};

// Deduplicated center cycles in generation order; raw = number before dedup.
std::vector<SymSeed> sym_seeds(const SymProblem &pr, int64_t *raw);
// The same, computed once per (n, center) in the process.
const std::vector<SymSeed> &sym_seed_list(const SymProblem &pr, int64_t *raw);

// Per-seed case counts at level `expanded` (cached per (n, center, expanded)
// in the process; computed on all host threads).
const std::vector<uint64_t> &sym_seed_case_counts(const SymProblem &pr, int expanded);

// Per-seed case counts of the seed slice [seed_start, seed_start+seed_count)
// (clipped to the seed list; the full list comes from the cache above).
std::vector<uint64_t> sym_slice_case_counts(const SymProblem &pr, int expanded, uint64_t seed_start,
                                            uint64_t seed_count);

// Cases [start, start+max) in stream order of the seed slice [seed_start,
// seed_start+seed_count) (sym_fold9.cpp:1071-1090: case index 0 is the first
// case of seed seed_start), pr.case_bytes() bytes per case.  Returns the
// number written.
int64_t sym_cases(const SymProblem &pr, int expanded, uint64_t start, uint64_t max_cases,
                  uint8_t *cycles, uint64_t *mults, uint64_t seed_start = 0, uint64_t seed_count = ~0ull);

// The library context as seen by its other translation units.
struct LibContext {
    int device = -1;
    int cus = 0;
    ::ihipStream_t *stream = nullptr;
    ::ihipEvent_t *ev0 = nullptr, *ev1 = nullptr;
    unsigned long long *host_acc = nullptr;   // pinned, ACC_SLOTS entries
    uint32_t *host_counters = nullptr;        // pinned, 4 entries
};
bool lib_current(LibContext *c);

}  // namespace mf
