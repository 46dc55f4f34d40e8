/*
 * mapfold.h -- C ABI of the MI355X map-folding engine (libmapfold.so).
 *
 * Plain C types only: pointers, sizes, integers.  Every entry point returns an
 * int status (MF_OK = 0) and never throws across the boundary; mf_status_string
 * names a status.  All device work runs on one HIP stream owned by the library
 * (default priority) on the device chosen by mf_init.  Calls are not
 * thread-safe: one host thread drives one library context.
 *
 * Reference interfaces replaced (johnowhitaker/map_folding, file:line):
 *   mf_gpu_folding        gpu_folding.cu:209-311  (main: frontier -> launch -> total)
 *   mf_lunnon_frontier    gpu_folding_v2.cu:220-256, mf_fast.cpp:185-221 (collect_states)
 *   mf_lunnon_check_states  mf.c:82-142 (the insertions a state must replay)
 *   mf_lunnon_solve       gpu_folding_v2.cu:381-453 (solve_kernel / solve_state_device),
 *                         gpu_folding.cu:123-207 (compute_counts)
 *   mf_sym_seeds          sym_fold.cpp:689-710 (recurse_generate + deduplicate_cases)
 *   mf_sym_cases          sym_fold.cpp:491-527 (recurse_stream_cases, case-index window)
 *   mf_sym_solve          sym_fold.cpp:219-233,410-489,529-620 (recurse_count over cases)
 *   mf_sym_fold           sym_fold.cpp:649-774 (main: the decimal answer);
 *                         for n = 9, experiments/9x9/sym_fold9.cpp:984-1198
 *   mf_sym_cases_slice,   experiments/9x9/sym_fold9.cpp:1071-1090,1139-1172
 *   mf_sym_fold_slice,      (MF_SEED_START / MF_SEED_COUNT seed slices and the
 *   mf_sym_seed_case_counts  per-seed count_generated_cases of MF_SEED_COUNTS)
 *   mf_wasm_*             mfah/wasm/sym_kernel.c:20-34,251-277 (batch case ABI)
 */
#ifndef MAPFOLD_H
#define MAPFOLD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
    MF_OK = 0,
    MF_ERR_ARG = 1,        /* bad dimensions / sizes (cf. sym_kernel.c status 1) */
    MF_ERR_DEPTH = 2,      /* bad depth / level window (cf. sym_kernel.c status 2) */
    MF_ERR_HIP = 3,        /* a HIP runtime call failed (message via mf_last_error) */
    MF_ERR_NO_DEVICE = 4,  /* no GPU visible / mf_init not called */
    MF_ERR_CAPACITY = 5,   /* output buffer too small or work queue overflow */
    MF_ERR_INTERNAL = 6    /* device-side consistency check failed */
};

const char *mf_status_string(int status);
/* Text of the last error (HIP error string or argument message); never NULL. */
const char *mf_last_error(void);

/* ---- library context ---------------------------------------------------- */
/* Select the device and create the library's stream.  Idempotent per device. */
int mf_init(int device);
/* Number of compute units of the current device (0 before mf_init). */
int mf_device_cus(void);
/* Release device buffers and the stream. */
void mf_shutdown(void);

/* ---- results ------------------------------------------------------------ */
/* Exact totals.  `count` is the sum over work items of multiplier x subtree
 * count as 128 bits (lo, hi).  Event counters are exact and independent of
 * scheduling:
 *   sym:    nodes = get_good calls, walks = page walks, steps = walk steps,
 *           leaves = sum of popcount at the last level
 *   Lunnon: nodes = leaf inserts, walks = gap computations, steps = walk
 *           steps, leaves = completions (each worth +n), cands = distinct
 *           gap candidates before the filter (mf.c:99-123)                   */
typedef struct {
    uint64_t count_lo;
    uint64_t count_hi;
    uint64_t nodes;
    uint64_t walks;
    uint64_t steps;
    uint64_t leaves;
    uint64_t cands;        /* Lunnon: gap candidates before the filter; sym: 0 */
    uint64_t items;        /* work items executed (>= input items when re-split) */
    uint32_t rounds;       /* kernel rounds (1 for the static schedule) */
    uint32_t status;
    double kernel_ms;      /* sum of the solve kernels' durations (HIP events per launch) */
    double span_ms;        /* device time from the first launch to the end of the last */
    double total_ms;       /* host wall time of the call */
    uint64_t inputs;       /* items the first round started from (after any device deepening);
                              items - inputs of them were written by the re-split rounds */
} mf_result;

/* Frozen int32-op weights of the algorithmic work model (roofline figure):
 * sym ops = 16*nodes + 6*walks + 8*steps;
 * Lunnon ops = 10*gapcomps + 8*steps + 6*cands + 8*inserts + 2*completions. */
enum {
    MF_W_SYM_NODE = 16, MF_W_SYM_WALK = 6, MF_W_SYM_STEP = 8,
    MF_W_LUN_GAPCOMP = 10, MF_W_LUN_STEP = 8, MF_W_LUN_CAND = 6, MF_W_LUN_INSERT = 8, MF_W_LUN_COMPLETION = 2
};

/* Scheduling options.  Zero-initialise and set what you need; zeros select
 * tuned defaults.  Scheduling never changes any count. */
enum { MF_SCHED_STATIC = 0, MF_SCHED_RESPLIT = 1 };
typedef struct {
    int32_t schedule;         /* MF_SCHED_STATIC: one lane per item, one launch;
                                 MF_SCHED_RESPLIT: persistent lanes + on-device re-split */
    int32_t block;            /* threads per workgroup (64..256) */
    int32_t waves_per_cu;     /* persistent schedule: resident waves per CU */
    int32_t budget;           /* re-split: DFS iterations between a wave's claim / drain checks */
    int32_t tail;             /* re-split: iterations a lane may still run once the input is
                                 drained before it splits its remaining work */
    int32_t count_events;     /* 1: count walks/steps too (nodes/leaves always counted) */
    int32_t core;             /* DFS stack and walk: 0 = default (compact), 1 = one u64 mask
                                 per level, 2 = compact 16-bit entries with recompute, 3 = compact
                                 with two lanes per subtree (sym: a lane pair shares one row and
                                 stack and splits each node's page walks; re-split schedule,
                                 up to 8x8), 4 = compact with unpaired walks (sym: one page walk
                                 at a time; Lunnon: one dimension at a time, where the default
                                 walks the two dimensions of two-dimensional maps in one loop) */
    int32_t graph;            /* re-split: 0 = default (on), 1 = on, 2 = off.  On: after the
                                 first round, rounds run as a captured HIP graph of 8 rounds
                                 (device-side item counts), replayed until a round spills
                                 nothing; off: one host-driven launch per round. */
    uint64_t queue_capacity;  /* re-split: work items the device queue can hold */
    const uint32_t *order;    /* optional host permutation of item indices (learned order);
                                 NULL = input order.  Never changes any count. */
    uint64_t order_len;       /* entries of `order`: must equal the frontier's item count */
    int32_t deepen;           /* sym re-split, up to 8x8: before the rounds, expand the frontier on
                                 the device until every item reaches this level (fewer stack levels
                                 per lane, more resident waves); 0 = default (46 for 8x8, else off),
                                 -1 = off.  Every expanded node is counted as the DFS counts it. */
    const uint8_t *deepen_levels;  /* optional per-item deepening targets (e.g. chosen by the cost
                                 model): item k is expanded to at least max(deepen_levels[k], the
                                 `deepen` target); entries past n*n - 2 are clipped.  NULL = none. */
    uint64_t deepen_levels_len;  /* entries of deepen_levels: must equal the frontier's item count */
} mf_solve_options;

/* ---- Lunnon DFS (mf.c / gpu_folding*.cu) --------------------------------- */
/* Canonical frontier at partition depth D (1 <= D <= n): every node at leaf
 * D+1 passing the normal-form filter, in DFS order.  State s is written as
 * b[0..D] (D+1 bytes) at out_b + s*(D+1).  *num_states receives the frontier
 * size even when it exceeds max_states (then only max_states rows are written
 * and MF_ERR_CAPACITY is returned).  out_b may be NULL to count only. */
int mf_lunnon_frontier(const int32_t *dims, int32_t ndim, int32_t depth,
                       uint8_t *out_b, int64_t max_states, int64_t *num_states);

/* Check caller-supplied states (no GPU needed): each must be a state the
 * folding DFS reaches (mf.c with its normal-form filter), i.e. leaves
 * depth, depth-1, ..., 2 each sit directly below a position the gap walks of
 * the smaller state allow, never below the root.  Every Lunnon entry point
 * that takes states applies this check; on failure *first_bad is the index
 * of the first rejected state. */
int mf_lunnon_check_states(const int32_t *dims, int32_t ndim, int32_t depth, const uint8_t *states_b,
                           int64_t num, int64_t *first_bad);

/* Resume `num` frontier states on the GPU and count all completions under
 * each (solve_state_device semantics: stop at the state's own level).
 * per_count / per_nodes (nullable) receive each state's count and inserts. */
int mf_lunnon_solve(const int32_t *dims, int32_t ndim, int32_t depth,
                    const uint8_t *states_b, int64_t num,
                    const mf_solve_options *opt,
                    uint64_t *per_count, uint64_t *per_nodes, mf_result *out);

/* gpu_folding <grid> <depth>: frontier + GPU solve; the total of `mf dims`. */
int mf_gpu_folding(const int32_t *dims, int32_t ndim, int32_t depth,
                   const mf_solve_options *opt, mf_result *out, int64_t *num_states);

/* Batched uneven instances: `nmaps` 2-D maps (dims2[2*k], dims2[2*k+1]), each
 * with its own frontier depth, all resumed in one queue.  per_map_count
 * receives each map's exact total (the value `mf p q` prints). */
int mf_lunnon_batch(const int32_t *dims2, const int32_t *depths, int32_t nmaps,
                    const mf_solve_options *opt, uint64_t *per_map_count, mf_result *out);

/* ---- symmetry-reduced cycle DFS (sym_fold.cpp) --------------------------- */
/* Case rows: the cycle successor of each cell, one byte per cell, in rows of
 * 64 bytes for n <= 8 and 128 bytes for n = 9 (cells past the row's level
 * are zero).  The seed center is c x c cells with c = 5 for n = 9, 4 for even
 * and 3 for odd n <= 8, and cells are numbered by the center-out spiral
 * (sym_fold.cpp:122-173, 704), unless mf_sym_set_numbering changed them.
 * The library reads no environment variable. */

/* Process-wide seed center and cell numbering of every sym entry point that
 * takes n (the sym_fold9 CLI sets them from MF_CENTER_N, MF_ORDER and
 * MF_LOCK_CENTER_N, experiments/9x9/sym_fold9.cpp:160-229,968-980).
 *   center: 0 = the default above, else 3..5 (checked against n: same parity,
 *           at most n; applies to n >= 6);
 *   order:  MF_ORDER_SPIRAL, or for odd n a ring-wise renumbering;
 *   lock_center_n: rings of radius <= (lock_center_n - 1) / 2 keep spiral
 *           order (sym_fold9's default is 3).
 * The mf_wasm_* entry points always use the spiral (their records are in
 * spiral order, mfah/wasm/sym_kernel.c:149-181). */
enum { MF_ORDER_SPIRAL = 0, MF_ORDER_REVERSE = 1, MF_ORDER_OPPOSITE = 2, MF_ORDER_AXES_FIRST = 3,
       MF_ORDER_CORNERS_FIRST = 4 };
int mf_sym_set_numbering(int32_t center, int32_t order, int32_t lock_center_n);
void mf_sym_get_numbering(int32_t *center, int32_t *order, int32_t *lock_center_n);

/* Seeds: deduplicated center cycles for 6 <= n <= 9 (orbit multipliers), in
 * generation order.  *raw = center cycles before dedup. */
int mf_sym_seeds(int32_t n, uint8_t *cycles, uint64_t *mults, int64_t max_seeds,
                 int64_t *num_seeds, int64_t *raw);

/* Expanded cases [start, start+max_cases) at level `expanded` in the global
 * stream order of sym_fold (seed order, then DFS order with children in
 * ascending insertion point).  Uses a per-seed case-count index, so the skip
 * costs O(#seeds + one seed's subtree) instead of O(start).  *num receives the
 * number written (fewer than max_cases at the end of the stream). */
int mf_sym_cases(int32_t n, int32_t expanded, uint64_t start, uint64_t max_cases,
                 uint8_t *cycles, uint64_t *mults, int64_t *num);

/* Total number of expanded cases at level `expanded`. */
int mf_sym_case_total(int32_t n, int32_t expanded, uint64_t *total);

/* A seed slice: seeds [seed_start, seed_start + seed_count) of the seed list
 * (clipped to it; MF_ALL_SEEDS = every seed from seed_start).  Case index 0
 * of a slice is the first case of seed seed_start, as with the reference's
 * MF_SEED_START / MF_SEED_COUNT. */
#define MF_ALL_SEEDS UINT64_MAX
int mf_sym_cases_slice(int32_t n, int32_t expanded, uint64_t seed_start, uint64_t seed_count,
                       uint64_t start, uint64_t max_cases, uint8_t *cycles, uint64_t *mults,
                       int64_t *num);
/* Cases at level `expanded` under each seed of the slice, in seed order.
 * *num = the slice's seed count; counts (nullable) needs room for it. */
int mf_sym_seed_case_counts(int32_t n, int32_t expanded, uint64_t seed_start, uint64_t seed_count,
                            uint64_t *counts, int64_t max_counts, int64_t *num);

/* Count `num` cases (case rows valid on cells [0, level)) on the GPU.
 * mults may be NULL (all 1).  out->count = sum mult*count (not times n^2). */
int mf_sym_solve(int32_t n, int32_t level, const uint8_t *cycles, const uint64_t *mults,
                 int64_t num, const mf_solve_options *opt,
                 uint64_t *per_count, uint64_t *per_nodes, mf_result *out);

/* sym_fold <n> [threads] [max_cases] [start_case] with MF_STREAM_CASES semantics
 * (expanded = MF_EXPANDED_N for n >= 7): the decimal answer as 128 bits. */
int mf_sym_fold(int32_t n, int32_t expanded, uint64_t start, uint64_t max_cases,
                const mf_solve_options *opt, uint64_t answer[2], mf_result *out);
/* mf_sym_fold over the case stream of a seed slice (sym_fold9 MF_SEED_START). */
int mf_sym_fold_slice(int32_t n, int32_t expanded, uint64_t seed_start, uint64_t seed_count,
                      uint64_t start, uint64_t max_cases, const mf_solve_options *opt,
                      uint64_t answer[2], mf_result *out);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The same solve as mf_sym_solve (re-split schedule,
 * host-driven rounds), but every device buffer beyond the caller's host
 * inputs and outputs lives in one static 64 MiB device array per GPU, whose
 * address mf_init looks up; the call allocates nothing.  Limits: at most
 * MF_NOALLOC_MAX_CASES cases; no opt->order; mf_init must have been called
 * (else MF_ERR_NO_DEVICE).  Calls that overlap on one GPU (several host
 * threads) race on the shared array. */
#define MF_NOALLOC_MAX_CASES 65536
int mf_sym_solve_noalloc(int32_t n, int32_t level, const uint8_t *cycles, const uint64_t *mults,
                         int64_t num, const mf_solve_options *opt,
                         uint64_t *per_count, uint64_t *per_nodes, mf_result *out);

/* ---- device-resident frontiers (inputs stay in HBM across solves) ------- */
typedef struct mf_frontier mf_frontier;
int mf_sym_frontier_create(int32_t n, int32_t level, const uint8_t *cycles, const uint64_t *mults,
                           int64_t num, mf_frontier **out);
int mf_lunnon_frontier_create(const int32_t *dims, int32_t ndim, int32_t depth,
                              const uint8_t *states_b, int64_t num, mf_frontier **out);
/* Batched uneven instances as a resident frontier (see mf_lunnon_batch). */
int mf_lunnon_batch_frontier_create(const int32_t *dims2, const int32_t *depths, int32_t nmaps,
                                    mf_frontier **out);
/* Split frontiers (used by the learned schedule, mapfold_cost.h).
 * mf_sym_expand: the children at level+1 of each case, in ascending child
 * order (one level of recurse_generate, sym_fold.cpp:252-265); *written =
 * total children, MF_ERR_CAPACITY if more than max_out.
 * mf_sym_pieces_create: a resident frontier of subtrees at mixed levels;
 * piece s (cycle through cells [0, levels[s])) adds mults[s] x its count to
 * per-item result origins[s] < num_origins.  The per_count / per_nodes arrays
 * of mf_frontier_solve then have num_origins entries. */
int mf_sym_expand(int32_t n, int32_t level, const uint8_t *cycles, int64_t count, uint8_t *out_cycles,
                  uint32_t *out_parent, int64_t max_out, int64_t *written);
int mf_sym_pieces_create(int32_t n, const uint8_t *cycles, const uint8_t *levels, const uint32_t *origins,
                         const uint64_t *mults, int64_t num, int64_t num_origins, mf_frontier **out);
/* Solve a resident frontier; per_count/per_nodes are host arrays (nullable);
 * opt->order, if given, must be a permutation of the items with
 * opt->order_len equal to their number. */
int mf_frontier_solve(mf_frontier *f, const mf_solve_options *opt,
                      uint64_t *per_count, uint64_t *per_nodes, mf_result *out);
/* Solve a resident Lunnon frontier and return each map's exact total. */
int mf_frontier_solve_maps(mf_frontier *f, const mf_solve_options *opt,
                           uint64_t *per_map_count, mf_result *out);
/* Work items resident in the frontier (cases, states or pieces). */
uint64_t mf_frontier_items(const mf_frontier *f);
void mf_frontier_destroy(mf_frontier *f);

/* ---- sym_kernel.c-compatible batch ABI (mfah wire format) --------------- */
/* 112-byte case records: cycle bytes [0, n^2) in spiral order, u64 LE
 * multiplier at byte 100.  Output words: u128 count LE, u64 nodes, u32 status
 * (mfah/wasm/sym_kernel.c:240-249).  The node definition matches that file:
 * one per recurse_count call, stop level included (sym_kernel.c:212-214). */
uint8_t *mf_wasm_input_ptr(void);
uint32_t *mf_wasm_output_ptr(void);
uint32_t mf_wasm_case_stride(void);
uint32_t mf_wasm_max_cases(void);
uint32_t mf_wasm_solve_cases(uint32_t n, uint32_t start_i, uint32_t stop_i, uint32_t case_count);

#ifdef __cplusplus
}
#endif

#endif /* MAPFOLD_H */
