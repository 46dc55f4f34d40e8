// abi.cpp -- libmapfold.so: the C ABI of include/mapfold.h.
//
// Host side of the hot path: builds work items from a frontier, keeps them in
// HBM, launches the HIP kernels on the library's stream and reduces the exact
// totals.  Re-split schedule: rounds of the persistent kernel; each round's
// spill queue is the next round's input until a round spills nothing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mapfold.h"
#include "mf_counters.hpp"
#include "mf_device.hpp"
#include "mf_host.hpp"

namespace mf {
hipError_t launch_sym(const SymArgs &, int core, bool persistent, uint32_t grid, uint32_t block, uint32_t lds,
                      bool events, uint32_t cells, hipStream_t);
uint32_t sym_lds_bytes(int core, uint32_t levels, uint32_t block, uint32_t cells);
const void *sym_kernel_ptr(int core, bool persistent, bool events, uint32_t cells);
hipError_t launch_lunnon(const LunnonArgs &, int core, bool persistent, uint32_t grid, uint32_t block,
                         uint32_t lds, bool events, hipStream_t);
uint32_t lunnon_lds_bytes(int core, uint32_t tables_bytes, uint32_t levels, uint32_t block, uint32_t rows);
const void *lunnon_kernel_ptr(int core, bool persistent, bool events, uint32_t dims);
uint32_t lunnon_dims(const LunnonArgs &a);
bool lunnon_paired(int core, const LunnonArgs &a);
hipError_t launch_round_reset(uint32_t *claim, uint32_t *out_count, hipStream_t s);
const void *sym_deepen_ptr(int core, bool events, uint32_t cells);
int sym_walk_core(int core, const SymArgs &a, uint32_t cells);
hipError_t launch_sym_features(const SymArgs &a, float *out, hipStream_t s);
}  // namespace mf

using namespace mf;

namespace {

std::string g_err = "";

int fail(int status, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return status;
}

#define HIP_TRY(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(MF_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));       \
    } while (0)

struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    int ensure(size_t want) {
        if (want <= bytes) return MF_OK;
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&ptr, want);
        if (e != hipSuccess) return fail(MF_ERR_HIP, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
        bytes = want;
        return MF_OK;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

struct Context {
    int device = -1;
    int cus = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;   // span of a solve
    hipEvent_t la = nullptr, lb = nullptr;     // one launch
    DevBuf items, spill_a, spill_b, acc, counters, per_count, per_nodes, per_map, tables;
    DevBuf deep;                               // deepened frontier (sym)
    DevBuf targets;                            // per-item deepening targets
    unsigned long long *host_acc = nullptr;   // pinned
    uint32_t *host_counters = nullptr;        // pinned: [claim, count of spill_a, count of spill_b]
    // Captured re-split rounds, reused while the launch configuration holds.
    hipGraphExec_t graph = nullptr;
    std::vector<uint8_t> graph_key;
};

Context g;

int require_init() {
    if (g.device < 0) {
        int rc = mf_init(0);
        if (rc != MF_OK) return rc;
    }
    return MF_OK;
}

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

constexpr uint32_t kGraphRounds = 8;   // rounds per captured graph (even: replays keep buffer parity)

struct Tuned {
    int schedule;
    int core;
    bool graph;
    uint32_t block;
    uint32_t waves_per_cu;
    uint32_t budget;
    uint32_t tail;
    bool events;
    uint64_t capacity;
    int deepen;   // sym re-split: target level of the device deepening (0 default, -1 off)
};

constexpr int kDefaultCore = 2;   // compact 16-bit stack

Tuned tune(const mf_solve_options *opt) {
    Tuned t;
    t.schedule = opt ? opt->schedule : MF_SCHED_RESPLIT;
    t.core = (opt && opt->core >= 1 && opt->core <= 4) ? opt->core : kDefaultCore;
    t.block = (opt && opt->block > 0) ? static_cast<uint32_t>(opt->block) : 256u;
    t.block = std::max(64u, std::min(256u, (t.block / 64u) * 64u));
    t.waves_per_cu = (opt && opt->waves_per_cu > 0) ? static_cast<uint32_t>(opt->waves_per_cu) : 32u;
    // A step expands one node (or recomputes one).
    t.budget = (opt && opt->budget > 0) ? static_cast<uint32_t>(opt->budget) : 64u;
    // Per-item event counters are folded into 64 bits every `budget` steps.
    t.budget = std::min(t.budget, kMaxFoldBudget);
    t.tail = (opt && opt->tail > 0) ? static_cast<uint32_t>(opt->tail) : t.budget;
    t.events = opt ? opt->count_events != 0 : true;
    t.capacity = (opt && opt->queue_capacity > 0) ? opt->queue_capacity : 0;
    t.graph = !(opt && opt->graph == 2);
    t.deepen = opt ? opt->deepen : 0;
    return t;
}

// Persistent grid: as many blocks as fit (LDS- and wave-limited) on every CU.
int persistent_grid(const void *kernel, uint32_t block, uint32_t lds, uint32_t waves_per_cu, uint32_t *grid) {
    int per_cu = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, static_cast<int>(block), lds));
    if (per_cu <= 0) return fail(MF_ERR_ARG, "kernel does not fit: block %u with %u bytes of LDS", block, lds);
    const uint32_t by_waves = std::max(1u, waves_per_cu / (block / 64u));
    const uint32_t per = std::min(static_cast<uint32_t>(per_cu), by_waves);
    *grid = per * static_cast<uint32_t>(g.cus);
    return MF_OK;
}

int set_lds_limit(const void *kernel, uint32_t lds) {
    HIP_TRY(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    return MF_OK;
}

void fill_result(mf_result *out, const unsigned long long *acc) {
    out->count_lo = acc[ACC_COUNT_LO];
    out->count_hi = acc[ACC_COUNT_HI];
    out->nodes = acc[ACC_NODES];
    out->walks = acc[ACC_WALKS];
    out->steps = acc[ACC_STEPS];
    out->leaves = acc[ACC_LEAVES];
    out->cands = acc[ACC_CANDS];
    out->items = acc[ACC_ITEMS];
}

// Generic solve driver over a device-resident item array.
//   launch(args, persistent, grid) runs one kernel launch on g.stream.
template <typename Item, typename Args, typename Launch>
int run_solve(Args args, const void *kstatic, const void *kpersist, uint32_t lds_bytes, const Tuned &t,
              const Item *dev_items, uint64_t num, uint64_t origins, uint64_t *per_count, uint64_t *per_nodes,
              uint64_t *per_map, uint32_t nmaps, mf_result *out, bool keep_acc, Launch launch) {
    const auto t0 = std::chrono::steady_clock::now();
    std::memset(out, 0, sizeof(*out));
    int rc;
    if ((rc = g.acc.ensure(ACC_SLOTS * sizeof(unsigned long long))) != MF_OK) return rc;
    if ((rc = g.counters.ensure(4 * sizeof(uint32_t))) != MF_OK) return rc;
    // keep_acc: a deepening pass has already counted into the accumulators
    // and the per-origin arrays (cleared by it).
    if (!keep_acc) HIP_TRY(hipMemsetAsync(g.acc.ptr, 0, ACC_SLOTS * sizeof(unsigned long long), g.stream));
    unsigned long long *dcount = nullptr, *dnodes = nullptr, *dmap = nullptr;
    if (per_count) {
        if ((rc = g.per_count.ensure(origins * 8)) != MF_OK) return rc;
        if (!keep_acc) HIP_TRY(hipMemsetAsync(g.per_count.ptr, 0, origins * 8, g.stream));
        dcount = static_cast<unsigned long long *>(g.per_count.ptr);
    }
    if (per_nodes) {
        if ((rc = g.per_nodes.ensure(origins * 8)) != MF_OK) return rc;
        if (!keep_acc) HIP_TRY(hipMemsetAsync(g.per_nodes.ptr, 0, origins * 8, g.stream));
        dnodes = static_cast<unsigned long long *>(g.per_nodes.ptr);
    }
    if (per_map) {
        if ((rc = g.per_map.ensure(static_cast<size_t>(nmaps) * 8)) != MF_OK) return rc;
        HIP_TRY(hipMemsetAsync(g.per_map.ptr, 0, static_cast<size_t>(nmaps) * 8, g.stream));
        dmap = static_cast<unsigned long long *>(g.per_map.ptr);
    }
    args.acc = static_cast<unsigned long long *>(g.acc.ptr);
    args.per_count = dcount;
    args.per_nodes = dnodes;
    args.budget = t.budget;
    args.tail = t.tail;
    uint32_t *claim = static_cast<uint32_t *>(g.counters.ptr);
    uint32_t *spill_count = claim + 1;
    args.claim = claim;
    args.spill_count = spill_count;

    uint32_t rounds = 0;
    double launch_ms = 0.0;
    if (num > 0) {
        out->inputs = num;
        if (t.schedule == MF_SCHED_STATIC) {
            if ((rc = set_lds_limit(kstatic, lds_bytes)) != MF_OK) return rc;
            args.items = dev_items;
            args.num_items = static_cast<uint32_t>(num);
            args.spill = nullptr;
            args.spill_cap = 0;
            const uint32_t grid = static_cast<uint32_t>((num + t.block - 1) / t.block);
            HIP_TRY(hipEventRecord(g.ev0, g.stream));
            HIP_TRY(launch(args, false, grid));
            HIP_TRY(hipEventRecord(g.ev1, g.stream));
            rounds = 1;
            // One launch bounded by its longest item: it can run for minutes,
            // so wait by polling its end event instead of a spinning
            // synchronize that would hold a host core the whole time.
            for (;;) {
                const hipError_t q = hipEventQuery(g.ev1);
                if (q == hipSuccess) break;
                if (q != hipErrorNotReady) return fail(MF_ERR_HIP, "hipEventQuery failed: %s", hipGetErrorString(q));
                std::this_thread::sleep_for(std::chrono::milliseconds(1));
            }
        } else {
            if ((rc = set_lds_limit(kpersist, lds_bytes)) != MF_OK) return rc;
            uint32_t grid = 0;
            if ((rc = persistent_grid(kpersist, t.block, lds_bytes, t.waves_per_cu, &grid)) != MF_OK) return rc;
            const uint64_t lanes = static_cast<uint64_t>(grid) * t.block;
            // A lane spills at most once per round: its current node plus one
            // sibling set per stack level.
            uint64_t cap = t.capacity ? t.capacity : lanes * 65ull;
            cap = std::max<uint64_t>(cap, lanes * 2);
            if (cap > 0xffffffffull) cap = 0xffffffffull;
            if ((rc = g.spill_a.ensure(cap * sizeof(Item))) != MF_OK) return rc;
            if ((rc = g.spill_b.ensure(cap * sizeof(Item))) != MF_OK) return rc;
            Item *bufs[2] = {static_cast<Item *>(g.spill_a.ptr), static_cast<Item *>(g.spill_b.ptr)};
            uint32_t *counts = claim + 1;   // counts[q]: items written to bufs[q] by the last round into it
            const Item *in = dev_items;
            uint64_t in_num = num;
            int next = 0;
            auto check_errors = [&]() -> int {
                if (g.host_acc[ACC_ERROR] & 6ull)
                    return fail(MF_ERR_INTERNAL, "device consistency check failed (%llu)", g.host_acc[ACC_ERROR]);
                if (g.host_acc[ACC_ERROR] != 0)
                    return fail(MF_ERR_CAPACITY, "spill queue overflow (capacity %llu items)",
                                static_cast<unsigned long long>(cap));
                return MF_OK;
            };
            HIP_TRY(hipEventRecord(g.ev0, g.stream));
            // Host-driven rounds: the first one always (its input is the
            // frontier), every one when graphs are off.
            while (in_num > 0 && (!t.graph || rounds == 0)) {
                float one = 0.0f;
                if (rounds >= 100000) return fail(MF_ERR_INTERNAL, "re-split did not converge");
                HIP_TRY(hipMemsetAsync(claim, 0, sizeof(uint32_t), g.stream));
                HIP_TRY(hipMemsetAsync(counts + next, 0, sizeof(uint32_t), g.stream));
                args.items = in;
                args.num_items = static_cast<uint32_t>(in_num);
                args.num_items_dev = nullptr;
                args.spill = bufs[next];
                args.spill_count = counts + next;
                args.spill_cap = static_cast<uint32_t>(cap);
                HIP_TRY(hipEventRecord(g.la, g.stream));
                HIP_TRY(launch(args, true, grid));
                HIP_TRY(hipEventRecord(g.lb, g.stream));
                HIP_TRY(hipMemcpyAsync(g.host_counters, counts + next, sizeof(uint32_t), hipMemcpyDeviceToHost, g.stream));
                HIP_TRY(hipMemcpyAsync(g.host_acc + ACC_ERROR, args.acc + ACC_ERROR, 8, hipMemcpyDeviceToHost, g.stream));
                HIP_TRY(hipStreamSynchronize(g.stream));
                HIP_TRY(hipEventElapsedTime(&one, g.la, g.lb));
                launch_ms += one;
                ++rounds;
                if ((rc = check_errors()) != MF_OK) return rc;
                in = bufs[next];
                in_num = g.host_counters[0];
                next ^= 1;
            }
            if (in_num > 0) {
                // Graph rounds: kGraphRounds launches captured once, each
                // preceded by a reset kernel and reading its item count from
                // the previous round's spill counter; replayed until the last
                // round of a replay spills nothing (trailing empty rounds exit
                // at once).  kGraphRounds is even, so every replay starts on
                // the same buffer.
                const int p0 = (in == bufs[0]) ? 0 : 1;
                Args ga[2];
                for (int q = 0; q < 2; ++q) {
                    ga[q] = args;
                    ga[q].items = bufs[q];
                    ga[q].num_items = 0;
                    ga[q].num_items_dev = counts + q;
                    ga[q].spill = bufs[q ^ 1];
                    ga[q].spill_count = counts + (q ^ 1);
                    ga[q].spill_cap = static_cast<uint32_t>(cap);
                }
                std::vector<uint8_t> key(2 * sizeof(Args) + sizeof(void *) + 4 * sizeof(uint32_t));
                uint8_t *kp = key.data();
                std::memcpy(kp, &ga[0], sizeof(Args));
                std::memcpy(kp + sizeof(Args), &ga[1], sizeof(Args));
                const uint32_t cfg[4] = {static_cast<uint32_t>(p0), grid, t.block, lds_bytes};
                std::memcpy(kp + 2 * sizeof(Args), &kpersist, sizeof(void *));
                std::memcpy(kp + 2 * sizeof(Args) + sizeof(void *), cfg, sizeof(cfg));
                if (g.graph == nullptr || key != g.graph_key) {
                    if (g.graph) (void)hipGraphExecDestroy(g.graph);
                    g.graph = nullptr;
                    g.graph_key.clear();
                    hipGraph_t graph = nullptr;
                    HIP_TRY(hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal));
                    for (uint32_t k = 0; k < kGraphRounds; ++k) {
                        const int q = p0 ^ static_cast<int>(k & 1u);
                        hipError_t e = launch_round_reset(claim, counts + (q ^ 1), g.stream);
                        if (e == hipSuccess) e = launch(ga[q], true, grid);
                        if (e != hipSuccess) {
                            (void)hipStreamEndCapture(g.stream, &graph);
                            if (graph) (void)hipGraphDestroy(graph);
                            return fail(MF_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
                        }
                    }
                    HIP_TRY(hipStreamEndCapture(g.stream, &graph));
                    hipGraphExec_t exec = nullptr;
                    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(graph);
                    if (e != hipSuccess) return fail(MF_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
                    g.graph = exec;
                    g.graph_key = key;
                }
                while (in_num > 0) {
                    float one = 0.0f;
                    if (rounds >= 100000) return fail(MF_ERR_INTERNAL, "re-split did not converge");
                    HIP_TRY(hipEventRecord(g.la, g.stream));
                    HIP_TRY(hipGraphLaunch(g.graph, g.stream));
                    HIP_TRY(hipEventRecord(g.lb, g.stream));
                    HIP_TRY(hipMemcpyAsync(g.host_counters, counts + p0, sizeof(uint32_t), hipMemcpyDeviceToHost,
                                           g.stream));
                    HIP_TRY(hipMemcpyAsync(g.host_acc + ACC_ERROR, args.acc + ACC_ERROR, 8, hipMemcpyDeviceToHost,
                                           g.stream));
                    HIP_TRY(hipStreamSynchronize(g.stream));
                    HIP_TRY(hipEventElapsedTime(&one, g.la, g.lb));
                    launch_ms += one;
                    rounds += kGraphRounds;
                    if ((rc = check_errors()) != MF_OK) return rc;
                    in_num = g.host_counters[0];
                }
            }
            HIP_TRY(hipEventRecord(g.ev1, g.stream));
        }
    }
    HIP_TRY(hipMemcpyAsync(g.host_acc, g.acc.ptr, ACC_SLOTS * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           g.stream));
    if (per_count) HIP_TRY(hipMemcpyAsync(per_count, dcount, origins * 8, hipMemcpyDeviceToHost, g.stream));
    if (per_nodes) HIP_TRY(hipMemcpyAsync(per_nodes, dnodes, origins * 8, hipMemcpyDeviceToHost, g.stream));
    if (per_map) HIP_TRY(hipMemcpyAsync(per_map, dmap, static_cast<size_t>(nmaps) * 8, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    if (g.host_acc[ACC_ERROR] & 6ull)
        return fail(MF_ERR_INTERNAL, "device consistency check failed (%llu)", g.host_acc[ACC_ERROR]);
    if (g.host_acc[ACC_ERROR] != 0) return fail(MF_ERR_CAPACITY, "work queue overflow");
    fill_result(out, g.host_acc);
    // Persistent launches that had input (graph replays also run empty rounds).
    out->rounds = (t.schedule == MF_SCHED_STATIC) ? rounds : static_cast<uint32_t>(g.host_acc[ACC_ROUNDS]);
    if (num > 0) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, g.ev0, g.ev1));
        out->span_ms = ms;
        out->kernel_ms = (t.schedule == MF_SCHED_STATIC) ? ms : launch_ms;
    }
    out->total_ms = ms_since(t0);
    return MF_OK;
}

// ------------------------------------------------------------------ frontiers

}  // namespace

struct mf_frontier {
    int kind = 0;                  // 1 = sym (up to 8x8), 2 = Lunnon, 3 = sym 9x9 (u128 masks)
    SymProblem sym;
    int level = 0;
    int stop = 0;                  // sym: paths are counted at this level (n2 = full foldings)
    std::vector<LunnonProblem> maps;
    std::vector<int> depths;
    DevBuf items;
    uint64_t num = 0;                   // work items
    uint64_t origins = 0;               // entries of the per-item result arrays (0 = num)
    std::vector<WorkItem> host_items;   // kept for learned reordering
    std::vector<WorkItem128> host_items128;   // kind 3 (9x9 sym)
};

namespace {

int sym_args(const mf_frontier &f, SymArgs &a, uint32_t *levels) {
    std::memset(&a, 0, sizeof(a));
    std::memcpy(a.edge_word, f.sym.edge_word, sizeof(a.edge_word));
    a.last = static_cast<uint32_t>((f.stop > 0 ? f.stop : f.sym.n2) - 1);
    a.base_level = static_cast<uint32_t>(f.level);
    *levels = a.last - a.base_level;
    if (*levels == 0) *levels = 1;
    a.levels = *levels;
    return MF_OK;
}

// Build the Lunnon table blob and map descriptors; returns stack levels.
int lunnon_args(const mf_frontier &f, LunnonArgs &a, std::vector<uint8_t> &blob, uint32_t *levels) {
    std::memset(&a, 0, sizeof(a));
    if (f.maps.size() > static_cast<size_t>(kMaxMaps)) return fail(MF_ERR_ARG, "at most %d maps per launch", kMaxMaps);
    blob.clear();
    uint32_t lv = 1;
    for (size_t k = 0; k < f.maps.size(); ++k) {
        const LunnonProblem &pr = f.maps[k];
        LunnonMap m;
        // Maps of the same shape share one table (LDS holds every table of
        // the launch, so duplicates would cost occupancy).
        size_t same = k;
        for (size_t q = 0; q < k; ++q)
            if (f.maps[q].d == pr.d) {
                same = q;
                break;
            }
        m.table_off = same < k ? a.maps[same].table_off : static_cast<uint32_t>(blob.size());
        m.n = static_cast<uint16_t>(pr.n);
        m.ndim = static_cast<uint8_t>(pr.ndim);
        m.base = static_cast<uint8_t>(f.depths[k] + 1);
        a.maps[k] = m;
        if (same == k)   // lower triangle of each dimension: d[i][l][m], m <= l
            for (int i = 0; i < pr.ndim; ++i)
                for (int l = 0; l <= pr.n; ++l)
                    for (int mm = 0; mm <= l; ++mm) blob.push_back(pr.at(i, l, mm));
        lv = std::max<uint32_t>(lv, static_cast<uint32_t>(pr.n - f.depths[k] - 1));
    }
    a.num_maps = static_cast<uint32_t>(f.maps.size());
    a.tables_bytes = static_cast<uint32_t>(blob.size());
    // Leaves 0..n-1 of the largest map: leaf n is never linked (inserting it
    // counts completions), so a row needs n entries (n + 1 when the launch
    // walks both dimensions in one loop, which points a leaf's own entry at
    // itself while it computes the leaf's gaps: solve_frontier).
    a.row_entries = 0;
    for (const LunnonProblem &pr : f.maps) a.row_entries = std::max<uint32_t>(a.row_entries, pr.n);
    a.levels = lv;
    *levels = lv;
    return MF_OK;
}

template <typename Item>
int upload_vec(mf_frontier *f, const std::vector<Item> &v) {
    int rc;
    if ((rc = f->items.ensure(std::max<size_t>(1, v.size()) * sizeof(Item))) != MF_OK) return rc;
    if (!v.empty())
        HIP_TRY(hipMemcpyAsync(f->items.ptr, v.data(), v.size() * sizeof(Item), hipMemcpyHostToDevice, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    f->num = v.size();
    return MF_OK;
}

int upload_items(mf_frontier *f) {
    return f->kind == 3 ? upload_vec(f, f->host_items128) : upload_vec(f, f->host_items);
}

// The caller's order as a permutation of the host items (a repeated item would
// be counted twice), uploaded to `dst`.
template <typename Item>
int reorder_items(const std::vector<Item> &src, const uint32_t *order, DevBuf &dst, const Item **dev) {
    const uint64_t num = src.size();
    std::vector<Item> tmp(num);
    std::vector<uint8_t> seen(num, 0);
    for (uint64_t k = 0; k < num; ++k) {
        const uint32_t from = order[k];
        if (from >= num || seen[from])
            return fail(MF_ERR_ARG, "order is not a permutation (entry %llu)", static_cast<unsigned long long>(k));
        seen[from] = 1;
        tmp[k] = src[from];
    }
    int rc;
    if ((rc = dst.ensure(num * sizeof(Item))) != MF_OK) return rc;
    HIP_TRY(hipMemcpyAsync(dst.ptr, tmp.data(), num * sizeof(Item), hipMemcpyHostToDevice, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    *dev = static_cast<const Item *>(dst.ptr);
    return MF_OK;
}

// Default deepening target of a sym frontier (0 = none): 8x8 frontiers are
// expanded to level 46, which leaves 17 stack levels per lane, so that six
// 4-wave workgroups fit the LDS of a CU instead of five at depth 36 (8x8
// chunk of 50,000 depth-36 cases: 908 ms at 46, 912 at 47, 917 at 48, 988 at
// 45, where the block is just too large for six, 1001 without).
int default_deepen(const mf_frontier &f) { return f.sym.n2 == 64 ? 46 : 0; }

constexpr int kDeepenFull = 100;   // internal: a deepening pass ran out of room

// Expand the frontier's expand items on the device until each reaches level
// `target` (sym_deepen_kernel, one launch per level), counting every expanded
// node into the accumulators and per-origin node array, which it clears
// first.  The deepened items land in g.deep.  Returns kDeepenFull if the
// buffers cannot hold a level (the caller then solves the frontier as it is).
int deepen_frontier(const SymArgs &base, int core, const Tuned &t, uint32_t cells, const WorkItem *in,
                    uint64_t num, uint64_t origins, bool want_count, bool want_nodes, uint32_t target,
                    const uint8_t *item_targets, const WorkItem **out_items, uint64_t *out_num, double *ms) {
    int rc;
    const void *kd = sym_deepen_ptr(core, t.events, cells);
    if (kd == nullptr) return kDeepenFull;
    if ((rc = g.acc.ensure(ACC_SLOTS * sizeof(unsigned long long))) != MF_OK) return rc;
    if ((rc = g.counters.ensure(4 * sizeof(uint32_t))) != MF_OK) return rc;
    HIP_TRY(hipMemsetAsync(g.acc.ptr, 0, ACC_SLOTS * sizeof(unsigned long long), g.stream));
    SymArgs a = base;
    a.acc = static_cast<unsigned long long *>(g.acc.ptr);
    a.per_count = nullptr;
    a.per_nodes = nullptr;
    if (want_count) {
        if ((rc = g.per_count.ensure(origins * 8)) != MF_OK) return rc;
        HIP_TRY(hipMemsetAsync(g.per_count.ptr, 0, origins * 8, g.stream));
    }
    if (want_nodes) {
        if ((rc = g.per_nodes.ensure(origins * 8)) != MF_OK) return rc;
        HIP_TRY(hipMemsetAsync(g.per_nodes.ptr, 0, origins * 8, g.stream));
        a.per_nodes = static_cast<unsigned long long *>(g.per_nodes.ptr);
    }
    // Room for the deepened frontier and for one level in flight (the spill
    // queues, which the rounds reuse afterwards).
    const uint64_t cap = std::min<uint64_t>(0xffffffffull, std::max<uint64_t>(num * 8, uint64_t{32} << 20));
    if ((rc = g.deep.ensure(cap * sizeof(WorkItem))) != MF_OK) return rc;
    if ((rc = g.spill_a.ensure(cap * sizeof(WorkItem))) != MF_OK) return rc;
    if ((rc = g.spill_b.ensure(cap * sizeof(WorkItem))) != MF_OK) return rc;
    uint32_t *counts = static_cast<uint32_t *>(g.counters.ptr);   // [1], [2]: next levels, [3]: final
    WorkItem *bufs[2] = {static_cast<WorkItem *>(g.spill_a.ptr), static_cast<WorkItem *>(g.spill_b.ptr)};
    const uint32_t lds = sym_lds_bytes(core, 1, 256, cells);
    HIP_TRY(hipFuncSetAttribute(kd, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    a.final_out = g.deep.ptr;
    a.final_count = counts + 3;
    a.final_cap = static_cast<uint32_t>(cap);
    a.spill_cap = static_cast<uint32_t>(cap);
    a.target = target;
    a.targets = nullptr;
    if (item_targets != nullptr) {
        if ((rc = g.targets.ensure(std::max<uint64_t>(num, 1))) != MF_OK) return rc;
        HIP_TRY(hipMemcpyAsync(g.targets.ptr, item_targets, num, hipMemcpyHostToDevice, g.stream));
        a.targets = static_cast<const uint8_t *>(g.targets.ptr);
    }
    HIP_TRY(hipMemsetAsync(counts + 3, 0, sizeof(uint32_t), g.stream));
    HIP_TRY(hipEventRecord(g.la, g.stream));
    const WorkItem *cur = in;
    uint64_t cur_num = num;
    int next = 0;
    for (uint32_t pass = 0; cur_num > 0; ++pass) {
        if (pass > 128) return fail(MF_ERR_INTERNAL, "deepening did not converge");
        HIP_TRY(hipMemsetAsync(counts + 1 + next, 0, sizeof(uint32_t), g.stream));
        a.items = cur;
        a.num_items = static_cast<uint32_t>(cur_num);
        a.spill = bufs[next];
        a.spill_count = counts + 1 + next;
        a.deep_pass = pass == 0 ? 0u : 1u;
        void *args[] = {&a};
        HIP_TRY(hipLaunchKernel(kd, dim3(static_cast<uint32_t>((cur_num + 255) / 256)), dim3(256), args, lds, g.stream));
        HIP_TRY(hipMemcpyAsync(g.host_counters, counts, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, g.stream));
        HIP_TRY(hipMemcpyAsync(g.host_acc + ACC_ERROR, a.acc + ACC_ERROR, 8, hipMemcpyDeviceToHost, g.stream));
        HIP_TRY(hipStreamSynchronize(g.stream));
        if (g.host_acc[ACC_ERROR] & 6ull) return fail(MF_ERR_INTERNAL, "device consistency check failed");
        if (g.host_acc[ACC_ERROR] & kErrDeepenFull) return kDeepenFull;
        cur = bufs[next];
        cur_num = g.host_counters[1 + next];
        next ^= 1;
    }
    HIP_TRY(hipEventRecord(g.lb, g.stream));
    HIP_TRY(hipEventSynchronize(g.lb));
    float one = 0.0f;
    HIP_TRY(hipEventElapsedTime(&one, g.la, g.lb));
    *ms = one;
    *out_items = static_cast<const WorkItem *>(g.deep.ptr);
    *out_num = g.host_counters[3];
    return MF_OK;
}

int solve_frontier(mf_frontier *f, const mf_solve_options *opt, uint64_t *per_count, uint64_t *per_nodes,
                   uint64_t *per_map, mf_result *out) {
    int rc = require_init();
    if (rc != MF_OK) return rc;
    const Tuned t = tune(opt);
    // Learned / caller order: permute the resident items on the host side;
    // origins stay attached to each item so results are unchanged.
    DevBuf reordered;
    const bool ordered = opt && opt->order && f->num > 0;
    if (opt && opt->order && opt->order_len != f->num)
        return fail(MF_ERR_ARG, "order has %llu entries, the frontier %llu items",
                    static_cast<unsigned long long>(opt->order_len), static_cast<unsigned long long>(f->num));
    if (f->kind == 1 || f->kind == 3) {
        const bool c96 = f->kind == 3;
        if (!c96 && t.core == 3 && t.schedule != MF_SCHED_RESPLIT)
            return fail(MF_ERR_ARG, "core 3 (two lanes per subtree) runs the re-split schedule only");
        SymArgs a;
        uint32_t levels = 0;
        sym_args(*f, a, &levels);
        // Numberings whose levels do not all walk one or two pages run one
        // page walk at a time.
        const int core = sym_walk_core(c96 && t.core != 4 ? 2 : t.core, a, static_cast<uint32_t>(f->sym.n2));
        const uint32_t cells = static_cast<uint32_t>(f->sym.n2);
        const uint32_t lds = sym_lds_bytes(core, levels, t.block, cells);
        auto launch = [&](const SymArgs &args, bool persistent, uint32_t grid) {
            return launch_sym(args, core, persistent, grid, t.block, lds, t.events, cells, g.stream);
        };
        const void *ks = sym_kernel_ptr(core, false, t.events, cells), *kp = sym_kernel_ptr(core, true, t.events, cells);
        const uint64_t origins = f->origins ? f->origins : f->num;
        if (c96) {
            const WorkItem128 *dev = static_cast<const WorkItem128 *>(f->items.ptr);
            if (ordered && (rc = reorder_items(f->host_items128, opt->order, reordered, &dev)) != MF_OK) return rc;
            rc = run_solve(a, ks, kp, lds, t, dev, f->num, origins, per_count, per_nodes, nullptr, 0, out, false, launch);
        } else {
            const WorkItem *dev = static_cast<const WorkItem *>(f->items.ptr);
            if (ordered && (rc = reorder_items(f->host_items, opt->order, reordered, &dev)) != MF_OK) return rc;
            // Re-split schedule: deepen the frontier on the device first, so
            // that a lane needs fewer stack levels and more waves fit a CU.
            int target = t.deepen == 0 ? default_deepen(*f) : t.deepen;
            // Per-item targets (in the order the items are solved), at least
            // the uniform one, at most n*n - 2.
            std::vector<uint8_t> item_targets;
            if (opt && opt->deepen_levels) {
                if (opt->deepen_levels_len != f->num)
                    return fail(MF_ERR_ARG, "deepen_levels has %llu entries, the frontier %llu items",
                                static_cast<unsigned long long>(opt->deepen_levels_len),
                                static_cast<unsigned long long>(f->num));
                if (target < f->level) target = f->level;
                item_targets.resize(f->num);
                for (uint64_t k = 0; k < f->num; ++k) {
                    const uint64_t from = ordered ? opt->order[k] : k;
                    const int want = std::max<int>(opt->deepen_levels[from], target);
                    item_targets[k] = static_cast<uint8_t>(std::min<int>(want, static_cast<int>(a.last) - 1));
                }
                if (t.deepen < 0) item_targets.clear();
            }
            bool done = false;
            const bool per_item = !item_targets.empty();
            if (t.schedule == MF_SCHED_RESPLIT && (target > f->level || per_item) &&
                target <= static_cast<int>(a.last) - 1 && f->num > 0) {
                const WorkItem *deep = nullptr;
                uint64_t deep_num = 0;
                double deep_ms = 0.0;
                // The deepening pass runs one lane per item (the compact-stack instance).
                const int dcore = core == 3 ? sym_walk_core(2, a, cells) : core;
                rc = deepen_frontier(a, dcore, t, cells, dev, f->num, origins, per_count != nullptr,
                                     per_nodes != nullptr, static_cast<uint32_t>(target),
                                     per_item ? item_targets.data() : nullptr, &deep, &deep_num, &deep_ms);
                if (rc == kDeepenFull && per_item && target > f->level) {
                    // The per-item targets outgrew the buffers: deepen every
                    // item to the uniform target alone (the pass starts over,
                    // clearing what the first attempt counted).
                    rc = deepen_frontier(a, dcore, t, cells, dev, f->num, origins, per_count != nullptr,
                                         per_nodes != nullptr, static_cast<uint32_t>(target), nullptr, &deep,
                                         &deep_num, &deep_ms);
                }
                if (rc == MF_OK) {
                    SymArgs a2 = a;
                    // Every deepened item is at or past the uniform target
                    // (or its own level, if that was higher).
                    a2.base_level = static_cast<uint32_t>(std::max(target, f->level));
                    a2.levels = std::max<uint32_t>(1u, a.last - a2.base_level);
                    const uint32_t lds2 = sym_lds_bytes(core, a2.levels, t.block, cells);
                    auto launch2 = [&](const SymArgs &args, bool persistent, uint32_t grid) {
                        return launch_sym(args, core, persistent, grid, t.block, lds2, t.events, cells, g.stream);
                    };
                    const auto t0 = std::chrono::steady_clock::now();
                    rc = run_solve(a2, ks, kp, lds2, t, deep, deep_num, origins, per_count, per_nodes, nullptr, 0, out,
                                   true, launch2);
                    if (rc == MF_OK) {
                        out->kernel_ms += deep_ms;
                        out->span_ms += deep_ms;
                        out->total_ms = ms_since(t0) + deep_ms;
                    }
                    done = true;
                } else if (rc != kDeepenFull) {
                    return rc;
                }
            }
            if (!done)
                rc = run_solve(a, ks, kp, lds, t, dev, f->num, origins, per_count, per_nodes, nullptr, 0, out, false,
                               launch);
        }
    } else {
        LunnonArgs a;
        std::vector<uint8_t> blob;
        uint32_t levels = 0;
        if ((rc = lunnon_args(*f, a, blob, &levels)) != MF_OK) return rc;
        if ((rc = g.tables.ensure(std::max<size_t>(blob.size(), 16))) != MF_OK) return rc;
        HIP_TRY(hipMemcpyAsync(g.tables.ptr, blob.data(), blob.size(), hipMemcpyHostToDevice, g.stream));
        a.tables = static_cast<const uint8_t *>(g.tables.ptr);
        a.per_map = nullptr;
        if (lunnon_paired(t.core, a)) ++a.row_entries;   // at most 64: paired launches have n <= 63
        const uint32_t lds = lunnon_lds_bytes(t.core, a.tables_bytes, levels, t.block, a.row_entries);
        if (per_map) {
            if ((rc = g.per_map.ensure(f->maps.size() * 8)) != MF_OK) return rc;
            a.per_map = static_cast<unsigned long long *>(g.per_map.ptr);
        }
        auto launch = [&](const LunnonArgs &args, bool persistent, uint32_t grid) {
            return launch_lunnon(args, t.core, persistent, grid, t.block, lds, t.events, g.stream);
        };
        const WorkItem *dev_items = static_cast<const WorkItem *>(f->items.ptr);
        if (ordered && (rc = reorder_items(f->host_items, opt->order, reordered, &dev_items)) != MF_OK) return rc;
        const uint32_t dims = lunnon_dims(a);
        rc = run_solve(a, lunnon_kernel_ptr(t.core, false, t.events, dims), lunnon_kernel_ptr(t.core, true, t.events, dims),
                       lds,
                       t, dev_items,
                       f->num, f->origins ? f->origins : f->num, per_count, per_nodes, per_map,
                       static_cast<uint32_t>(f->maps.size()), out, false,
                       launch);
    }
    reordered.release();
    return rc;
}

// A case must be one cycle through exactly the cells [0, level): then every
// get_good walk is an orbit of a permutation and terminates.
bool valid_cycle(const uint8_t *cycle, int level) {
    using u128 = unsigned __int128;
    if (level < 1 || level > 127) return false;
    u128 seen = 0;
    int c = 0;
    for (int k = 0; k < level; ++k) {
        if (cycle[c] >= level) return false;
        seen |= static_cast<u128>(1) << c;
        c = cycle[c];
    }
    return c == 0 && seen == ((static_cast<u128>(1) << level) - 1);
}

WorkItem sym_item(const uint8_t *cycle, uint64_t mult, int level, uint32_t origin) {
    WorkItem it;
    std::memset(&it, 0, sizeof(it));
    std::memcpy(it.state, cycle, 64);
    it.mask = 0;
    it.origin = origin;
    it.mult = static_cast<uint16_t>(mult);
    it.level = static_cast<uint8_t>(level);
    it.map = 0;
    return it;
}

WorkItem128 sym_item128(const uint8_t *cycle, uint64_t mult, int level, uint32_t origin) {
    WorkItem128 it;
    std::memset(&it, 0, sizeof(it));
    std::memcpy(it.state, cycle, sizeof(it.state));
    it.origin = origin;
    it.mult = static_cast<uint16_t>(mult);
    it.level = static_cast<uint8_t>(level);
    return it;
}

WorkItem lunnon_item(const uint8_t *b, int depth, int n, uint32_t origin, int map) {
    WorkItem it;
    std::memset(&it, 0, sizeof(it));
    std::memcpy(it.state, b, static_cast<size_t>(depth + 1));
    it.mask = 0;
    it.origin = origin;
    it.mult = static_cast<uint16_t>(n);
    it.level = static_cast<uint8_t>(depth + 1);
    it.map = static_cast<uint8_t>(map);
    return it;
}

}  // namespace

// Shared with the other translation units of the library (cost model,
// the no-allocation solve).
namespace mf {
int noalloc_setup(int device);
bool valid_case_cycle(const uint8_t *cycle, int level) { return valid_cycle(cycle, level); }
WorkItem make_sym_item(const uint8_t *cycle, uint64_t mult, int level, uint32_t origin) {
    return sym_item(cycle, mult, level, origin);
}
WorkItem128 make_sym_item128(const uint8_t *cycle, uint64_t mult, int level, uint32_t origin) {
    return sym_item128(cycle, mult, level, origin);
}
// The initialised context as it is (never initialises: mf_init allocates).
bool lib_current(LibContext *c) {
    if (g.device < 0) return false;
    c->device = g.device;
    c->cus = g.cus;
    c->stream = g.stream;
    c->ev0 = g.ev0;
    c->ev1 = g.ev1;
    c->host_acc = g.host_acc;
    c->host_counters = g.host_counters;
    return true;
}
int lib_error(int status, const char *msg) { return fail(status, "%s", msg); }
// Cost-model feature rows (80 floats each, device memory) of a resident sym
// frontier of up to 8x8, on the library stream.
int frontier_sym_features(mf_frontier *f, float *dev_out, uint64_t *rows) {
    if (!f || f->kind != 1) return fail(MF_ERR_ARG, "features on the device take a sym frontier of n <= 8");
    SymArgs a;
    uint32_t levels = 0;
    sym_args(*f, a, &levels);
    a.items = f->items.ptr;
    a.num_items = static_cast<uint32_t>(f->num);
    *rows = f->num;
    if (f->num == 0) return MF_OK;
    HIP_TRY(launch_sym_features(a, dev_out, g.stream));
    return MF_OK;
}
int lib_context(hipStream_t *stream, int *cus) {
    const int rc = require_init();
    if (rc != MF_OK) return rc;
    *stream = g.stream;
    *cus = g.cus;
    return MF_OK;
}
}  // namespace mf

// ====================================================================== C ABI

extern "C" {

const char *mf_status_string(int status) {
    switch (status) {
        case MF_OK: return "ok";
        case MF_ERR_ARG: return "invalid argument";
        case MF_ERR_DEPTH: return "invalid depth";
        case MF_ERR_HIP: return "HIP runtime error";
        case MF_ERR_NO_DEVICE: return "no GPU device";
        case MF_ERR_CAPACITY: return "capacity exceeded";
        case MF_ERR_INTERNAL: return "internal consistency failure";
        default: return "unknown status";
    }
}

const char *mf_last_error(void) { return g_err.c_str(); }

int mf_init(int device) {
    if (g.device == device) return MF_OK;
    if (g.device >= 0) mf_shutdown();
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) return fail(MF_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= count) return fail(MF_ERR_NO_DEVICE, "device %d out of range (%d visible)", device, count);
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipDeviceGetAttribute(&g.cus, hipDeviceAttributeMultiprocessorCount, device));
    HIP_TRY(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&g.ev0));
    HIP_TRY(hipEventCreate(&g.ev1));
    HIP_TRY(hipEventCreate(&g.la));
    HIP_TRY(hipEventCreate(&g.lb));
    HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&g.host_acc), ACC_SLOTS * sizeof(unsigned long long)));
    HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&g.host_counters), 4 * sizeof(uint32_t)));
    // The device address of the no-allocation solve's static scratch array.
    const int rc = noalloc_setup(device);
    if (rc != MF_OK) return rc;
    g.device = device;
    return MF_OK;
}

int mf_device_cus(void) { return g.device >= 0 ? g.cus : 0; }

void mf_shutdown(void) {
    if (g.device < 0) return;
    (void)hipSetDevice(g.device);
    if (g.stream) (void)hipStreamSynchronize(g.stream);
    if (g.graph) (void)hipGraphExecDestroy(g.graph);
    g.graph = nullptr;
    g.graph_key.clear();
    for (DevBuf *b : {&g.items, &g.spill_a, &g.spill_b, &g.acc, &g.counters, &g.per_count, &g.per_nodes, &g.per_map,
                      &g.tables, &g.deep, &g.targets})
        b->release();
    if (g.host_acc) (void)hipHostFree(g.host_acc);
    if (g.host_counters) (void)hipHostFree(g.host_counters);
    if (g.ev0) (void)hipEventDestroy(g.ev0);
    if (g.ev1) (void)hipEventDestroy(g.ev1);
    if (g.la) (void)hipEventDestroy(g.la);
    if (g.lb) (void)hipEventDestroy(g.lb);
    if (g.stream) (void)hipStreamDestroy(g.stream);
    g = Context();
}

// ----------------------------------------------------------------- Lunnon

int mf_lunnon_frontier(const int32_t *dims, int32_t ndim, int32_t depth, uint8_t *out_b, int64_t max_states,
                       int64_t *num_states) {
    LunnonProblem pr;
    std::string err;
    if (!dims || !lunnon_problem(dims, ndim, pr, err)) return fail(MF_ERR_ARG, "%s", err.c_str());
    if (depth < 1 || depth > pr.n) return fail(MF_ERR_DEPTH, "depth must be between 1 and %d", pr.n);
    const int64_t total = lunnon_frontier(pr, depth, out_b, out_b ? max_states : 0);
    if (num_states) *num_states = total;
    if (out_b && total > max_states) return fail(MF_ERR_CAPACITY, "frontier has %lld states", static_cast<long long>(total));
    return MF_OK;
}

int mf_lunnon_check_states(const int32_t *dims, int32_t ndim, int32_t depth, const uint8_t *states_b, int64_t num,
                           int64_t *first_bad) {
    if (first_bad) *first_bad = -1;
    if (num < 0 || (num > 0 && !states_b)) return fail(MF_ERR_ARG, "bad state arguments");
    LunnonProblem pr;
    std::string err;
    if (!dims || !lunnon_problem(dims, ndim, pr, err)) return fail(MF_ERR_ARG, "%s", err.c_str());
    if (depth < 1 || depth >= pr.n) return fail(MF_ERR_DEPTH, "depth must be between 1 and %d", pr.n - 1);
    for (int64_t s = 0; s < num; ++s)
        if (!lunnon_reachable(pr, states_b + s * (depth + 1), depth)) {
            if (first_bad) *first_bad = s;
            return fail(MF_ERR_ARG, "state %lld is not a folding-DFS state over leaves 0..%d",
                        static_cast<long long>(s), depth);
        }
    return MF_OK;
}

int mf_lunnon_frontier_create(const int32_t *dims, int32_t ndim, int32_t depth, const uint8_t *states_b, int64_t num,
                              mf_frontier **out) {
    int rc = require_init();
    if (rc != MF_OK) return rc;
    if (!out || num < 0 || (num > 0 && !states_b)) return fail(MF_ERR_ARG, "bad frontier arguments");
    LunnonProblem pr;
    std::string err;
    if (!dims || !lunnon_problem(dims, ndim, pr, err)) return fail(MF_ERR_ARG, "%s", err.c_str());
    if (depth < 1 || depth >= pr.n) return fail(MF_ERR_DEPTH, "depth must be between 1 and %d", pr.n - 1);
    if (num > 0xffffffffll) return fail(MF_ERR_CAPACITY, "too many states for one frontier");
    if ((rc = mf_lunnon_check_states(dims, ndim, depth, states_b, num, nullptr)) != MF_OK) return rc;
    auto *f = new mf_frontier();
    f->kind = 2;
    f->maps.push_back(pr);
    f->depths.push_back(depth);
    f->host_items.reserve(static_cast<size_t>(num));
    for (int64_t s = 0; s < num; ++s)
        f->host_items.push_back(lunnon_item(states_b + s * (depth + 1), depth, pr.n, static_cast<uint32_t>(s), 0));
    if ((rc = upload_items(f)) != MF_OK) {
        delete f;
        return rc;
    }
    *out = f;
    return MF_OK;
}

int mf_lunnon_solve(const int32_t *dims, int32_t ndim, int32_t depth, const uint8_t *states_b, int64_t num,
                    const mf_solve_options *opt, uint64_t *per_count, uint64_t *per_nodes, mf_result *out) {
    if (!out) return fail(MF_ERR_ARG, "out is NULL");
    mf_frontier *f = nullptr;
    int rc = mf_lunnon_frontier_create(dims, ndim, depth, states_b, num, &f);
    if (rc != MF_OK) return rc;
    rc = mf_frontier_solve(f, opt, per_count, per_nodes, out);
    mf_frontier_destroy(f);
    return rc;
}

int mf_gpu_folding(const int32_t *dims, int32_t ndim, int32_t depth, const mf_solve_options *opt, mf_result *out,
                   int64_t *num_states) {
    if (!out) return fail(MF_ERR_ARG, "out is NULL");
    LunnonProblem pr;
    std::string err;
    if (!dims || !lunnon_problem(dims, ndim, pr, err)) return fail(MF_ERR_ARG, "%s", err.c_str());
    if (depth < 1 || depth >= pr.n) return fail(MF_ERR_DEPTH, "partition depth must be between 1 and %d", pr.n - 1);
    const int64_t total = lunnon_frontier(pr, depth, nullptr, 0);
    if (num_states) *num_states = total;
    std::vector<uint8_t> states(static_cast<size_t>(total) * (depth + 1));
    lunnon_frontier(pr, depth, states.data(), total);
    return mf_lunnon_solve(dims, ndim, depth, states.data(), total, opt, nullptr, nullptr, out);
}

int mf_lunnon_batch_frontier_create(const int32_t *dims2, const int32_t *depths, int32_t nmaps, mf_frontier **out) {
    int rc = require_init();
    if (rc != MF_OK) return rc;
    if (!dims2 || !depths || !out || nmaps < 1 || nmaps > kMaxMaps) return fail(MF_ERR_ARG, "bad batch arguments");
    auto *f = new mf_frontier();
    f->kind = 2;
    for (int k = 0; k < nmaps; ++k) {
        LunnonProblem pr;
        std::string err;
        if (!lunnon_problem(dims2 + 2 * k, 2, pr, err)) {
            delete f;
            return fail(MF_ERR_ARG, "map %d: %s", k, err.c_str());
        }
        const int d = depths[k];
        if (d < 1 || d >= pr.n) {
            delete f;
            return fail(MF_ERR_DEPTH, "map %d: depth must be between 1 and %d", k, pr.n - 1);
        }
        f->maps.push_back(pr);
        f->depths.push_back(d);
    }
    // Frontiers of distinct (shape, depth) pairs, collected on host threads;
    // items are then laid out map by map, in map order.
    std::vector<int> first(nmaps);
    for (int k = 0; k < nmaps; ++k) {
        first[k] = k;
        for (int q = 0; q < k; ++q)
            if (f->depths[q] == f->depths[k] && dims2[2 * q] == dims2[2 * k] && dims2[2 * q + 1] == dims2[2 * k + 1]) {
                first[k] = q;
                break;
            }
    }
    std::vector<std::vector<uint8_t>> states(nmaps);
    std::atomic<int> next{0};
    auto work = [&]() {
        for (int k = next++; k < nmaps; k = next++) {
            if (first[k] != k) continue;
            const int d = f->depths[k];
            const int64_t cnt = lunnon_frontier(f->maps[k], d, nullptr, 0);
            states[k].resize(static_cast<size_t>(cnt) * (d + 1));
            lunnon_frontier(f->maps[k], d, states[k].data(), cnt);
        }
    };
    const int nthreads = std::max(1, std::min<int>(std::min<int>(nmaps, 16),
                                                   static_cast<int>(std::thread::hardware_concurrency())));
    std::vector<std::thread> pool;
    for (int t = 1; t < nthreads; ++t) pool.emplace_back(work);
    work();
    for (auto &th : pool) th.join();
    for (int k = 0; k < nmaps; ++k) {
        const int d = f->depths[k];
        const std::vector<uint8_t> &st = states[first[k]];
        const int64_t cnt = static_cast<int64_t>(st.size() / (d + 1));
        for (int64_t s2 = 0; s2 < cnt; ++s2)
            f->host_items.push_back(lunnon_item(st.data() + s2 * (d + 1), d, f->maps[k].n,
                                                static_cast<uint32_t>(f->host_items.size()), k));
    }
    if (f->host_items.size() > 0xffffffffull) {
        delete f;
        return fail(MF_ERR_CAPACITY, "too many states for one frontier");
    }
    if ((rc = upload_items(f)) != MF_OK) {
        delete f;
        return rc;
    }
    *out = f;
    return MF_OK;
}

int mf_frontier_solve_maps(mf_frontier *f, const mf_solve_options *opt, uint64_t *per_map_count, mf_result *out) {
    if (!f || !out || !per_map_count) return fail(MF_ERR_ARG, "frontier/out/per_map_count is NULL");
    if (f->kind != 2) return fail(MF_ERR_ARG, "per-map totals exist for Lunnon frontiers only");
    return solve_frontier(f, opt, nullptr, nullptr, per_map_count, out);
}

int mf_lunnon_batch(const int32_t *dims2, const int32_t *depths, int32_t nmaps, const mf_solve_options *opt,
                    uint64_t *per_map_count, mf_result *out) {
    if (!out || !per_map_count) return fail(MF_ERR_ARG, "out/per_map_count is NULL");
    mf_frontier *f = nullptr;
    int rc = mf_lunnon_batch_frontier_create(dims2, depths, nmaps, &f);
    if (rc != MF_OK) return rc;
    rc = mf_frontier_solve_maps(f, opt, per_map_count, out);
    mf_frontier_destroy(f);
    return rc;
}

// ------------------------------------------------------------------- sym

int mf_sym_set_numbering(int32_t center, int32_t order, int32_t lock_center_n) {
    SymNumbering num;
    num.center = center;
    num.order = order;
    num.lock_center_n = lock_center_n;
    std::string err;
    if (!set_sym_numbering(num, err)) return fail(MF_ERR_ARG, "%s", err.c_str());
    return MF_OK;
}

void mf_sym_get_numbering(int32_t *center, int32_t *order, int32_t *lock_center_n) {
    const SymNumbering num = sym_numbering();
    if (center) *center = num.center;
    if (order) *order = num.order;
    if (lock_center_n) *lock_center_n = num.lock_center_n;
}

int mf_sym_seeds(int32_t n, uint8_t *cycles, uint64_t *mults, int64_t max_seeds, int64_t *num_seeds, int64_t *raw) {
    SymProblem pr;
    std::string err;
    if (!sym_problem(n, pr, err)) return fail(MF_ERR_ARG, "%s", err.c_str());
    if (n < 6) return fail(MF_ERR_ARG, "seeds exist for 6 <= n <= 9");
    int64_t r = 0;
    const auto &seeds = sym_seed_list(pr, &r);
    if (raw) *raw = r;
    if (num_seeds) *num_seeds = static_cast<int64_t>(seeds.size());
    const int64_t k = std::min<int64_t>(max_seeds, static_cast<int64_t>(seeds.size()));
    const int cb = pr.case_bytes();
    for (int64_t s = 0; s < k; ++s) {
        if (cycles) std::memcpy(cycles + s * cb, seeds[s].cycle, cb);
        if (mults) mults[s] = seeds[s].mult;
    }
    if ((cycles || mults) && max_seeds < static_cast<int64_t>(seeds.size())) return fail(MF_ERR_CAPACITY, "too many seeds");
    return MF_OK;
}

static int check_expanded(const SymProblem &pr, int expanded) {
    const int c2 = pr.center * pr.center;
    if (expanded < c2 || expanded > pr.n2 - 1)
        return fail(MF_ERR_DEPTH, "MF_EXPANDED_N must be between %d and %d", c2, pr.n2 - 1);
    return MF_OK;
}

int mf_sym_cases_slice(int32_t n, int32_t expanded, uint64_t seed_start, uint64_t seed_count, uint64_t start,
                       uint64_t max_cases, uint8_t *cycles, uint64_t *mults, int64_t *num) {
    SymProblem pr;
    std::string err;
    if (!sym_problem(n, pr, err)) return fail(MF_ERR_ARG, "%s", err.c_str());
    if (n < 6) return fail(MF_ERR_ARG, "the case stream exists for 6 <= n <= 9");
    int rc = check_expanded(pr, expanded);
    if (rc != MF_OK) return rc;
    if (max_cases > 0 && (!cycles || !mults)) return fail(MF_ERR_ARG, "output buffers are NULL");
    const int64_t got = sym_cases(pr, expanded, start, max_cases, cycles, mults, seed_start, seed_count);
    if (num) *num = got;
    return MF_OK;
}

int mf_sym_cases(int32_t n, int32_t expanded, uint64_t start, uint64_t max_cases, uint8_t *cycles, uint64_t *mults,
                 int64_t *num) {
    return mf_sym_cases_slice(n, expanded, 0, MF_ALL_SEEDS, start, max_cases, cycles, mults, num);
}

int mf_sym_seed_case_counts(int32_t n, int32_t expanded, uint64_t seed_start, uint64_t seed_count, uint64_t *counts,
                            int64_t max_counts, int64_t *num) {
    SymProblem pr;
    std::string err;
    if (!sym_problem(n, pr, err)) return fail(MF_ERR_ARG, "%s", err.c_str());
    if (n < 6) return fail(MF_ERR_ARG, "the case stream exists for 6 <= n <= 9");
    int rc = check_expanded(pr, expanded);
    if (rc != MF_OK) return rc;
    const auto got = sym_slice_case_counts(pr, expanded, seed_start, seed_count);
    if (num) *num = static_cast<int64_t>(got.size());
    if (!counts) return MF_OK;
    if (max_counts < static_cast<int64_t>(got.size())) return fail(MF_ERR_CAPACITY, "counts has room for %lld of %zu seeds",
                                                                 static_cast<long long>(max_counts), got.size());
    std::memcpy(counts, got.data(), got.size() * sizeof(uint64_t));
    return MF_OK;
}

int mf_sym_case_total(int32_t n, int32_t expanded, uint64_t *total) {
    SymProblem pr;
    std::string err;
    if (!sym_problem(n, pr, err)) return fail(MF_ERR_ARG, "%s", err.c_str());
    if (n < 6) return fail(MF_ERR_ARG, "the case stream exists for 6 <= n <= 9");
    int rc = check_expanded(pr, expanded);
    if (rc != MF_OK) return rc;
    const auto &counts = sym_seed_case_counts(pr, expanded);
    uint64_t t = 0;
    for (uint64_t c : counts) t += c;
    if (total) *total = t;
    return MF_OK;
}

// num: the cell numbering (NULL = the process-wide one).
static int sym_frontier_create_stop(int32_t n, int32_t level, int32_t stop, const uint8_t *cycles,
                                    const uint64_t *mults, int64_t num, mf_frontier **out,
                                    const SymNumbering *numbering);

int mf_sym_frontier_create(int32_t n, int32_t level, const uint8_t *cycles, const uint64_t *mults, int64_t num,
                           mf_frontier **out) {
    return sym_frontier_create_stop(n, level, n * n, cycles, mults, num, out, nullptr);
}

static int sym_frontier_create_stop(int32_t n, int32_t level, int32_t stop, const uint8_t *cycles,
                                    const uint64_t *mults, int64_t num, mf_frontier **out,
                                    const SymNumbering *numbering) {
    int rc = require_init();
    if (rc != MF_OK) return rc;
    if (!out || num < 0 || (num > 0 && !cycles)) return fail(MF_ERR_ARG, "bad frontier arguments");
    if (num > 0xffffffffll) return fail(MF_ERR_CAPACITY, "too many cases for one frontier");
    SymProblem pr;
    std::string err;
    const bool ok = numbering ? sym_problem(n, *numbering, pr, err) : sym_problem(n, pr, err);
    if (!ok) return fail(MF_ERR_ARG, "%s", err.c_str());
    if (n < 2) return fail(MF_ERR_ARG, "n must be at least 2 for the cycle DFS");
    if (stop < 2 || stop > pr.n2) return fail(MF_ERR_DEPTH, "stop level must be between 2 and %d", pr.n2);
    if (level < 1 || level > stop - 1) return fail(MF_ERR_DEPTH, "level must be between 1 and %d", stop - 1);
    const int cb = pr.case_bytes();
    for (int64_t s = 0; s < num; ++s) {
        const uint64_t m = mults ? mults[s] : 1;
        if (m == 0 || m > 0xffff) return fail(MF_ERR_ARG, "multiplier out of range at case %lld", static_cast<long long>(s));
        if (!valid_cycle(cycles + s * cb, level))
            return fail(MF_ERR_ARG, "case %lld is not one cycle through cells 0..%d", static_cast<long long>(s), level - 1);
    }
    auto *f = new mf_frontier();
    f->kind = pr.n2 <= 64 ? 1 : 3;
    f->sym = pr;
    f->level = level;
    f->stop = stop;
    if (f->kind == 1) {
        f->host_items.reserve(static_cast<size_t>(num));
        for (int64_t s = 0; s < num; ++s)
            f->host_items.push_back(sym_item(cycles + s * cb, mults ? mults[s] : 1, level, static_cast<uint32_t>(s)));
    } else {
        f->host_items128.reserve(static_cast<size_t>(num));
        for (int64_t s = 0; s < num; ++s)
            f->host_items128.push_back(
                sym_item128(cycles + s * cb, mults ? mults[s] : 1, level, static_cast<uint32_t>(s)));
    }
    if ((rc = upload_items(f)) != MF_OK) {
        delete f;
        return rc;
    }
    *out = f;
    return MF_OK;
}

int mf_sym_solve(int32_t n, int32_t level, const uint8_t *cycles, const uint64_t *mults, int64_t num,
                 const mf_solve_options *opt, uint64_t *per_count, uint64_t *per_nodes, mf_result *out) {
    if (!out) return fail(MF_ERR_ARG, "out is NULL");
    mf_frontier *f = nullptr;
    int rc = mf_sym_frontier_create(n, level, cycles, mults, num, &f);
    if (rc != MF_OK) return rc;
    rc = mf_frontier_solve(f, opt, per_count, per_nodes, out);
    mf_frontier_destroy(f);
    return rc;
}

int mf_sym_expand(int32_t n, int32_t level, const uint8_t *cycles, int64_t count, uint8_t *out_cycles,
                  uint32_t *out_parent, int64_t max_out, int64_t *written) {
    if (!written || count < 0 || (count > 0 && !cycles) || max_out < 0 || (max_out > 0 && (!out_cycles || !out_parent)))
        return fail(MF_ERR_ARG, "bad expand arguments");
    if (count > 0xffffffffll) return fail(MF_ERR_CAPACITY, "too many cases");
    SymProblem pr;
    std::string err;
    if (!sym_problem(n, pr, err)) return fail(MF_ERR_ARG, "%s", err.c_str());
    if (n > 8) return fail(MF_ERR_ARG, "split frontiers take n <= 8");
    if (level < 1 || level > pr.n2 - 2) return fail(MF_ERR_DEPTH, "level must be between 1 and %d", pr.n2 - 2);
    int64_t w = 0;
    for (int64_t s = 0; s < count; ++s) {
        const uint8_t *src = cycles + s * 64;
        if (!valid_cycle(src, level))
            return fail(MF_ERR_ARG, "case %lld is not one cycle through cells 0..%d", static_cast<long long>(s), level - 1);
        const uint64_t good = sym_good(pr, src, level);
        for (uint64_t gm = good; gm; gm &= gm - 1, ++w) {
            if (w >= max_out) continue;
            const int j = __builtin_ctzll(gm);
            uint8_t *dst = out_cycles + w * 64;
            std::memcpy(dst, src, 64);
            dst[level] = dst[j];
            dst[j] = static_cast<uint8_t>(level);
            out_parent[w] = static_cast<uint32_t>(s);
        }
    }
    *written = w;
    if (w > max_out) return fail(MF_ERR_CAPACITY, "%lld children, room for %lld", static_cast<long long>(w),
                                 static_cast<long long>(max_out));
    return MF_OK;
}

int mf_sym_pieces_create(int32_t n, const uint8_t *cycles, const uint8_t *levels, const uint32_t *origins,
                         const uint64_t *mults, int64_t num, int64_t num_origins, mf_frontier **out) {
    int rc = require_init();
    if (rc != MF_OK) return rc;
    if (!out || num < 0 || num_origins < 0 || (num > 0 && (!cycles || !levels || !origins)))
        return fail(MF_ERR_ARG, "bad pieces arguments");
    if (num > 0xffffffffll || num_origins > 0xffffffffll) return fail(MF_ERR_CAPACITY, "too many pieces");
    SymProblem pr;
    std::string err;
    if (!sym_problem(n, pr, err)) return fail(MF_ERR_ARG, "%s", err.c_str());
    if (n < 2 || n > 8) return fail(MF_ERR_ARG, "split frontiers take 2 <= n <= 8");
    int lo = pr.n2;
    for (int64_t s = 0; s < num; ++s) {
        const int level = levels[s];
        const uint64_t m = mults ? mults[s] : 1;
        if (level < 1 || level > pr.n2 - 1)
            return fail(MF_ERR_DEPTH, "piece %lld: level %d outside [1, %d]", static_cast<long long>(s), level, pr.n2 - 1);
        if (origins[s] >= static_cast<uint64_t>(num_origins))
            return fail(MF_ERR_ARG, "piece %lld: origin out of range", static_cast<long long>(s));
        if (m == 0 || m > 0xffff) return fail(MF_ERR_ARG, "multiplier out of range at piece %lld", static_cast<long long>(s));
        if (!valid_cycle(cycles + s * 64, level))
            return fail(MF_ERR_ARG, "piece %lld is not one cycle through cells 0..%d", static_cast<long long>(s), level - 1);
        lo = std::min(lo, level);
    }
    auto *f = new mf_frontier();
    f->kind = 1;
    f->sym = pr;
    f->level = num > 0 ? lo : pr.n2 - 1;
    f->stop = pr.n2;
    f->host_items.reserve(static_cast<size_t>(num));
    for (int64_t s = 0; s < num; ++s)
        f->host_items.push_back(sym_item(cycles + s * 64, mults ? mults[s] : 1, levels[s], origins[s]));
    if ((rc = upload_items(f)) != MF_OK) {
        delete f;
        return rc;
    }
    f->origins = static_cast<uint64_t>(num_origins);
    *out = f;
    return MF_OK;
}

int mf_sym_fold(int32_t n, int32_t expanded, uint64_t start, uint64_t max_cases, const mf_solve_options *opt,
                uint64_t answer[2], mf_result *out) {
    return mf_sym_fold_slice(n, expanded, 0, MF_ALL_SEEDS, start, max_cases, opt, answer, out);
}

int mf_sym_fold_slice(int32_t n, int32_t expanded, uint64_t seed_start, uint64_t seed_count, uint64_t start,
                      uint64_t max_cases, const mf_solve_options *opt, uint64_t answer[2], mf_result *out) {
    if (!answer || !out) return fail(MF_ERR_ARG, "answer/out is NULL");
    std::memset(out, 0, sizeof(*out));
    if (n < 1 || n > 9) return fail(MF_ERR_ARG, "n must be between 1 and 9");
    unsigned __int128 ans = 0;
    if (n == 1) {
        ans = 1;
    } else if (n == 2) {
        ans = 8;
    } else if (n <= 5) {
        // One case: the initial cycle 0 -> 1 -> 2 -> 0 at cell 3; both
        // orientations and n^2 starting cells (sym_fold.cpp:691-702).
        uint8_t cycle[64] = {1, 2, 0};
        int rc = mf_sym_solve(n, 3, cycle, nullptr, 1, opt, nullptr, nullptr, out);
        if (rc != MF_OK) return rc;
        ans = static_cast<unsigned __int128>(2) * static_cast<unsigned>(n * n) * out->count_lo;
    } else {
        SymProblem pr;
        std::string err;
        if (!sym_problem(n, pr, err)) return fail(MF_ERR_ARG, "%s", err.c_str());
        const int level = (n >= 7) ? expanded : pr.center * pr.center;
        if (n >= 7) {
            int rc = check_expanded(pr, expanded);
            if (rc != MF_OK) return rc;
        }
        // Stream the window in batches that bound host and device memory; a
        // host thread generates batch k+1 while the GPU counts batch k.
        const uint64_t batch = 1ull << 22;
        uint64_t pos = start, left = max_cases == 0 ? ~0ull : max_cases;
        struct Batch {
            std::vector<uint8_t> cyc;
            std::vector<uint64_t> mul;
            uint64_t want = 0;
            int64_t got = 0;
        } cur, nxt;
        auto generate = [&pr, level, seed_start, seed_count](Batch &b, uint64_t at, uint64_t want) {
            b.want = want;
            b.cyc.resize(want * pr.case_bytes());
            b.mul.resize(want);
            b.got = sym_cases(pr, level, at, want, b.cyc.data(), b.mul.data(), seed_start, seed_count);
        };
        unsigned __int128 sum = 0;
        mf_result part;
        generate(cur, pos, std::min(batch, left));
        while (left > 0) {
            const int64_t got = cur.got;
            if (got <= 0) break;
            const bool more = static_cast<uint64_t>(got) == cur.want && left > static_cast<uint64_t>(got);
            std::thread producer;
            if (more)
                producer = std::thread(generate, std::ref(nxt), pos + static_cast<uint64_t>(got),
                                       std::min(batch, left - static_cast<uint64_t>(got)));
            int rc = mf_sym_solve(n, level, cur.cyc.data(), cur.mul.data(), got, opt, nullptr, nullptr, &part);
            if (producer.joinable()) producer.join();
            if (rc != MF_OK) return rc;
            sum += (static_cast<unsigned __int128>(part.count_hi) << 64) | part.count_lo;
            out->nodes += part.nodes;
            out->walks += part.walks;
            out->steps += part.steps;
            out->leaves += part.leaves;
            out->items += part.items;
            out->rounds += part.rounds;
            out->inputs += part.inputs;
            out->kernel_ms += part.kernel_ms;
            out->span_ms += part.span_ms;
            pos += static_cast<uint64_t>(got);
            left -= static_cast<uint64_t>(got);
            if (!more) break;
            std::swap(cur, nxt);
        }
        ans = sum * static_cast<unsigned>(n * n);
    }
    answer[0] = static_cast<uint64_t>(ans);
    answer[1] = static_cast<uint64_t>(ans >> 64);
    return MF_OK;
}

int mf_frontier_solve(mf_frontier *f, const mf_solve_options *opt, uint64_t *per_count, uint64_t *per_nodes,
                      mf_result *out) {
    if (!f || !out) return fail(MF_ERR_ARG, "frontier/out is NULL");
    return solve_frontier(f, opt, per_count, per_nodes, nullptr, out);
}

uint64_t mf_frontier_items(const mf_frontier *f) { return f ? f->num : 0; }

void mf_frontier_destroy(mf_frontier *f) {
    if (!f) return;
    f->items.release();
    delete f;
}

// ------------------------------------------- sym_kernel.c-compatible batch ABI
// Record layout and status words of mfah/wasm/sym_kernel.c:1-34,240-277:
// 112-byte records (cycle bytes, u64 LE multiplier at byte 100), output =
// u128 count LE, u64 nodes, u32 status.  Node definition of that file: one
// per recurse_count call, the calls that reach stop_i included.

static uint8_t g_wasm_cases[256 * 112];
static uint32_t g_wasm_out[8];

uint8_t *mf_wasm_input_ptr(void) { return g_wasm_cases; }
uint32_t *mf_wasm_output_ptr(void) { return g_wasm_out; }
uint32_t mf_wasm_case_stride(void) { return 112; }
uint32_t mf_wasm_max_cases(void) { return 256; }

static uint32_t wasm_write(unsigned __int128 count, uint64_t nodes, uint32_t status) {
    const uint64_t lo = static_cast<uint64_t>(count), hi = static_cast<uint64_t>(count >> 64);
    g_wasm_out[0] = static_cast<uint32_t>(lo);
    g_wasm_out[1] = static_cast<uint32_t>(lo >> 32);
    g_wasm_out[2] = static_cast<uint32_t>(hi);
    g_wasm_out[3] = static_cast<uint32_t>(hi >> 32);
    g_wasm_out[4] = static_cast<uint32_t>(nodes);
    g_wasm_out[5] = static_cast<uint32_t>(nodes >> 32);
    g_wasm_out[6] = status;
    g_wasm_out[7] = 0;
    return status;
}

uint32_t mf_wasm_solve_cases(uint32_t n, uint32_t start_i, uint32_t stop_i, uint32_t case_count) {
    // Squares up to 9x9 (MAX_SIDE): u64-mask kernels to 8x8, u128 for 9x9.
    if (n < 2 || n > 9 || case_count > 256) {
        fail(MF_ERR_ARG, "n must be 2..9 and case_count <= 256");
        return wasm_write(0, 0, 1);
    }
    const uint32_t n2 = n * n;
    if (start_i > stop_i || stop_i > n2) {
        fail(MF_ERR_DEPTH, "need start_i <= stop_i <= n*n");
        return wasm_write(0, 0, 2);
    }
    const size_t cb = n2 <= 64 ? 64 : 128;   // case row bytes
    std::vector<uint8_t> cycles(static_cast<size_t>(case_count) * cb, 0);
    std::vector<uint64_t> mults(case_count);
    unsigned __int128 total = 0;
    for (uint32_t k = 0; k < case_count; ++k) {
        const uint8_t *rec = g_wasm_cases + k * 112;
        std::memcpy(cycles.data() + k * cb, rec, n2);
        uint64_t m = 0;
        for (int b = 7; b >= 0; --b) m = (m << 8) | rec[100 + b];
        mults[k] = m;
    }
    if (start_i == stop_i || case_count == 0) {
        for (uint32_t k = 0; k < case_count; ++k) total += mults[k];
        return wasm_write(total, case_count, 0);
    }
    // Level 0 below the stop level: get_good(cycle, 0) is empty (no cell lies
    // below cell 0), so every case is one childless node (sym_kernel.c:183,
    // 212-230).
    if (start_i == 0) return wasm_write(0, case_count, 0);
    // Multipliers can exceed the kernel's 16-bit field: count with 1 and
    // weight on the host from the per-case results.  Records are in spiral
    // order (sym_kernel.c make_edges), whatever the process-wide numbering.
    mf_frontier *f = nullptr;
    const SymNumbering spiral_numbering;
    if (sym_frontier_create_stop(static_cast<int32_t>(n), static_cast<int32_t>(start_i), static_cast<int32_t>(stop_i),
                                 cycles.data(), nullptr, case_count, &f, &spiral_numbering) != MF_OK)
        return wasm_write(0, 0, 3);
    std::vector<uint64_t> per(case_count, 0);
    mf_result r;
    mf_solve_options opt;
    std::memset(&opt, 0, sizeof(opt));
    opt.schedule = MF_SCHED_RESPLIT;
    opt.count_events = 0;
    const int rc = solve_frontier(f, &opt, per.data(), nullptr, nullptr, &r);
    mf_frontier_destroy(f);
    if (rc != MF_OK) return wasm_write(0, 0, 3);
    for (uint32_t k = 0; k < case_count; ++k) total += static_cast<unsigned __int128>(per[k]) * mults[k];
    return wasm_write(total, r.nodes + r.leaves, 0);
}

}  // extern "C"
