// noalloc.hip -- mf_sym_solve_noalloc: the symmetry-DFS solve of
// mf_sym_solve without any device allocation.
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid it).  Everything the
// re-split schedule needs beyond the caller's inputs and outputs -- the
// accumulators, the claim / spill counters, the per-case logs, the resident
// input items and both spill queues -- lives in one file-scope __device__
// array, `scratch`, of kScratchBytes.  Each GPU the library runs on holds its
// own copy until the process ends; mf_init looks up its device address once
// per device (noalloc_setup) and every launch uses that address, never the
// array's host-side name.
//
// The kernels are the persistent re-split kernels of sym_kernels.hip.  The
// two spill queues share the rest of the array (about 360,000 items of 80
// bytes each), and the grid is cut so that the worst-case spill of a round
// fits them (8x8 depth 36: 50 workgroups of 256 lanes, a twenty-fifth of the
// full grid; the allocating mf_sym_solve runs the full grid).
// Rounds are launched from the host (no graph: instantiating one
// allocates).  Every call clears what must start at zero; the array keeps the
// last call's contents otherwise, and calls that overlap on one GPU (from
// several host threads) race on it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/mapfold.h"
#include "mf_counters.hpp"
#include "mf_device.hpp"
#include "mf_host.hpp"

namespace mf {
hipError_t launch_sym(const SymArgs &, int core, bool persistent, uint32_t grid, uint32_t block, uint32_t lds,
                      bool events, uint32_t cells, hipStream_t);
uint32_t sym_lds_bytes(int core, uint32_t levels, uint32_t block, uint32_t cells);
const void *sym_kernel_ptr(int core, bool persistent, bool events, uint32_t cells);
int sym_walk_core(int core, const SymArgs &a, uint32_t cells);
int lib_error(int status, const char *msg);
bool valid_case_cycle(const uint8_t *cycle, int level);
WorkItem make_sym_item(const uint8_t *cycle, uint64_t mult, int level, uint32_t origin);
WorkItem128 make_sym_item128(const uint8_t *cycle, uint64_t mult, int level, uint32_t origin);

namespace {

constexpr size_t kScratchBytes = size_t{64} << 20;   // 64 MiB per GPU
constexpr uint64_t kMaxCases = MF_NOALLOC_MAX_CASES;
constexpr int kMaxDevices = 64;

// Layout of the array (byte offsets, 256-aligned).
constexpr size_t kOffAcc = 0;                                  // ACC_SLOTS u64
constexpr size_t kOffCounters = 256;                           // claim, spill count A, spill count B
constexpr size_t kOffPerCount = 512;                           // kMaxCases u64
constexpr size_t kOffPerNodes = kOffPerCount + kMaxCases * 8;  // kMaxCases u64
constexpr size_t kOffItems = kOffPerNodes + kMaxCases * 8;     // kMaxCases items of up to 128 bytes
constexpr size_t kOffSpill = kOffItems + kMaxCases * sizeof(WorkItem128);
static_assert(kOffSpill % 256 == 0 && kOffSpill < kScratchBytes / 4, "scratch layout");

__device__ __align__(256) uint8_t scratch[kScratchBytes];

// Device address of `scratch` on each device (set by noalloc_setup).
uint8_t *g_scratch_dev[kMaxDevices] = {};

// Host staging of the input items (static: nothing is allocated per call).
alignas(64) uint8_t g_stage[kMaxCases * sizeof(WorkItem128)];

int hip_fail(const char *what, hipError_t e) {
    std::string m = std::string(what) + " failed: " + hipGetErrorString(e);
    return lib_error(MF_ERR_HIP, m.c_str());
}

#define NA_TRY(call)                                   \
    do {                                               \
        const hipError_t e_ = (call);                  \
        if (e_ != hipSuccess) return hip_fail(#call, e_); \
    } while (0)

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

// With `device` current: look up this device's copy of the scratch array.
int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return lib_error(MF_ERR_NO_DEVICE, "device index beyond 63");
    if (g_scratch_dev[device] != nullptr) return MF_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail("hipGetSymbolAddress(scratch)", e);
    if (p == nullptr || (reinterpret_cast<uintptr_t>(p) & 255u) != 0)
        return lib_error(MF_ERR_INTERNAL, "scratch array has no 256-byte aligned device address");
    g_scratch_dev[device] = static_cast<uint8_t *>(p);
    return MF_OK;
}

}  // namespace mf

using namespace mf;

extern "C" int mf_sym_solve_noalloc(int32_t n, int32_t level, const uint8_t *cycles, const uint64_t *mults,
                                    int64_t num, const mf_solve_options *opt, uint64_t *per_count,
                                    uint64_t *per_nodes, mf_result *out) {
    const auto t0 = std::chrono::steady_clock::now();
    if (!out) return lib_error(MF_ERR_ARG, "out is NULL");
    std::memset(out, 0, sizeof(*out));
    LibContext ctx;
    if (!lib_current(&ctx)) return lib_error(MF_ERR_NO_DEVICE, "mf_sym_solve_noalloc needs mf_init first");
    uint8_t *base = g_scratch_dev[ctx.device];
    if (base == nullptr) return lib_error(MF_ERR_INTERNAL, "scratch address of this device unknown");
    if (num < 0 || (num > 0 && !cycles)) return lib_error(MF_ERR_ARG, "bad case arguments");
    if (static_cast<uint64_t>(num) > kMaxCases) {
        char msg[160];
        std::snprintf(msg, sizeof(msg), "mf_sym_solve_noalloc takes at most %llu cases (got %lld)",
                      static_cast<unsigned long long>(kMaxCases), static_cast<long long>(num));
        return lib_error(MF_ERR_CAPACITY, msg);
    }
    if (opt && opt->order) return lib_error(MF_ERR_ARG, "mf_sym_solve_noalloc takes no item order");
    SymProblem pr;
    std::string err;
    if (!sym_problem(n, pr, err)) return lib_error(MF_ERR_ARG, err.c_str());
    if (n < 2) return lib_error(MF_ERR_ARG, "n must be at least 2 for the cycle DFS");
    if (level < 1 || level > pr.n2 - 1) return lib_error(MF_ERR_DEPTH, "level outside [1, n*n - 1]");
    const int cb = pr.case_bytes();
    const bool c96 = pr.n2 > 64;
    const size_t item_bytes = c96 ? sizeof(WorkItem128) : sizeof(WorkItem);
    for (int64_t s = 0; s < num; ++s) {
        const uint64_t m = mults ? mults[s] : 1;
        if (m == 0 || m > 0xffff) return lib_error(MF_ERR_ARG, "multiplier out of range");
        if (!valid_case_cycle(cycles + s * cb, level)) return lib_error(MF_ERR_ARG, "a case is not one cycle through cells 0..level-1");
    }
    for (int64_t s = 0; s < num; ++s) {
        const uint64_t m = mults ? mults[s] : 1;
        if (c96) {
            const WorkItem128 it = make_sym_item128(cycles + s * cb, m, level, static_cast<uint32_t>(s));
            std::memcpy(g_stage + s * item_bytes, &it, item_bytes);
        } else {
            const WorkItem it = make_sym_item(cycles + s * cb, m, level, static_cast<uint32_t>(s));
            std::memcpy(g_stage + s * item_bytes, &it, item_bytes);
        }
    }

    // Launch configuration: the mf_sym_solve defaults.
    int core = c96 ? 2 : ((opt && opt->core == 1) ? 1 : 2);
    uint32_t block = (opt && opt->block > 0) ? static_cast<uint32_t>(opt->block) : 256u;
    block = std::max(64u, std::min(256u, (block / 64u) * 64u));
    const uint32_t waves_per_cu = (opt && opt->waves_per_cu > 0) ? static_cast<uint32_t>(opt->waves_per_cu) : 32u;
    uint32_t budget = (opt && opt->budget > 0) ? static_cast<uint32_t>(opt->budget) : 64u;
    budget = std::min(budget, kMaxFoldBudget);
    const uint32_t tail = (opt && opt->tail > 0) ? static_cast<uint32_t>(opt->tail) : budget;
    const bool events = opt ? opt->count_events != 0 : true;

    SymArgs a;
    std::memset(&a, 0, sizeof(a));
    std::memcpy(a.edge_word, pr.edge_word, sizeof(a.edge_word));
    a.last = static_cast<uint32_t>(pr.n2 - 1);
    a.base_level = static_cast<uint32_t>(level);
    a.levels = std::max<uint32_t>(1u, a.last - a.base_level);
    a.budget = budget;
    a.tail = tail;
    const uint32_t cells = static_cast<uint32_t>(pr.n2);
    core = sym_walk_core(core, a, cells);
    const uint32_t lds = sym_lds_bytes(core, a.levels, block, cells);
    const void *kp = sym_kernel_ptr(core, true, events, cells);
    NA_TRY(hipFuncSetAttribute(kp, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    int per_cu = 0;
    NA_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kp, static_cast<int>(block), lds));
    if (per_cu <= 0) return lib_error(MF_ERR_ARG, "kernel does not fit a CU");
    uint32_t grid = std::min<uint32_t>(static_cast<uint32_t>(per_cu), std::max(1u, waves_per_cu / (block / 64u))) *
                    static_cast<uint32_t>(ctx.cus);
    // Two spill queues in the rest of the array.  A lane spills at most
    // (levels + 1) items per round (its pending node and one sibling set per
    // stack level), so the grid is cut to what the queues hold in the worst
    // case: no wave's spill reservation ever fails (one that failed would run
    // its whole remaining subtree alone, which is correct but can take
    // minutes on the largest cases).
    const size_t queue_bytes = ((kScratchBytes - kOffSpill) / 2) & ~size_t{255};
    const uint64_t cap = queue_bytes / item_bytes;
    const uint64_t fit_blocks = cap / (static_cast<uint64_t>(a.levels + 1) * block);
    if (fit_blocks == 0) return lib_error(MF_ERR_CAPACITY, "scratch too small for one workgroup at this level");
    grid = static_cast<uint32_t>(std::min<uint64_t>(grid, fit_blocks));

    unsigned long long *acc = reinterpret_cast<unsigned long long *>(base + kOffAcc);
    uint32_t *claim = reinterpret_cast<uint32_t *>(base + kOffCounters);
    uint32_t *counts = claim + 1;
    unsigned long long *dcount = reinterpret_cast<unsigned long long *>(base + kOffPerCount);
    unsigned long long *dnodes = reinterpret_cast<unsigned long long *>(base + kOffPerNodes);
    void *items = base + kOffItems;
    void *bufs[2] = {base + kOffSpill, base + kOffSpill + queue_bytes};
    hipStream_t st = ctx.stream;

    NA_TRY(hipMemsetAsync(acc, 0, ACC_SLOTS * sizeof(unsigned long long), st));
    if (per_count && num > 0) NA_TRY(hipMemsetAsync(dcount, 0, static_cast<size_t>(num) * 8, st));
    if (per_nodes && num > 0) NA_TRY(hipMemsetAsync(dnodes, 0, static_cast<size_t>(num) * 8, st));
    if (num > 0) NA_TRY(hipMemcpyAsync(items, g_stage, static_cast<size_t>(num) * item_bytes, hipMemcpyHostToDevice, st));
    a.acc = acc;
    a.per_count = per_count ? dcount : nullptr;
    a.per_nodes = per_nodes ? dnodes : nullptr;
    a.claim = claim;
    a.spill_cap = static_cast<uint32_t>(std::min<uint64_t>(cap, 0xffffffffull));

    const void *in = items;
    uint64_t in_num = static_cast<uint64_t>(num);
    int next = 0;
    uint32_t rounds = 0;
    NA_TRY(hipEventRecord(ctx.ev0, st));
    while (in_num > 0) {
        if (rounds >= 100000) return lib_error(MF_ERR_INTERNAL, "re-split did not converge");
        NA_TRY(hipMemsetAsync(claim, 0, sizeof(uint32_t), st));
        NA_TRY(hipMemsetAsync(counts + next, 0, sizeof(uint32_t), st));
        a.items = in;
        a.num_items = static_cast<uint32_t>(in_num);
        a.num_items_dev = nullptr;
        a.spill = bufs[next];
        a.spill_count = counts + next;
        NA_TRY(launch_sym(a, core, true, grid, block, lds, events, cells, st));
        NA_TRY(hipMemcpyAsync(ctx.host_counters, counts + next, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        NA_TRY(hipMemcpyAsync(ctx.host_acc + ACC_ERROR, acc + ACC_ERROR, 8, hipMemcpyDeviceToHost, st));
        NA_TRY(hipStreamSynchronize(st));
        ++rounds;
        if (ctx.host_acc[ACC_ERROR] & 6ull) return lib_error(MF_ERR_INTERNAL, "device consistency check failed");
        if (ctx.host_acc[ACC_ERROR] != 0) return lib_error(MF_ERR_CAPACITY, "spill queue overflow");
        in = bufs[next];
        in_num = ctx.host_counters[0];
        next ^= 1;
    }
    NA_TRY(hipEventRecord(ctx.ev1, st));
    NA_TRY(hipMemcpyAsync(ctx.host_acc, acc, ACC_SLOTS * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    if (per_count && num > 0)
        NA_TRY(hipMemcpyAsync(per_count, dcount, static_cast<size_t>(num) * 8, hipMemcpyDeviceToHost, st));
    if (per_nodes && num > 0)
        NA_TRY(hipMemcpyAsync(per_nodes, dnodes, static_cast<size_t>(num) * 8, hipMemcpyDeviceToHost, st));
    NA_TRY(hipStreamSynchronize(st));
    if (ctx.host_acc[ACC_ERROR] != 0) return lib_error(MF_ERR_INTERNAL, "device error flag set");
    const unsigned long long *h = ctx.host_acc;
    out->count_lo = h[ACC_COUNT_LO];
    out->count_hi = h[ACC_COUNT_HI];
    out->nodes = h[ACC_NODES];
    out->walks = h[ACC_WALKS];
    out->steps = h[ACC_STEPS];
    out->leaves = h[ACC_LEAVES];
    out->items = h[ACC_ITEMS];
    out->rounds = rounds;
    out->inputs = static_cast<uint64_t>(num);
    float ms = 0.0f;
    NA_TRY(hipEventElapsedTime(&ms, ctx.ev0, ctx.ev1));
    out->kernel_ms = ms;
    out->span_ms = ms;
    out->total_ms = ms_since(t0);
    return MF_OK;
}
