// cost.cpp -- C ABI of the learned subtree-cost model (include/mapfold_cost.h).
//
// Host half: feature extraction from sym cases / Lunnon states (CPU, the same
// tables the generators use), parameter initialisation, the training loop
// (batches gathered on the device from a resident copy of the data set, one
// Adam step per batch) and prediction.  Every product and reduction runs in
// the kernels of cost_mlp.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mapfold.h"
#include "../../include/mapfold_cost.h"
#include "mf_cost.hpp"
#include "mf_host.hpp"

namespace mf {
int lib_error(int status, const char *msg);
int lib_context(hipStream_t *stream, int *cus);
int frontier_sym_features(mf_frontier *f, float *dev_out, uint64_t *rows);
}  // namespace mf

using namespace mf;

static_assert(MF_COST_FEATURES == kFeatures && MF_COST_HIDDEN == kHidden && MF_COST_PARAMS == kParams,
              "C header and kernel shapes disagree");

namespace {

#define COST_TRY(call)                                                               \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess) {                                                      \
            std::string m_ = std::string(#call) + " failed: " + hipGetErrorString(e_); \
            return lib_error(MF_ERR_HIP, m_.c_str());                                \
        }                                                                            \
    } while (0)

struct Buf {
    void *p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t want) {
        if (want <= bytes) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&p, want);
        if (e == hipSuccess) bytes = want;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    float *f() const { return static_cast<float *>(p); }
};

uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

float uniform01(uint64_t &s) { return static_cast<float>(splitmix(s) >> 40) * (1.0f / 16777216.0f); }

// Activations of one batch (rows x 128 per layer) and their gradients.
struct Work {
    Buf x, y, h1, h2, h3, pred, g, d1, d2, d3, partial, idx, loss, part;
    int rows = 0;
    hipError_t ensure(int r, int nbatch_loss) {
        hipError_t e;
        const size_t R = static_cast<size_t>(std::max(r, 1));
        if ((e = x.ensure(R * kFeatures * 4)) != hipSuccess) return e;
        for (Buf *b : {&h1, &h2, &h3, &d1, &d2, &d3})
            if ((e = b->ensure(R * kHidden * 4)) != hipSuccess) return e;
        for (Buf *b : {&y, &pred, &g})
            if ((e = b->ensure(R * 4)) != hipSuccess) return e;
        if ((e = partial.ensure((R / 256 + 1) * 4)) != hipSuccess) return e;
        if ((e = loss.ensure(static_cast<size_t>(std::max(nbatch_loss, 1)) * 4)) != hipSuccess) return e;
        if ((e = part.ensure(static_cast<size_t>(kSplitK) * kHidden * kHidden * 4)) != hipSuccess) return e;
        return hipSuccess;
    }
    void release() {
        for (Buf *b : {&x, &y, &h1, &h2, &h3, &pred, &g, &d1, &d2, &d3, &partial, &idx, &loss, &part}) b->release();
    }
};

}  // namespace

struct mf_cost_model {
    Buf params, grad, m, v, frag, small;
    Work w;
    Buf data_x, data_y, pred;   // resident training set / prediction scratch
    int step = 0;
    bool frag_valid = false;
};

namespace {

int context(hipStream_t *s, int *cus) { return lib_context(s, cus); }

// FP32 forward of `rows` rows already in w.x; leaves h1..h3 and pred.
hipError_t forward(mf_cost_model *m, int rows, hipStream_t s) {
    const float *P = m->params.f();
    hipError_t e;
    if ((e = cost_linear_fwd(rows, kFeatures, m->w.x.f(), P + kOffW1, P + kOffB1, m->w.h1.f(), s)) != hipSuccess)
        return e;
    if ((e = cost_linear_fwd(rows, kHidden, m->w.h1.f(), P + kOffW2, P + kOffB2, m->w.h2.f(), s)) != hipSuccess)
        return e;
    if ((e = cost_linear_fwd(rows, kHidden, m->w.h2.f(), P + kOffW3, P + kOffB3, m->w.h3.f(), s)) != hipSuccess)
        return e;
    return cost_head_fwd(rows, m->w.h3.f(), P + kOffW4, m->w.pred.f(), s);
}

// MSE against w.y and the full gradient into m->grad; batch loss to *loss_dev.
hipError_t backward(mf_cost_model *m, int rows, float *loss_dev, hipStream_t s) {
    const float *P = m->params.f();
    float *G = m->grad.f();
    Work &w = m->w;
    hipError_t e;
    if ((e = cost_mse(rows, w.pred.f(), w.y.f(), w.g.f(), w.partial.f(), loss_dev, s)) != hipSuccess) return e;
    if ((e = cost_head_bwd(rows, w.g.f(), P + kOffW4, w.h3.f(), G + kOffW4, w.d3.f(), s)) != hipSuccess) return e;
    if ((e = cost_linear_bwd_weight(rows, kHidden, w.d3.f(), w.h2.f(), G + kOffW3, G + kOffB3, w.part.f(), s)) != hipSuccess)
        return e;
    if ((e = cost_linear_bwd_input(rows, kHidden, w.d3.f(), P + kOffW3, w.h2.f(), w.d2.f(), s)) != hipSuccess)
        return e;
    if ((e = cost_linear_bwd_weight(rows, kHidden, w.d2.f(), w.h1.f(), G + kOffW2, G + kOffB2, w.part.f(), s)) != hipSuccess)
        return e;
    if ((e = cost_linear_bwd_input(rows, kHidden, w.d2.f(), P + kOffW2, w.h1.f(), w.d1.f(), s)) != hipSuccess)
        return e;
    return cost_linear_bwd_weight(rows, kFeatures, w.d1.f(), w.x.f(), G + kOffW1, G + kOffB1, w.part.f(), s);
}

// ------------------------------------------------------------------ features
//
// Row layout (kFeatures = 80):
//   [0, 64)  state encoding, one slot per cell / leaf:
//              placed (c < level): rank of c along the stacking cycle from
//                                  cell 0, (rank + 1) / 64;
//              sym, not yet placed: how many insertion points the cell would
//                                  have against the current stack, i.e. the
//                                  intersection of the walks from its placed
//                                  page partners, / level (an upper bound on
//                                  its future options; 1 = unconstrained).
//   64 level / 64        65 level / leaves      66 gap count (children) / 16
//   67 leaves / 64       68 kind (1 sym, 0 Lunnon)
//   69 constrained pages of the next cell / 4 (sym), dimensions / 4 (Lunnon)
//   70 levels left / 64  71 n / 8 (sym), first dimension / 16 (Lunnon)
//   [72, 80) lookahead profile: log2(1 + descendants at level + k) / 24,
//            k = 1..8 (estimate_work of sym_fold.cpp:235-250 at depth k;
//            the gap count is k = 1).

inline float lg(double v) { return static_cast<float>(std::log2(1.0 + v)); }

constexpr int kProfile = 8;

// Walk of get_good for cell i on page p from a placed anchor, against a stack
// holding the cells below `placed` (sym_fold.cpp:179-204).
uint64_t page_reach(const SymProblem &pr, const uint8_t *cyc, int p, int anchor, int placed) {
    uint64_t reach = 1ull << anchor;
    for (int j = cyc[anchor]; j != anchor; j = cyc[j]) {
        const int partner = pr.edge[p][j];
        if (partner < placed) j = partner;
        reach |= 1ull << j;
    }
    return reach;
}

// Descendants of the node at level i that reach level i + k, for k = 1..kmax
// (added into prof[k-1]); one DFS to depth kmax.
void sym_profile(const SymProblem &pr, uint8_t *cyc, int i, int k, int kmax, double *prof) {
    if (i >= pr.n2) return;
    const uint64_t good = sym_good(pr, cyc, i);
    prof[k] += static_cast<double>(__builtin_popcountll(good));
    if (k + 1 >= kmax || i + 1 >= pr.n2) return;
    for (uint64_t gm = good; gm; gm &= gm - 1) {
        const int j = __builtin_ctzll(gm);
        cyc[i] = cyc[j];
        cyc[j] = static_cast<uint8_t>(i);
        sym_profile(pr, cyc, i + 1, k + 1, kmax, prof);
        cyc[j] = cyc[i];
    }
}

void sym_row(const SymProblem &pr, int level, const uint8_t *in, float *f) {
    std::memset(f, 0, sizeof(float) * kFeatures);
    uint8_t cyc[kCells];
    std::memcpy(cyc, in, kCells);
    const int n2 = pr.n2;
    int c = 0;
    for (int r = 0; r < level; ++r) {
        f[c] = static_cast<float>(r + 1) / 64.0f;
        c = cyc[c];
        if (c == 0) break;
    }
    for (int cell = level; cell < n2; ++cell) {
        uint64_t opts = level >= 64 ? ~0ull : ((1ull << level) - 1ull);
        for (int p = 0; p < 4 && opts; ++p) {
            const int anchor = pr.edge[p][cell];
            if (anchor < level) opts &= page_reach(pr, cyc, p, anchor, level);
        }
        f[cell] = static_cast<float>(__builtin_popcountll(opts)) / static_cast<float>(level);
    }
    double prof[kProfile] = {0};
    sym_profile(pr, cyc, level, 0, kProfile, prof);
    int pages = 0;
    for (int p = 0; p < 4; ++p) pages += (level < n2 && pr.edge[p][level] < level) ? 1 : 0;
    f[64] = level / 64.0f;
    f[65] = static_cast<float>(level) / n2;
    f[66] = static_cast<float>(prof[0]) / 16.0f;
    f[67] = n2 / 64.0f;
    f[68] = 1.0f;
    f[69] = pages / 4.0f;
    f[70] = (n2 - level) / 64.0f;
    f[71] = pr.n / 8.0f;
    for (int k = 0; k < kProfile; ++k) f[72 + k] = lg(prof[k]) / 24.0f;
}

// Legal positions for leaf l given the circular list b (push_gaps of the
// frontier walker, mf.c:88-125), with the normal-form filter b[0] == 1.
int lunnon_gaps(const LunnonProblem &pr, const uint8_t *b, int l, uint8_t *out) {
    if (l > 1 && b[0] != 1) return 0;
    uint8_t tally[kMaxLeaves + 2] = {0};
    uint8_t cand[kMaxLeaves * kMaxDims + 2];
    int nc = 0, free_dims = 0;
    for (int i = 0; i < pr.ndim; ++i) {
        int m = pr.at(i, l, l);
        if (m == l) {
            ++free_dims;
            continue;
        }
        do {
            if (tally[m]++ == 0) cand[nc++] = static_cast<uint8_t>(m);
            m = pr.at(i, l, b[m]);
        } while (m != l);
    }
    int k = 0;
    if (free_dims == pr.ndim) {
        for (int m = 0; m < l; ++m) out[k++] = static_cast<uint8_t>(m);
        return k;
    }
    const int need = pr.ndim - free_dims;
    for (int q = 0; q < nc; ++q)
        if (tally[cand[q]] == need) out[k++] = cand[q];
    return k;
}

// Lunnon lookahead: children of the node at leaf l counted into prof[k]
// (a completion, l > n, counts as one), recursing to depth kmax.
void lunnon_profile(const LunnonProblem &pr, uint8_t *a, uint8_t *b, int l, int k, int kmax, double *prof) {
    if (l > pr.n) return;
    uint8_t gaps[kMaxLeaves + 2];
    const int ng = lunnon_gaps(pr, b, l, gaps);
    prof[k] += ng;
    if (k + 1 >= kmax) return;
    for (int q = 0; q < ng; ++q) {
        const int below = gaps[q];
        a[l] = static_cast<uint8_t>(below);
        b[l] = b[below];
        b[below] = static_cast<uint8_t>(l);
        a[b[l]] = static_cast<uint8_t>(l);
        lunnon_profile(pr, a, b, l + 1, k + 1, kmax, prof);
        b[a[l]] = b[l];
        a[b[l]] = a[l];
    }
}

void lunnon_row(const LunnonProblem &pr, int depth, const uint8_t *state, float *f) {
    std::memset(f, 0, sizeof(float) * kFeatures);
    uint8_t a[kMaxLeaves + 2] = {0}, b[kMaxLeaves + 2] = {0};
    for (int m = 0; m <= depth; ++m) b[m] = state[m];
    for (int m = 0; m <= depth; ++m) a[b[m]] = static_cast<uint8_t>(m);
    int c = 0;
    for (int r = 0; r <= depth && r < 64; ++r) {
        f[c] = static_cast<float>(r + 1) / 64.0f;
        c = b[c];
        if (c == 0) break;
    }
    const int n = pr.n, l = depth + 1;
    double prof[kProfile] = {0};
    lunnon_profile(pr, a, b, l, 0, kProfile, prof);
    f[64] = l / 64.0f;
    f[65] = static_cast<float>(l) / n;
    f[66] = static_cast<float>(prof[0]) / 16.0f;
    f[67] = n / 64.0f;
    f[68] = 0.0f;
    f[69] = pr.ndim / 4.0f;
    f[70] = (n - l) / 64.0f;
    f[71] = pr.dims[0] / 16.0f;
    for (int k = 0; k < kProfile; ++k) f[72 + k] = lg(prof[k]) / 24.0f;
}

// estimate_work (sym_fold.cpp:235-250, sym_fold9.cpp:359-374): descendants
// of the node at level i that reach level stop.  M = uint64_t up to 8x8,
// unsigned __int128 for 9x9.
template <typename M>
M good_of(const SymProblem &pr, const uint8_t *cyc, int i);
template <>
uint64_t good_of<uint64_t>(const SymProblem &pr, const uint8_t *cyc, int i) { return sym_good(pr, cyc, i); }
template <>
unsigned __int128 good_of<unsigned __int128>(const SymProblem &pr, const uint8_t *cyc, int i) {
    return sym_good128(pr, cyc, i);
}
inline int lowest(uint64_t v) { return __builtin_ctzll(v); }
inline int lowest(unsigned __int128 v) {
    const uint64_t lo = static_cast<uint64_t>(v);
    return lo ? __builtin_ctzll(lo) : 64 + __builtin_ctzll(static_cast<uint64_t>(v >> 64));
}
inline uint64_t bits(uint64_t v) { return static_cast<uint64_t>(__builtin_popcountll(v)); }
inline uint64_t bits(unsigned __int128 v) {
    return static_cast<uint64_t>(__builtin_popcountll(static_cast<uint64_t>(v)) +
                                 __builtin_popcountll(static_cast<uint64_t>(v >> 64)));
}

template <typename M>
uint64_t descendants(const SymProblem &pr, uint8_t *cyc, int i, int stop) {
    if (i >= stop) return 1;
    const M good = good_of<M>(pr, cyc, i);
    if (i + 1 == stop) return bits(good);
    uint64_t out = 0;
    for (M gm = good; gm; gm &= gm - 1) {
        const int j = lowest(gm);
        cyc[i] = cyc[j];
        cyc[j] = static_cast<uint8_t>(i);
        out += descendants<M>(pr, cyc, i + 1, stop);
        cyc[j] = cyc[i];
    }
    return out;
}

template <typename Row>
void parallel_rows(int64_t count, Row &&row) {
    const int64_t threads = std::max<int64_t>(1, std::min<int64_t>(std::thread::hardware_concurrency(), 16));
    if (count < 256 || threads == 1) {
        for (int64_t k = 0; k < count; ++k) row(k);
        return;
    }
    std::vector<std::thread> pool;
    for (int64_t t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            for (int64_t k = t; k < count; k += threads) row(k);
        });
    for (auto &th : pool) th.join();
}

}  // namespace

// ====================================================================== C ABI

extern "C" {

int mf_cost_create(uint64_t seed, mf_cost_model **out) {
    if (!out) return lib_error(MF_ERR_ARG, "out is NULL");
    *out = nullptr;
    hipStream_t s;
    int cus, rc;
    if ((rc = context(&s, &cus)) != MF_OK) return rc;
    std::vector<float> p(kParams, 0.0f);
    uint64_t st = seed ^ 0x6d66636f73746d6cull;
    auto fill = [&](int off, int rows, int cols) {
        const float bound = std::sqrt(6.0f / static_cast<float>(cols));
        for (int k = 0; k < rows * cols; ++k) p[off + k] = (2.0f * uniform01(st) - 1.0f) * bound;
    };
    fill(kOffW1, kHidden, kFeatures);
    fill(kOffW2, kHidden, kHidden);
    fill(kOffW3, kHidden, kHidden);
    fill(kOffW4, 1, kHidden);
    auto *m = new mf_cost_model();
    const size_t pb = static_cast<size_t>(kParams) * 4;
    hipError_t e = hipSuccess;
    for (Buf *b : {&m->params, &m->grad, &m->m, &m->v})
        if (e == hipSuccess) e = b->ensure(pb);
    if (e == hipSuccess) e = m->frag.ensure(static_cast<size_t>(kFragHalves) * 2);
    if (e == hipSuccess) e = m->small.ensure(static_cast<size_t>(kSmallFloats) * 4);
    if (e == hipSuccess) e = hipMemcpyAsync(m->params.p, p.data(), pb, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemsetAsync(m->m.p, 0, pb, s);
    if (e == hipSuccess) e = hipMemsetAsync(m->v.p, 0, pb, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        mf_cost_destroy(m);
        std::string msg = std::string("cost model allocation failed: ") + hipGetErrorString(e);
        return lib_error(MF_ERR_HIP, msg.c_str());
    }
    *out = m;
    return MF_OK;
}

void mf_cost_destroy(mf_cost_model *m) {
    if (!m) return;
    for (Buf *b : {&m->params, &m->grad, &m->m, &m->v, &m->frag, &m->small, &m->data_x, &m->data_y, &m->pred})
        b->release();
    m->w.release();
    delete m;
}

int mf_cost_get_params(const mf_cost_model *m, float *params) {
    if (!m || !params) return lib_error(MF_ERR_ARG, "NULL argument");
    hipStream_t s;
    int cus, rc;
    if ((rc = context(&s, &cus)) != MF_OK) return rc;
    COST_TRY(hipMemcpyAsync(params, m->params.p, static_cast<size_t>(kParams) * 4, hipMemcpyDeviceToHost, s));
    COST_TRY(hipStreamSynchronize(s));
    return MF_OK;
}

int mf_cost_set_params(mf_cost_model *m, const float *params) {
    if (!m || !params) return lib_error(MF_ERR_ARG, "NULL argument");
    hipStream_t s;
    int cus, rc;
    if ((rc = context(&s, &cus)) != MF_OK) return rc;
    const size_t pb = static_cast<size_t>(kParams) * 4;
    COST_TRY(hipMemcpyAsync(m->params.p, params, pb, hipMemcpyHostToDevice, s));
    COST_TRY(hipMemsetAsync(m->m.p, 0, pb, s));
    COST_TRY(hipMemsetAsync(m->v.p, 0, pb, s));
    COST_TRY(hipStreamSynchronize(s));
    m->step = 0;
    m->frag_valid = false;
    return MF_OK;
}

int mf_cost_features_sym(int32_t n, int32_t level, const uint8_t *cycles, int64_t count, float *out) {
    if (count < 0 || (count > 0 && (!cycles || !out))) return lib_error(MF_ERR_ARG, "bad buffers");
    SymProblem pr;
    std::string err;
    if (!sym_problem(n, pr, err)) return lib_error(MF_ERR_ARG, err.c_str());
    if (n > 8) return lib_error(MF_ERR_ARG, "cost features take n <= 8");
    if (level < 1 || level >= pr.n2) return lib_error(MF_ERR_DEPTH, "level must be in [1, n*n)");
    parallel_rows(count, [&](int64_t k) { sym_row(pr, level, cycles + 64 * k, out + kFeatures * k); });
    return MF_OK;
}

int mf_cost_features_lunnon(const int32_t *dims, int32_t ndim, int32_t depth, const uint8_t *states,
                            int64_t count, float *out) {
    if (count < 0 || (count > 0 && (!states || !out))) return lib_error(MF_ERR_ARG, "bad buffers");
    LunnonProblem pr;
    std::string err;
    if (!lunnon_problem(dims, ndim, pr, err)) return lib_error(MF_ERR_ARG, err.c_str());
    if (depth < 1 || depth >= pr.n) return lib_error(MF_ERR_DEPTH, "depth must be in [1, n)");
    const int64_t stride = depth + 1;
    parallel_rows(count, [&](int64_t k) { lunnon_row(pr, depth, states + stride * k, out + kFeatures * k); });
    return MF_OK;
}

int mf_sym_estimate_work(int32_t n, int32_t level, const uint8_t *cycles, int64_t count, int32_t depth,
                         uint64_t *scores) {
    if (count < 0 || (count > 0 && (!cycles || !scores))) return lib_error(MF_ERR_ARG, "bad buffers");
    SymProblem pr;
    std::string err;
    if (!sym_problem(n, pr, err)) return lib_error(MF_ERR_ARG, err.c_str());
    if (level < 1 || level > pr.n2 || depth < 0) return lib_error(MF_ERR_DEPTH, "need 1 <= level <= n*n, depth >= 0");
    const int stop = std::min(pr.n2, level + depth);
    const int cb = pr.case_bytes();
    parallel_rows(count, [&](int64_t k) {
        uint8_t cyc[kMaxCells] = {};
        std::memcpy(cyc, cycles + static_cast<int64_t>(cb) * k, cb);
        scores[k] = pr.n2 <= 64 ? descendants<uint64_t>(pr, cyc, level, stop)
                                : descendants<unsigned __int128>(pr, cyc, level, stop);
    });
    return MF_OK;
}

int mf_cost_gradient(mf_cost_model *m, const float *x, const float *y, int64_t count, float *grad, float *loss) {
    if (!m || !x || !y || !grad || count <= 0 || count > (1 << 24)) return lib_error(MF_ERR_ARG, "bad arguments");
    hipStream_t s;
    int cus, rc;
    if ((rc = context(&s, &cus)) != MF_OK) return rc;
    const int rows = static_cast<int>(count);
    COST_TRY(m->w.ensure(rows, 1));
    COST_TRY(hipMemcpyAsync(m->w.x.p, x, static_cast<size_t>(rows) * kFeatures * 4, hipMemcpyHostToDevice, s));
    COST_TRY(hipMemcpyAsync(m->w.y.p, y, static_cast<size_t>(rows) * 4, hipMemcpyHostToDevice, s));
    COST_TRY(forward(m, rows, s));
    COST_TRY(backward(m, rows, m->w.loss.f(), s));
    COST_TRY(hipMemcpyAsync(grad, m->grad.p, static_cast<size_t>(kParams) * 4, hipMemcpyDeviceToHost, s));
    float l = 0;
    COST_TRY(hipMemcpyAsync(&l, m->w.loss.p, 4, hipMemcpyDeviceToHost, s));
    COST_TRY(hipStreamSynchronize(s));
    if (loss) *loss = l;
    return MF_OK;
}

int mf_cost_train(mf_cost_model *m, const float *x, const float *y, int64_t count, const mf_cost_train_options *opt,
                  float *epoch_loss) {
    if (!m || !x || !y || count <= 0 || count > (1 << 26)) return lib_error(MF_ERR_ARG, "bad arguments");
    hipStream_t s;
    int cus, rc;
    if ((rc = context(&s, &cus)) != MF_OK) return rc;
    const int epochs = (opt && opt->epochs > 0) ? opt->epochs : 20;
    const int batch = (opt && opt->batch > 0) ? std::min(opt->batch, 1 << 16) : 512;
    const float lr = (opt && opt->lr > 0) ? opt->lr : 1e-3f;
    const float b1 = (opt && opt->beta1 > 0) ? opt->beta1 : 0.9f;
    const float b2 = (opt && opt->beta2 > 0) ? opt->beta2 : 0.999f;
    const float eps = (opt && opt->eps > 0) ? opt->eps : 1e-8f;
    uint64_t st = opt ? opt->seed : 0;
    const int64_t N = count;
    const int nb = static_cast<int>((N + batch - 1) / batch);
    COST_TRY(m->data_x.ensure(static_cast<size_t>(N) * kFeatures * 4));
    COST_TRY(m->data_y.ensure(static_cast<size_t>(N) * 4));
    COST_TRY(m->w.ensure(batch, nb));
    COST_TRY(m->w.idx.ensure(static_cast<size_t>(N) * 4));
    COST_TRY(hipMemcpyAsync(m->data_x.p, x, static_cast<size_t>(N) * kFeatures * 4, hipMemcpyHostToDevice, s));
    COST_TRY(hipMemcpyAsync(m->data_y.p, y, static_cast<size_t>(N) * 4, hipMemcpyHostToDevice, s));
    std::vector<uint32_t> perm(N);
    std::vector<float> bl(nb);
    for (int ep = 0; ep < epochs; ++ep) {
        for (int64_t k = 0; k < N; ++k) perm[k] = static_cast<uint32_t>(k);
        for (int64_t k = N - 1; k > 0; --k) std::swap(perm[k], perm[splitmix(st) % static_cast<uint64_t>(k + 1)]);
        COST_TRY(hipMemcpyAsync(m->w.idx.p, perm.data(), static_cast<size_t>(N) * 4, hipMemcpyHostToDevice, s));
        for (int bi = 0; bi < nb; ++bi) {
            const int64_t lo = static_cast<int64_t>(bi) * batch;
            const int rows = static_cast<int>(std::min<int64_t>(batch, N - lo));
            const uint32_t *idx = static_cast<const uint32_t *>(m->w.idx.p) + lo;
            COST_TRY(cost_gather(rows, kFeatures, idx, m->data_x.f(), m->w.x.f(), s));
            COST_TRY(cost_gather(rows, 1, idx, m->data_y.f(), m->w.y.f(), s));
            COST_TRY(forward(m, rows, s));
            COST_TRY(backward(m, rows, m->w.loss.f() + bi, s));
            ++m->step;
            COST_TRY(cost_adam(kParams, m->params.f(), m->grad.f(), m->m.f(), m->v.f(), lr, b1, b2, eps, m->step, s));
        }
        COST_TRY(hipMemcpyAsync(bl.data(), m->w.loss.p, static_cast<size_t>(nb) * 4, hipMemcpyDeviceToHost, s));
        COST_TRY(hipStreamSynchronize(s));
        if (epoch_loss) {
            double acc = 0;
            for (int bi = 0; bi < nb; ++bi) {
                const int64_t lo = static_cast<int64_t>(bi) * batch;
                acc += static_cast<double>(bl[bi]) * static_cast<double>(std::min<int64_t>(batch, N - lo));
            }
            epoch_loss[ep] = static_cast<float>(acc / static_cast<double>(N));
        }
    }
    m->frag_valid = false;
    return MF_OK;
}

int mf_cost_predict(mf_cost_model *m, const float *x, int64_t count, float *pred) {
    if (!m || count < 0 || (count > 0 && (!x || !pred)) || count > (1 << 28)) return lib_error(MF_ERR_ARG, "bad arguments");
    if (count == 0) return MF_OK;
    hipStream_t s;
    int cus, rc;
    if ((rc = context(&s, &cus)) != MF_OK) return rc;
    if (!m->frag_valid) {
        COST_TRY(cost_pack_bf16(m->params.f(), static_cast<uint16_t *>(m->frag.p), m->small.f(), s));
        m->frag_valid = true;
    }
    const size_t N = static_cast<size_t>(count);
    COST_TRY(m->data_x.ensure(N * kFeatures * 4));
    COST_TRY(m->pred.ensure(N * 4));
    COST_TRY(hipMemcpyAsync(m->data_x.p, x, N * kFeatures * 4, hipMemcpyHostToDevice, s));
    COST_TRY(cost_infer_bf16(static_cast<int>(count), m->data_x.f(), static_cast<const uint16_t *>(m->frag.p),
                             m->small.f(), m->pred.f(), cus, s));
    COST_TRY(hipMemcpyAsync(pred, m->pred.p, N * 4, hipMemcpyDeviceToHost, s));
    COST_TRY(hipStreamSynchronize(s));
    return MF_OK;
}

int mf_cost_predict_frontier(mf_cost_model *m, mf_frontier *f, float *pred, int64_t max_rows, int64_t *rows) {
    if (!m || !f || !rows) return lib_error(MF_ERR_ARG, "bad arguments");
    hipStream_t s;
    int cus, rc;
    if ((rc = context(&s, &cus)) != MF_OK) return rc;
    uint64_t n = mf_frontier_items(f);
    *rows = static_cast<int64_t>(n);
    if (n == 0) return MF_OK;
    if (static_cast<int64_t>(n) > max_rows || !pred)
        return lib_error(MF_ERR_CAPACITY, "pred has room for fewer rows than the frontier has items");
    if (!m->frag_valid) {
        COST_TRY(cost_pack_bf16(m->params.f(), static_cast<uint16_t *>(m->frag.p), m->small.f(), s));
        m->frag_valid = true;
    }
    COST_TRY(m->data_x.ensure(n * kFeatures * 4));
    COST_TRY(m->pred.ensure(n * 4));
    uint64_t got = 0;
    if ((rc = frontier_sym_features(f, m->data_x.f(), &got)) != MF_OK) return rc;
    COST_TRY(cost_infer_bf16(static_cast<int>(n), m->data_x.f(), static_cast<const uint16_t *>(m->frag.p),
                             m->small.f(), m->pred.f(), cus, s));
    COST_TRY(hipMemcpyAsync(pred, m->pred.p, n * 4, hipMemcpyDeviceToHost, s));
    COST_TRY(hipStreamSynchronize(s));
    return MF_OK;
}

int mf_cost_features_frontier(mf_frontier *f, float *x, int64_t max_rows, int64_t *rows) {
    if (!f || !rows) return lib_error(MF_ERR_ARG, "bad arguments");
    hipStream_t s;
    int cus, rc;
    if ((rc = context(&s, &cus)) != MF_OK) return rc;
    const uint64_t n = mf_frontier_items(f);
    *rows = static_cast<int64_t>(n);
    if (n == 0) return MF_OK;
    if (static_cast<int64_t>(n) > max_rows || !x)
        return lib_error(MF_ERR_CAPACITY, "x has room for fewer rows than the frontier has items");
    Buf dev;
    COST_TRY(dev.ensure(n * kFeatures * 4));
    uint64_t got = 0;
    if ((rc = frontier_sym_features(f, dev.f(), &got)) != MF_OK) {
        dev.release();
        return rc;
    }
    hipError_t e = hipMemcpyAsync(x, dev.p, n * kFeatures * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    dev.release();
    COST_TRY(e);
    return MF_OK;
}

int mf_cost_predict_fp32(mf_cost_model *m, const float *x, int64_t count, float *pred) {
    if (!m || count < 0 || (count > 0 && (!x || !pred)) || count > (1 << 24)) return lib_error(MF_ERR_ARG, "bad arguments");
    if (count == 0) return MF_OK;
    hipStream_t s;
    int cus, rc;
    if ((rc = context(&s, &cus)) != MF_OK) return rc;
    const int chunk = 1 << 16;
    COST_TRY(m->w.ensure(chunk, 1));
    for (int64_t lo = 0; lo < count; lo += chunk) {
        const int rows = static_cast<int>(std::min<int64_t>(chunk, count - lo));
        COST_TRY(hipMemcpyAsync(m->w.x.p, x + lo * kFeatures, static_cast<size_t>(rows) * kFeatures * 4,
                                hipMemcpyHostToDevice, s));
        COST_TRY(forward(m, rows, s));
        COST_TRY(hipMemcpyAsync(pred + lo, m->w.pred.p, static_cast<size_t>(rows) * 4, hipMemcpyDeviceToHost, s));
    }
    COST_TRY(hipStreamSynchronize(s));
    return MF_OK;
}

}  // extern "C"
