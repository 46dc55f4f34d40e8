// host_lunnon.cpp -- Lunnon tables and the canonical frontier collector.
//
// Tables follow mf.c:29-75 (bigP, coordinates c, connection table d); the
// collector follows the canonical list encoding of mf_fast.cpp:185-221 /
// gpu_folding_v2.cu:220-256: a circular doubly-linked list over the placed
// leaves 0..l-1 rooted at leaf 0, gap candidates on a LIFO stack, and the
// normal-form filter b[0] == 1 (mf.c:83).  A frontier state is the node at
// leaf D+1; it is stored as its "below" links b[0..D] (a[] is the inverse).
#include <cstring>
#include <vector>

#include "mf_host.hpp"

namespace mf {

bool lunnon_problem(const int32_t *dims, int ndim, LunnonProblem &pr, std::string &err) {
    if (ndim < 1 || ndim > kMaxDims) {
        err = "number of dimensions must be 1.." + std::to_string(kMaxDims);
        return false;
    }
    long n = 1;
    for (int i = 0; i < ndim; ++i) {
        if (dims[i] < 1) {
            err = "dimensions must be positive";
            return false;
        }
        n *= dims[i];
        if (n > kMaxLeaves) {
            err = "map too large: at most 64 leaves";
            return false;
        }
    }
    pr.ndim = ndim;
    pr.n = static_cast<int>(n);
    for (int i = 0; i < kMaxDims; ++i) pr.dims[i] = i < ndim ? dims[i] : 1;
    const int s = pr.n + 1;
    pr.d.assign(static_cast<size_t>(ndim) * s * s, 0);

    // Leaf m (1-based) has mixed-radix digits; digit i is its coordinate in
    // dimension i.  big[i] is the place value of digit i.
    int big[kMaxDims + 1];
    big[0] = 1;
    for (int i = 0; i < ndim; ++i) big[i + 1] = big[i] * dims[i];
    auto coord = [&](int i, int m) { return ((m - 1) / big[i]) % dims[i] + 1; };

    for (int i = 0; i < ndim; ++i) {
        for (int l = 1; l <= pr.n; ++l) {
            const int cl = coord(i, l);
            for (int m = 1; m <= l; ++m) {
                const int cm = coord(i, m);
                int v = m;
                if (((cl - cm) & 1) == 0) {
                    if (cm != 1) v = m - big[i];
                } else if (cm != dims[i] && m + big[i] <= l) {
                    v = m + big[i];
                }
                pr.d[(static_cast<size_t>(i) * s + l) * s + m] = static_cast<uint8_t>(v);
            }
        }
    }
    return true;
}

namespace {

// Explicit-stack Lunnon DFS over leaves 1..depth that reports each node at
// leaf depth+1 (after the normal-form filter) to `emit`.
class FrontierWalker {
  public:
    FrontierWalker(const LunnonProblem &pr, int depth) : pr_(pr), depth_(depth) {
        std::memset(a_, 0, sizeof(a_));
        std::memset(b_, 0, sizeof(b_));
        std::memset(tally_, 0, sizeof(tally_));
        std::memset(base_, 0, sizeof(base_));
        gap_.resize(static_cast<size_t>(pr.n) * pr.n + 2);
    }

    template <typename Emit>
    void run(Emit &&emit) {
        int leaf = 1;
        int top = 0;
        for (;;) {
            if (leaf <= 1 || b_[0] == 1) {
                if (leaf > depth_) {
                    emit(b_);
                    top = base_[leaf - 1];
                } else {
                    top = push_gaps(leaf, base_[leaf - 1]);
                }
            }
            while (top == base_[leaf - 1]) {
                if (--leaf == 0) return;
                unlink(leaf);
            }
            link(leaf, gap_[--top]);
            base_[leaf] = top;
            ++leaf;
        }
    }

  private:
    // Push the legal positions for `leaf` (those reached by the walk of every
    // dimension that constrains it) on the gap stack; return the new top.
    int push_gaps(int leaf, int top) {
        const int ndim = pr_.ndim;
        int free_dims = 0;
        int end = top;
        for (int i = 0; i < ndim; ++i) {
            int m = pr_.at(i, leaf, leaf);
            if (m == leaf) {
                ++free_dims;
                continue;
            }
            do {
                gap_[end] = static_cast<uint8_t>(m);
                if (tally_[m]++ == 0) ++end;
                m = pr_.at(i, leaf, b_[m]);
            } while (m != leaf);
        }
        if (free_dims == ndim)
            for (int m = 0; m < leaf; ++m) gap_[end++] = static_cast<uint8_t>(m);
        const int need = ndim - free_dims;
        int out = top;
        for (int k = top; k < end; ++k) {
            const int m = gap_[k];
            if (tally_[m] == need) gap_[out++] = static_cast<uint8_t>(m);
            tally_[m] = 0;
        }
        return out;
    }

    void link(int leaf, int below_of) {
        a_[leaf] = static_cast<uint8_t>(below_of);
        b_[leaf] = b_[below_of];
        b_[below_of] = static_cast<uint8_t>(leaf);
        a_[b_[leaf]] = static_cast<uint8_t>(leaf);
    }

    void unlink(int leaf) {
        b_[a_[leaf]] = b_[leaf];
        a_[b_[leaf]] = a_[leaf];
    }

    const LunnonProblem &pr_;
    int depth_;
    uint8_t a_[kMaxLeaves + 2];
    uint8_t b_[kMaxLeaves + 2];
    uint8_t tally_[kMaxLeaves + 2];
    int base_[kMaxLeaves + 2];
    std::vector<uint8_t> gap_;
};

}  // namespace

bool lunnon_reachable(const LunnonProblem &pr, const uint8_t *state, int depth) {
    if (depth < 1 || depth >= pr.n || state[0] != 1) return false;
    uint8_t b[kMaxLeaves + 2];
    std::memcpy(b, state, static_cast<size_t>(depth) + 1);
    // One circular list through exactly the leaves 0..depth.
    uint64_t seen = 0;
    int c = 0;
    for (int k = 0; k <= depth; ++k) {
        if (b[c] > depth) return false;
        seen |= 1ull << c;
        c = b[c];
    }
    if (c != 0 || seen != (depth >= 63 ? ~0ull : (2ull << depth) - 1)) return false;
    for (int l = depth; l >= 2; --l) {
        int m = 0;   // the leaf directly above l
        while (b[m] != l) ++m;
        if (m == 0) return false;   // below the root: the normal form never inserts there
        b[m] = b[l];
        // Gap walks of leaf l in the list of leaves 0..l-1 (push_gaps above),
        // each bounded by l steps.
        uint64_t meet = ~0ull;
        for (int i = 0; i < pr.ndim; ++i) {
            int x = pr.at(i, l, l);
            if (x == l) continue;
            uint64_t reach = 0;
            int steps = 0;
            do {
                reach |= 1ull << x;
                x = pr.at(i, l, b[x]);
                if (++steps > l) return false;
            } while (x != l);
            meet &= reach;
        }
        if (((meet >> m) & 1ull) == 0) return false;
    }
    return true;
}

int64_t lunnon_frontier(const LunnonProblem &pr, int depth, uint8_t *out, int64_t max_states) {
    FrontierWalker w(pr, depth);
    int64_t count = 0;
    const int row = depth + 1;
    w.run([&](const uint8_t *b) {
        if (out != nullptr && count < max_states) std::memcpy(out + count * row, b, static_cast<size_t>(row));
        ++count;
    });
    return count;
}

}  // namespace mf
