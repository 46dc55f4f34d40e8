// host_sym.cpp -- symmetry-reduced cycle DFS: tables, seeds and the case stream.
//
// Cell numbering, edge pages, get_good and child order follow sym_fold.cpp
// (build_map :122-145, make_edges :147-173, get_good :179-204, iter_good
// :206-217).  Seeds are the center cycles deduplicated under the 8 symmetries
// of the square and stack reversal, first representative kept, in generation
// order (:316-408, :689-710).  The case stream enumerates each seed's subtree
// to level `expanded` in seed order, children in ascending insertion point
// (recurse_stream_cases :491-527); the position in that stream is the case
// index that every chunk [start, start+max) refers to.
//
// Unlike the reference stream, which walks every case before `start`, the
// stream here first looks up a per-seed case-count index, skips whole seeds,
// then enumerates only inside the seed that holds `start`.
#include <algorithm>
#include <array>
#include <atomic>
#include <cstring>
#include <map>
#include <mutex>
#include <set>
#include <thread>
#include <utility>
#include <vector>

#include "mf_host.hpp"

namespace mf {

namespace {

void spiral(int w, int h, int *map) {
    // Walk from corner (0,0) along +x, turning left at the border or at an
    // already numbered cell; numbers count down from w*h-1 so that the center
    // ends at 0.
    const int unset = w * h;
    for (int k = 0; k < w * h; ++k) map[k] = unset;
    int x = 0, y = 0, dx = 1, dy = 0;
    for (int num = w * h - 1; num >= 0; --num) {
        map[x * h + y] = num;
        const int nx = x + dx, ny = y + dy;
        const bool blocked = nx < 0 || nx >= w || ny < 0 || ny >= h || map[nx * h + ny] != unset;
        if (blocked) {
            const int t = dy;
            dy = dx;
            dx = -t;
        }
        x += dx;
        y += dy;
    }
}

inline uint64_t below(int i) { return i >= 64 ? ~0ull : ((1ull << i) - 1ull); }

}  // namespace

bool sym_problem(int n, SymProblem &pr, std::string &err) {
    if (n < 1 || n > 8) {
        err = "n must be between 1 and 8";
        return false;
    }
    pr.n = n;
    pr.n2 = n * n;
    std::memset(pr.edge, pr.n2, sizeof(pr.edge));
    spiral(n, n, pr.map);
    for (int x = 0; x < n; ++x) {
        for (int y = 0; y < n; ++y) {
            const int c = pr.map[x * n + y];
            if (x > 0) {  // x-neighbour pair on page x%2
                const int o = pr.map[(x - 1) * n + y];
                pr.edge[x & 1][c] = static_cast<uint8_t>(o);
                pr.edge[x & 1][o] = static_cast<uint8_t>(c);
            }
            if (y > 0) {  // y-neighbour pair on page 2 + y%2
                const int o = pr.map[x * n + y - 1];
                pr.edge[2 + (y & 1)][c] = static_cast<uint8_t>(o);
                pr.edge[2 + (y & 1)][o] = static_cast<uint8_t>(c);
            }
        }
    }
    for (int c = 0; c < kCells; ++c) {
        uint32_t w = 0;
        for (int p = 0; p < 4; ++p) {
            const uint32_t e = c < pr.n2 ? pr.edge[p][c] : static_cast<uint32_t>(pr.n2);
            w |= e << (8 * p);
        }
        pr.edge_word[c] = w;
    }
    return true;
}

uint64_t sym_good(const SymProblem &pr, const uint8_t *cycle, int i) {
    uint64_t good = below(i);
    for (int p = 0; p < 4 && good; ++p) {
        const int anchor = pr.edge[p][i];
        if (anchor >= i) continue;
        // Walk the stack cycle from the anchor; a cell whose page-p partner is
        // already placed is replaced by that partner (a closed crease).
        uint64_t reach = 1ull << anchor;
        for (int j = cycle[anchor]; j != anchor; j = cycle[j]) {
            const int partner = pr.edge[p][j];
            if (partner < i) j = partner;
            reach |= 1ull << j;
        }
        good &= reach;
    }
    return good;
}

// --------------------------------------------------------------- seeds

namespace {

using Key = std::array<uint8_t, 16>;

Key center_key(const uint8_t *cycle, int c2) {
    Key k;
    k.fill(255);
    for (int q = 0; q < c2; ++q) k[q] = cycle[q];
    return k;
}

// The 8 symmetries of the c x c center, as maps on spiral numbers, in the
// reference's order: identity, transpose, y-flip, rotations and flips.
std::vector<std::array<uint8_t, 16>> center_symmetries(int c) {
    int map[16];
    spiral(c, c, map);
    int px[16], py[16];
    for (int x = 0; x < c; ++x)
        for (int y = 0; y < c; ++y) {
            px[map[x * c + y]] = x;
            py[map[x * c + y]] = y;
        }
    const int e = c - 1;
    std::vector<std::array<uint8_t, 16>> out;
    for (int t = 0; t < 8; ++t) {
        std::array<uint8_t, 16> perm;
        perm.fill(255);
        for (int q = 0; q < c * c; ++q) {
            const int x = px[q], y = py[q];
            int tx, ty;
            switch (t) {
                case 0: tx = x, ty = y; break;
                case 1: tx = y, ty = x; break;
                case 2: tx = x, ty = e - y; break;
                case 3: tx = y, ty = e - x; break;
                case 4: tx = e - x, ty = y; break;
                case 5: tx = e - y, ty = x; break;
                case 6: tx = e - x, ty = e - y; break;
                default: tx = e - y, ty = e - x; break;
            }
            perm[q] = static_cast<uint8_t>(map[tx * c + ty]);
        }
        out.push_back(perm);
    }
    return out;
}

void grow_centers(const SymProblem &pr, uint8_t *cycle, int i, int stop, std::vector<std::array<uint8_t, kCells>> &out) {
    if (i == stop) {
        std::array<uint8_t, kCells> c;
        std::memcpy(c.data(), cycle, kCells);
        out.push_back(c);
        return;
    }
    for (uint64_t g = sym_good(pr, cycle, i); g; g &= g - 1) {
        const int j = __builtin_ctzll(g);
        const uint8_t nx = cycle[j];
        cycle[i] = nx;
        cycle[j] = static_cast<uint8_t>(i);
        grow_centers(pr, cycle, i + 1, stop, out);
        cycle[j] = nx;
    }
}

}  // namespace

std::vector<SymSeed> sym_seeds(const SymProblem &pr, int64_t *raw) {
    std::vector<SymSeed> seeds;
    if (pr.n < 6) {
        if (raw) *raw = 0;
        return seeds;
    }
    const int c = (pr.n % 2 == 0) ? 4 : 3;
    const int c2 = c * c;
    uint8_t cycle[kCells] = {1, 2, 0};
    std::vector<std::array<uint8_t, kCells>> centers;
    grow_centers(pr, cycle, 3, c2, centers);
    if (raw) *raw = static_cast<int64_t>(centers.size());

    const auto syms = center_symmetries(c);
    std::set<Key> seen;
    for (const auto &cyc : centers) {
        if (seen.count(center_key(cyc.data(), c2))) continue;
        // Orbit: each symmetry image, and each image with the stack reversed.
        std::set<Key> orbit;
        for (const auto &perm : syms) {
            uint8_t img[kCells], rev[kCells];
            std::memcpy(img, cyc.data(), kCells);
            for (int q = 0; q < c2; ++q) img[perm[q]] = perm[cyc[q]];
            std::memcpy(rev, img, kCells);
            int q = 0;
            do {
                rev[img[q]] = static_cast<uint8_t>(q);
                q = img[q];
            } while (q != 0);
            orbit.insert(center_key(img, c2));
            orbit.insert(center_key(rev, c2));
        }
        SymSeed s;
        std::memcpy(s.cycle, cyc.data(), kCells);
        s.mult = orbit.size();
        seeds.push_back(s);
        seen.insert(orbit.begin(), orbit.end());
    }
    return seeds;
}

// --------------------------------------------------------------- case index

namespace {

uint64_t count_below(const SymProblem &pr, uint8_t *cycle, int i, int stop) {
    if (i == stop) return 1;
    uint64_t total = 0;
    for (uint64_t g = sym_good(pr, cycle, i); g; g &= g - 1) {
        const int j = __builtin_ctzll(g);
        const uint8_t nx = cycle[j];
        cycle[i] = nx;
        cycle[j] = static_cast<uint8_t>(i);
        total += count_below(pr, cycle, i + 1, stop);
        cycle[j] = nx;
    }
    return total;
}

std::mutex g_index_mu;
std::map<std::pair<int, int>, std::vector<uint64_t>> g_index;
std::map<int, std::vector<SymSeed>> g_seeds;

const std::vector<SymSeed> &cached_seeds(const SymProblem &pr) {
    auto it = g_seeds.find(pr.n);
    if (it == g_seeds.end()) it = g_seeds.emplace(pr.n, sym_seeds(pr, nullptr)).first;
    return it->second;
}

// Extend counts (cases per seed at `expanded`, in seed order) until they cover
// case index `upto` or every seed is counted.  Seeds are counted in parallel
// batches; only the prefix of the stream that is needed is ever enumerated.
void extend_counts(const SymProblem &pr, int expanded, uint64_t upto, std::vector<uint64_t> &counts) {
    const auto &seeds = cached_seeds(pr);
    const int c2 = (pr.n % 2 == 0) ? 16 : 9;
    uint64_t covered = 0;
    for (uint64_t c : counts) covered += c;
    unsigned hw = std::thread::hardware_concurrency();
    const size_t threads = std::max<size_t>(1, std::min<size_t>(hw == 0 ? 1 : hw, 64));
    while (counts.size() < seeds.size() && covered <= upto) {
        const size_t first = counts.size();
        const size_t batch = std::min(seeds.size() - first, threads * 4);
        std::vector<uint64_t> part(batch, 0);
        std::atomic<size_t> next{0};
        std::vector<std::thread> pool;
        for (size_t t = 0; t < std::min(threads, batch); ++t) {
            pool.emplace_back([&]() {
                for (;;) {
                    const size_t k = next.fetch_add(1);
                    if (k >= batch) break;
                    uint8_t cyc[kCells];
                    std::memcpy(cyc, seeds[first + k].cycle, kCells);
                    part[k] = count_below(pr, cyc, c2, expanded);
                }
            });
        }
        for (auto &th : pool) th.join();
        for (uint64_t c : part) {
            counts.push_back(c);
            covered += c;
        }
    }
}

}  // namespace

const std::vector<uint64_t> &sym_seed_case_counts(const SymProblem &pr, int expanded) {
    std::lock_guard<std::mutex> lock(g_index_mu);
    auto &counts = g_index[std::make_pair(pr.n, expanded)];
    extend_counts(pr, expanded, ~0ull, counts);
    return counts;
}

namespace {

struct Stream {
    const SymProblem &pr;
    int stop;
    uint64_t skip;      // cases still to pass over before the window
    uint64_t want;      // cases still to write
    uint64_t mult;
    uint8_t *cycles;
    uint64_t *mults;
    int64_t written = 0;

    // Returns true when the window is full.
    bool visit(uint8_t *cycle, int i) {
        if (i == stop) {
            if (skip > 0) {
                --skip;
                return false;
            }
            std::memcpy(cycles + written * kCells, cycle, kCells);
            mults[written] = mult;
            ++written;
            return --want == 0;
        }
        for (uint64_t g = sym_good(pr, cycle, i); g; g &= g - 1) {
            const int j = __builtin_ctzll(g);
            const uint8_t nx = cycle[j];
            cycle[i] = nx;
            cycle[j] = static_cast<uint8_t>(i);
            const bool full = visit(cycle, i + 1);
            cycle[j] = nx;
            if (full) return true;
        }
        return false;
    }
};

}  // namespace

int64_t sym_cases(const SymProblem &pr, int expanded, uint64_t start, uint64_t max_cases, uint8_t *cycles,
                  uint64_t *mults) {
    if (max_cases == 0) return 0;
    std::vector<uint64_t> counts;
    const std::vector<SymSeed> *seeds;
    {
        std::lock_guard<std::mutex> lock(g_index_mu);
        auto &cached = g_index[std::make_pair(pr.n, expanded)];
        extend_counts(pr, expanded, start, cached);
        counts = cached;
        seeds = &cached_seeds(pr);
    }
    const int c2 = (pr.n % 2 == 0) ? 16 : 9;
    size_t s = 0;
    uint64_t skip = start;
    while (s < counts.size() && counts[s] <= skip) skip -= counts[s++];
    if (s == counts.size() && s == seeds->size()) return 0;
    Stream st{pr, expanded, skip, max_cases, 0, cycles, mults};
    for (; s < seeds->size(); ++s) {
        uint8_t cyc[kCells];
        std::memcpy(cyc, (*seeds)[s].cycle, kCells);
        st.mult = (*seeds)[s].mult;
        if (st.visit(cyc, c2)) break;
    }
    return st.written;
}

}  // namespace mf
