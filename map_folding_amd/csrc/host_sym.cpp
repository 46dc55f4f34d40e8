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
//
// Up to 8x8 the insertion masks are u64 and a case row is 64 bytes; 9x9
// (experiments/9x9/sym_fold9.cpp) uses u128 masks, 128-byte rows and, by
// default, a 5x5 center (MF_CENTER_N selects 3 or 5).
#include <algorithm>
#include <array>
#include <atomic>
#include <cstring>
#include <map>
#include <cstdlib>
#include <mutex>
#include <set>
#include <thread>
#include <tuple>
#include <unordered_set>
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

// Cell numbering (build_map, sym_fold9.cpp:160-229): the spiral above or,
// for odd squares under MF_ORDER, a ring-wise renumbering.  Rings of
// Chebyshev radius r around the center come inner first, each in spiral
// order, except that rings beyond the locked radius reorder: reverse;
// opposite (first and second half interleaved, a leftover last); axes_first
// (stable by Manhattan distance from the center); corners_first (the four
// corners first, stable).  The spiral itself is this construction with every
// ring kept in order, since spiral numbers grow ring by ring.
void cell_map(int w, int h, int order, int lock_radius, int *map) {
    spiral(w, h, map);
    if (order == 0 || w != h || w % 2 == 0) return;
    const int n = w, c = n / 2;
    std::vector<std::pair<int, int>> coords(n * n), ordered;
    for (int x = 0; x < n; ++x)
        for (int y = 0; y < n; ++y) coords[map[x * n + y]] = {x, y};
    for (int r = 0; r <= c; ++r) {
        std::vector<std::pair<int, int>> ring;
        for (const auto &xy : coords)
            if (std::max(std::abs(xy.first - c), std::abs(xy.second - c)) == r) ring.push_back(xy);
        if (r > lock_radius) {
            if (order == 1) {
                std::reverse(ring.begin(), ring.end());
            } else if (order == 2) {
                std::vector<std::pair<int, int>> mixed;
                const size_t half = ring.size() / 2;
                for (size_t k = 0; k < half; ++k) {
                    mixed.push_back(ring[k]);
                    if (k + half < ring.size()) mixed.push_back(ring[k + half]);
                }
                if (ring.size() % 2 != 0) mixed.push_back(ring.back());
                ring = std::move(mixed);
            } else if (order == 3) {
                std::stable_sort(ring.begin(), ring.end(), [c](const auto &a, const auto &b) {
                    return std::abs(a.first - c) + std::abs(a.second - c) < std::abs(b.first - c) + std::abs(b.second - c);
                });
            } else if (order == 4) {
                auto corner = [c, r](const std::pair<int, int> &a) {
                    return std::abs(a.first - c) == r && std::abs(a.second - c) == r ? 0 : 1;
                };
                std::stable_sort(ring.begin(), ring.end(),
                                 [&](const auto &a, const auto &b) { return corner(a) < corner(b); });
            }
        }
        ordered.insert(ordered.end(), ring.begin(), ring.end());
    }
    for (int k = 0; k < n * n; ++k) map[ordered[k].first * n + ordered[k].second] = k;
}

std::mutex g_numbering_mu;
SymNumbering g_numbering;

using u128 = unsigned __int128;

template <typename M>
inline M below(int i) {
    return i >= static_cast<int>(8 * sizeof(M)) ? ~static_cast<M>(0) : ((static_cast<M>(1) << i) - 1);
}
template <typename M>
inline int ctz_mask(M v) {
    if (sizeof(M) == 8) return __builtin_ctzll(static_cast<uint64_t>(v));
    const uint64_t lo = static_cast<uint64_t>(v);
    return lo ? __builtin_ctzll(lo) : 64 + __builtin_ctzll(static_cast<uint64_t>(static_cast<u128>(v) >> 64));
}

// get_good (sym_fold.cpp:179-204; u128 as sym_fold9.cpp:286-311).
template <typename M>
M good_t(const SymProblem &pr, const uint8_t *cycle, int i) {
    M good = below<M>(i);
    for (int p = 0; p < 4 && good; ++p) {
        const int anchor = pr.edge[p][i];
        if (anchor >= i) continue;
        // Walk the stack cycle from the anchor; a cell whose page-p partner is
        // already placed is replaced by that partner (a closed crease).
        M reach = static_cast<M>(1) << anchor;
        for (int j = cycle[anchor]; j != anchor; j = cycle[j]) {
            const int partner = pr.edge[p][j];
            if (partner < i) j = partner;
            reach |= static_cast<M>(1) << j;
        }
        good &= reach;
    }
    return good;
}

int default_center(int n) { return n >= 9 ? 5 : ((n % 2 == 0) ? 4 : 3); }

}  // namespace

SymNumbering sym_numbering() {
    std::lock_guard<std::mutex> lock(g_numbering_mu);
    return g_numbering;
}

bool set_sym_numbering(const SymNumbering &num, std::string &err) {
    if (num.center != 0 && (num.center < 3 || num.center > 5)) {
        err = "center must be 0 (default) or between 3 and 5";
        return false;
    }
    if (num.order < 0 || num.order > 4) {
        err = "order must be 0 (spiral) .. 4 (corners_first)";
        return false;
    }
    if (num.lock_center_n < 0) {
        err = "lock_center_n must not be negative";
        return false;
    }
    std::lock_guard<std::mutex> lock(g_numbering_mu);
    g_numbering = num;
    return true;
}

bool sym_problem(int n, SymProblem &pr, std::string &err) { return sym_problem(n, sym_numbering(), pr, err); }

bool sym_problem(int n, const SymNumbering &num, SymProblem &pr, std::string &err) {
    if (n < 1 || n > 9) {
        err = "n must be between 1 and 9";
        return false;
    }
    pr.n = n;
    pr.n2 = n * n;
    pr.center = default_center(n);
    if (num.center != 0 && n >= 6) {
        const int c = num.center;
        if (c < 3 || c > n || c * c > 25 || (c % 2) != (n % 2)) {
            err = "the seed center must be between 3 and n, have the same parity as n, and be at most 5";
            return false;
        }
        pr.center = c;
    }
    pr.order = num.order;
    pr.lock_radius = std::max(0, (num.lock_center_n - 1) / 2);
    if (pr.order == 0 || n % 2 == 0) pr.order = pr.lock_radius = 0;   // numbering is the spiral
    std::memset(pr.edge, pr.n2, sizeof(pr.edge));
    cell_map(n, n, pr.order, pr.lock_radius, pr.map);
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
    for (int c = 0; c < kMaxCells; ++c) {
        uint32_t w = 0;
        for (int p = 0; p < 4; ++p) {
            const uint32_t e = c < pr.n2 ? pr.edge[p][c] : static_cast<uint32_t>(pr.n2);
            w |= e << (8 * p);
        }
        pr.edge_word[c] = w;
    }
    return true;
}

uint64_t sym_good(const SymProblem &pr, const uint8_t *cycle, int i) { return good_t<uint64_t>(pr, cycle, i); }
unsigned __int128 sym_good128(const SymProblem &pr, const uint8_t *cycle, int i) { return good_t<u128>(pr, cycle, i); }

// --------------------------------------------------------------- seeds

namespace {

using Key = std::array<uint8_t, 25>;   // center cells (c*c <= 25)
using Row = std::array<uint8_t, kMaxCells>;

Key center_key(const uint8_t *cycle, int c2) {
    Key k;
    k.fill(255);
    for (int q = 0; q < c2; ++q) k[q] = cycle[q];
    return k;
}

// The 8 symmetries of the c x c center, as maps on cell numbers (the c x c
// grid numbered as the whole board is, generate_permutations,
// sym_fold9.cpp:506-537), in the reference's order: identity, transpose,
// y-flip, rotations and flips.
std::vector<std::array<uint8_t, 25>> center_symmetries(const SymProblem &pr, int c) {
    int map[25];
    cell_map(c, c, pr.order, pr.lock_radius, map);
    int px[25], py[25];
    for (int x = 0; x < c; ++x)
        for (int y = 0; y < c; ++y) {
            px[map[x * c + y]] = x;
            py[map[x * c + y]] = y;
        }
    const int e = c - 1;
    std::vector<std::array<uint8_t, 25>> out;
    for (int t = 0; t < 8; ++t) {
        std::array<uint8_t, 25> perm;
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

// Center cycles in generation order, as keys (cells >= c*c of a center are
// never written, so the key is the whole cycle).
template <typename M>
void grow_centers(const SymProblem &pr, uint8_t *cycle, int i, int stop, std::vector<Key> &out) {
    if (i == stop) {
        out.push_back(center_key(cycle, stop));
        return;
    }
    for (M g = good_t<M>(pr, cycle, i); g; g &= g - 1) {
        const int j = ctz_mask(g);
        const uint8_t nx = cycle[j];
        cycle[i] = nx;
        cycle[j] = static_cast<uint8_t>(i);
        grow_centers<M>(pr, cycle, i + 1, stop, out);
        cycle[j] = nx;
    }
}

// Prefixes of the center DFS at level `split`, in DFS order.
template <typename M>
void split_prefixes(const SymProblem &pr, uint8_t *cycle, int i, int split, std::vector<Row> &out) {
    if (i == split) {
        Row r;
        std::memcpy(r.data(), cycle, kMaxCells);
        out.push_back(r);
        return;
    }
    for (M g = good_t<M>(pr, cycle, i); g; g &= g - 1) {
        const int j = ctz_mask(g);
        const uint8_t nx = cycle[j];
        cycle[i] = nx;
        cycle[j] = static_cast<uint8_t>(i);
        split_prefixes<M>(pr, cycle, i + 1, split, out);
        cycle[j] = nx;
    }
}

template <typename Fn>
void parallel_for(size_t count, Fn &&fn) {
    const unsigned hw = std::thread::hardware_concurrency();
    const size_t threads = std::max<size_t>(1, std::min<size_t>({hw == 0 ? 1 : hw, 64, count}));
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (size_t t = 0; t < threads; ++t)
        pool.emplace_back([&]() {
            for (size_t k; (k = next.fetch_add(1)) < count;) fn(k);
        });
    for (auto &th : pool) th.join();
}

// Every center cycle at level c*c, in the order of the sequential DFS: the
// subtrees under the prefixes at a split level are grown in parallel and
// concatenated in prefix order.
template <typename M>
std::vector<Key> all_centers(const SymProblem &pr, int c2) {
    uint8_t cycle[kMaxCells] = {1, 2, 0};
    const int split = std::min(c2, 3 + 8);
    std::vector<Row> prefixes;
    split_prefixes<M>(pr, cycle, 3, split, prefixes);
    std::vector<std::vector<Key>> parts(prefixes.size());
    parallel_for(prefixes.size(), [&](size_t k) {
        Row r = prefixes[k];
        grow_centers<M>(pr, r.data(), split, c2, parts[k]);
    });
    size_t total = 0;
    for (const auto &p : parts) total += p.size();
    std::vector<Key> out;
    out.reserve(total);
    for (auto &p : parts) {
        out.insert(out.end(), p.begin(), p.end());
        std::vector<Key>().swap(p);
    }
    return out;
}

struct KeyHash {
    size_t operator()(const Key &k) const {
        uint64_t h = 1469598103934665603ull;
        for (uint8_t b : k) h = (h ^ b) * 1099511628211ull;
        return static_cast<size_t>(h);
    }
};

std::mutex g_seed_mu;
// (n, center, order, lock radius) -> (seeds, raw centers)
std::map<std::tuple<int, int, int, int>, std::pair<std::vector<SymSeed>, int64_t>> g_seed_sets;

}  // namespace

// The 8 symmetries with and without stack reversal form a group (the
// dihedral group of the square times reversal), so the images of a center
// form its orbit and orbits partition the centers.  Keeping the first center
// of each orbit in generation order (with the orbit size as multiplier) is
// then what the reference's seen-set pass keeps: a center is skipped exactly
// when an earlier one lies in its orbit.  Each center's orbit (its canonical
// member = the smallest image, and the orbit size) is found in parallel; one
// sequential pass over the canonical keys picks the first of each.
std::vector<SymSeed> sym_seeds(const SymProblem &pr, int64_t *raw) {
    std::vector<SymSeed> seeds;
    if (pr.n < 6) {
        if (raw) *raw = 0;
        return seeds;
    }
    const int c = pr.center;
    const int c2 = c * c;
    const std::vector<Key> centers = pr.n2 <= 64 ? all_centers<uint64_t>(pr, c2) : all_centers<u128>(pr, c2);
    if (raw) *raw = static_cast<int64_t>(centers.size());

    const auto syms = center_symmetries(pr, c);
    std::vector<Key> canon(centers.size());
    std::vector<uint8_t> orbit_size(centers.size());
    constexpr size_t kBlock = 4096;
    parallel_for((centers.size() + kBlock - 1) / kBlock, [&](size_t blk) {
        const size_t end = std::min(centers.size(), (blk + 1) * kBlock);
        for (size_t k = blk * kBlock; k < end; ++k) {
            const Key &cyc = centers[k];
            Key orbit[16];
            int no = 0;
            auto add = [&](const Key &key) {
                for (int q = 0; q < no; ++q)
                    if (orbit[q] == key) return;
                orbit[no++] = key;
            };
            for (const auto &perm : syms) {
                Key img, rev;
                img.fill(255);
                rev.fill(255);
                for (int q = 0; q < c2; ++q) img[perm[q]] = perm[cyc[q]];
                int q = 0;
                do {
                    rev[img[q]] = static_cast<uint8_t>(q);
                    q = img[q];
                } while (q != 0);
                add(img);
                add(rev);
            }
            canon[k] = *std::min_element(orbit, orbit + no);
            orbit_size[k] = static_cast<uint8_t>(no);
        }
    });

    std::unordered_set<Key, KeyHash> seen;
    seen.reserve(centers.size() / 4 + 16);
    for (size_t k = 0; k < centers.size(); ++k) {
        if (!seen.insert(canon[k]).second) continue;
        SymSeed s;
        std::memset(s.cycle, 0, kMaxCells);
        std::memcpy(s.cycle, centers[k].data(), c2);
        s.mult = orbit_size[k];
        seeds.push_back(s);
    }
    return seeds;
}

const std::vector<SymSeed> &sym_seed_list(const SymProblem &pr, int64_t *raw) {
    std::lock_guard<std::mutex> lock(g_seed_mu);
    const auto key = std::make_tuple(pr.n, pr.center, pr.order, pr.lock_radius);
    auto it = g_seed_sets.find(key);
    if (it == g_seed_sets.end()) {
        int64_t r = 0;
        auto seeds = sym_seeds(pr, &r);
        it = g_seed_sets.emplace(key, std::make_pair(std::move(seeds), r)).first;
    }
    if (raw) *raw = it->second.second;
    return it->second.first;
}

// --------------------------------------------------------------- case index

namespace {

// Cases of the subtree at level `stop`, counting stops once it exceeds `cap`
// (then the result is only known to be > cap).
template <typename M>
uint64_t count_below(const SymProblem &pr, uint8_t *cycle, int i, int stop, uint64_t cap) {
    if (i == stop) return 1;
    uint64_t total = 0;
    for (M g = good_t<M>(pr, cycle, i); g; g &= g - 1) {
        const int j = ctz_mask(g);
        const uint8_t nx = cycle[j];
        cycle[i] = nx;
        cycle[j] = static_cast<uint8_t>(i);
        total += count_below<M>(pr, cycle, i + 1, stop, cap - total);
        cycle[j] = nx;
        if (total > cap) break;
    }
    return total;
}

std::mutex g_index_mu;
std::map<std::tuple<int, int, int, int, int>, std::vector<uint64_t>> g_index;   // (n, center, order, lock, expanded)

const std::vector<SymSeed> &cached_seeds(const SymProblem &pr) { return sym_seed_list(pr, nullptr); }

// Extend counts (cases per seed at `expanded`; counts[k] belongs to seed
// first + k of the slice [first, end)) until they cover case index `upto` of
// the slice or every seed of it is counted.  Seeds are counted in parallel
// batches; only the prefix of the stream that is needed is ever enumerated,
// and a seed is counted only up to what is left to cover: one whose count
// passes that bound holds index `upto`, so it is not recorded (the stream
// enumerates inside it instead; its count may be far beyond what any window
// needs, e.g. 9x9 seeds at deep levels).
void extend_counts(const SymProblem &pr, const std::vector<SymSeed> &seeds, size_t first, size_t end, int expanded,
                   uint64_t upto, std::vector<uint64_t> &counts) {
    const int c2 = pr.center * pr.center;
    uint64_t covered = 0;
    for (uint64_t c : counts) covered += c;
    unsigned hw = std::thread::hardware_concurrency();
    const size_t threads = std::max<size_t>(1, std::min<size_t>(hw == 0 ? 1 : hw, 64));
    while (first + counts.size() < end && covered <= upto) {
        const size_t at = first + counts.size();
        const size_t batch = std::min(end - at, threads * 4);
        const uint64_t cap = upto == ~0ull ? ~0ull : upto - covered;   // a seed with more holds `upto`
        std::vector<uint64_t> part(batch, 0);
        std::atomic<size_t> next{0};
        std::vector<std::thread> pool;
        for (size_t t = 0; t < std::min(threads, batch); ++t) {
            pool.emplace_back([&]() {
                for (;;) {
                    const size_t k = next.fetch_add(1);
                    if (k >= batch) break;
                    uint8_t cyc[kMaxCells];
                    std::memcpy(cyc, seeds[at + k].cycle, kMaxCells);
                    part[k] = pr.n2 <= 64 ? count_below<uint64_t>(pr, cyc, c2, expanded, cap)
                                          : count_below<u128>(pr, cyc, c2, expanded, cap);
                }
            });
        }
        for (auto &th : pool) th.join();
        size_t kept = 0;
        for (uint64_t c : part) {
            if (c > cap) break;   // counted only past the bound: not exact
            counts.push_back(c);
            covered += c;
            ++kept;
        }
        if (kept < batch) break;
    }
}

// The seed slice [first, end) clipped to the seed list.
void clip_slice(size_t nseeds, uint64_t seed_start, uint64_t seed_count, size_t &first, size_t &end) {
    first = static_cast<size_t>(std::min<uint64_t>(seed_start, nseeds));
    end = seed_count >= nseeds - first ? nseeds : first + static_cast<size_t>(seed_count);
}

}  // namespace

const std::vector<uint64_t> &sym_seed_case_counts(const SymProblem &pr, int expanded) {
    std::lock_guard<std::mutex> lock(g_index_mu);
    const auto &seeds = cached_seeds(pr);
    auto &counts = g_index[std::make_tuple(pr.n, pr.center, pr.order, pr.lock_radius, expanded)];
    extend_counts(pr, seeds, 0, seeds.size(), expanded, ~0ull, counts);
    return counts;
}

std::vector<uint64_t> sym_slice_case_counts(const SymProblem &pr, int expanded, uint64_t seed_start,
                                            uint64_t seed_count) {
    const std::vector<SymSeed> *seeds;
    {
        std::lock_guard<std::mutex> lock(g_index_mu);
        seeds = &cached_seeds(pr);
    }
    size_t first, end;
    clip_slice(seeds->size(), seed_start, seed_count, first, end);
    if (first == 0 && end == seeds->size()) return sym_seed_case_counts(pr, expanded);
    std::vector<uint64_t> counts;
    extend_counts(pr, *seeds, first, end, expanded, ~0ull, counts);
    return counts;
}

namespace {

template <typename M>
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
            const int cb = pr.case_bytes();
            std::memcpy(cycles + written * cb, cycle, cb);
            mults[written] = mult;
            ++written;
            return --want == 0;
        }
        for (M g = good_t<M>(pr, cycle, i); g; g &= g - 1) {
            const int j = ctz_mask(g);
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

template <typename M>
int64_t stream_window(const SymProblem &pr, const std::vector<SymSeed> &seeds, size_t s, size_t end, int expanded,
                      uint64_t skip, uint64_t max_cases, uint8_t *cycles, uint64_t *mults) {
    Stream<M> st{pr, expanded, skip, max_cases, 0, cycles, mults};
    const int c2 = pr.center * pr.center;
    for (; s < end; ++s) {
        uint8_t cyc[kMaxCells];
        std::memcpy(cyc, seeds[s].cycle, kMaxCells);
        st.mult = seeds[s].mult;
        if (st.visit(cyc, c2)) break;
    }
    return st.written;
}

}  // namespace

int64_t sym_cases(const SymProblem &pr, int expanded, uint64_t start, uint64_t max_cases, uint8_t *cycles,
                  uint64_t *mults, uint64_t seed_start, uint64_t seed_count) {
    if (max_cases == 0) return 0;
    std::vector<uint64_t> counts;
    const std::vector<SymSeed> *seeds;
    size_t first, end;
    {
        std::lock_guard<std::mutex> lock(g_index_mu);
        seeds = &cached_seeds(pr);
        clip_slice(seeds->size(), seed_start, seed_count, first, end);
        if (first == 0 && end == seeds->size()) {
            auto &cached = g_index[std::make_tuple(pr.n, pr.center, pr.order, pr.lock_radius, expanded)];
            extend_counts(pr, *seeds, 0, end, expanded, start, cached);
            counts = cached;
        }
    }
    if (first != 0 || end != seeds->size()) extend_counts(pr, *seeds, first, end, expanded, start, counts);
    size_t s = 0;
    uint64_t skip = start;
    while (s < counts.size() && counts[s] <= skip) skip -= counts[s++];
    if (first + s >= end) return 0;
    return pr.n2 <= 64 ? stream_window<uint64_t>(pr, *seeds, first + s, end, expanded, skip, max_cases, cycles, mults)
                       : stream_window<u128>(pr, *seeds, first + s, end, expanded, skip, max_cases, cycles, mults);
}

}  // namespace mf
