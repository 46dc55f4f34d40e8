// sym_fold n [threads] [max_cases] [start_case] -- drop-in for sym_fold.cpp:642-774.
// Built a second time with MF_CLI_MAX_N=9 as sym_fold9, the drop-in for
// experiments/9x9/sym_fold9.cpp:984-1198.
//
// stdout carries exactly one decimal line (the answer); diagnostics go to
// stderr.  Environment as in the reference: MF_EXPANDED_N (n >= 7, default 24,
// for sym_fold9 max(24, c*c)), MF_STREAM_CASES, MF_SETUP_ONLY.  `threads` is
// accepted and validated; the counting runs on the GPU.  The case-index window
// [start_case, start_case+max_cases) selects the same cases as the reference
// in both the streaming and the materialised mode (they enumerate in the same
// order).  MF_EXPANDED_N = n*n is rejected: cases would be complete foldings
// enumerated on the host (the 8x8 reference would also read past its 64-entry
// edge tables there).
//
// Materialised mode (no MF_STREAM_CASES) also honours MF_COST_DEPTH and
// MF_SORT_CASES (sym_fold.cpp:267-314): cases are scored with the lookahead
// proxy and, if asked, reordered before the [start, start+max) window is
// taken -- so, as in the reference, sorting changes which cases a window
// holds.  The order is std::sort with the reference's comparator (score, ties
// by multiplier ascending) over the generation order, the same algorithm and
// comparisons as the reference's host code.
//
// sym_fold9 only (sym_fold9.cpp:160-229,968-980,1040-1172): MF_CENTER_N
// (center side, default 5 for n = 9), MF_ORDER / MF_LOCK_CENTER_N (ring-wise
// cell numberings of odd squares), MF_CENTER_HIST, MF_SEED_START /
// MF_SEED_COUNT (a seed slice; case 0 is the first case of seed
// MF_SEED_START), MF_SEED_COUNTS / MF_SEED_COUNTS_SUMMARY with
// MF_SETUP_ONLY, and MF_PROFILE_WORK (per-case
// nodes / leaves / weighted leaves of the window, counted on the GPU; a node
// is one get_good call as in recurse_count_work, sym_fold9.cpp:342-357).
// MF_CENTER_STREAM_DEDUP is accepted: it changes how the reference holds the
// center cycles while deduplicating, not the seeds.
#include <utility>
#include <vector>

#include "../../include/mapfold_cost.h"
#include "cli_common.hpp"

#ifndef MF_CLI_MAX_N
#define MF_CLI_MAX_N 8
#endif

namespace {

constexpr bool kNine = MF_CLI_MAX_N >= 9;

int case_bytes(int n) { return n * n <= 64 ? 64 : 128; }

struct Slice {
    uint64_t start = 0;
    uint64_t count = MF_ALL_SEEDS;
};

// Cases of the materialised list (generation order) of the seed slice: the
// seeds themselves for n = 6, their expansion to `level` for n >= 7.
int all_cases(int n, int level, const Slice &sl, std::vector<uint8_t> &cyc, std::vector<uint64_t> &mul) {
    int64_t seeds = 0;
    int rc = mf_sym_seed_case_counts(n, level, sl.start, sl.count, nullptr, 0, &seeds);
    if (rc != MF_OK) return rc;
    std::vector<uint64_t> per(static_cast<size_t>(seeds));
    if ((rc = mf_sym_seed_case_counts(n, level, sl.start, sl.count, per.data(), seeds, &seeds)) != MF_OK) return rc;
    uint64_t total = 0;
    for (uint64_t c : per) total += c;
    cyc.assign(static_cast<size_t>(total) * case_bytes(n), 0);
    mul.assign(static_cast<size_t>(total), 0);
    int64_t got = 0;
    return mf_sym_cases_slice(n, level, sl.start, sl.count, 0, total, cyc.data(), mul.data(), &got);
}

struct Sorted {
    std::vector<uint8_t> cyc;
    std::vector<uint64_t> mul;
    std::vector<size_t> order;   // case list order after scoring / sorting
};

// score_and_sort_cases (sym_fold.cpp:267-314): false when MF_COST_DEPTH is unset.
bool score_and_sort(int n, int level, const Slice &sl, Sorted &st) {
    const char *depth_text = std::getenv("MF_COST_DEPTH");
    const int depth = depth_text == nullptr ? 0 : std::atoi(depth_text);
    if (depth <= 0) return false;
    int rc = all_cases(n, level, sl, st.cyc, st.mul);
    if (rc != MF_OK) mfcli::die(rc, "case generation");
    const size_t k = st.mul.size();
    std::vector<uint64_t> score(k);
    const auto t0 = std::chrono::steady_clock::now();
    if ((rc = mf_sym_estimate_work(n, level, st.cyc.data(), static_cast<int64_t>(k), depth, score.data())) != MF_OK)
        mfcli::die(rc, "mf_sym_estimate_work");
    uint64_t lo = UINT64_MAX, hi = 0;
    unsigned __int128 tot = 0;
    for (uint64_t v : score) {
        lo = std::min(lo, v);
        hi = std::max(hi, v);
        tot += v;
    }
    std::cerr << "cost_depth=" << depth << " min_score=" << lo << " max_score=" << hi
              << " total_score=" << mfcli::to_decimal(tot) << " score_s=" << mfcli::seconds_since(t0) << "\n";
    st.order.resize(k);
    for (size_t i = 0; i < k; ++i) st.order[i] = i;
    const char *sort_text = std::getenv("MF_SORT_CASES");
    if (sort_text != nullptr && sort_text[0] != '\0' && std::string(sort_text) != "0") {
        const bool asc = std::string(sort_text) == "asc";
        std::vector<std::pair<uint64_t, size_t>> sc(k);
        for (size_t i = 0; i < k; ++i) sc[i] = {score[i], i};
        const std::vector<uint64_t> &mul = st.mul;
        std::sort(sc.begin(), sc.end(), [&](const auto &x, const auto &y) {
            if (x.first == y.first) return mul[x.second] < mul[y.second];
            return asc ? x.first < y.first : x.first > y.first;
        });
        for (size_t i = 0; i < k; ++i) st.order[i] = sc[i].second;
        std::cerr << "sorted_cases=" << (asc ? "asc" : "desc") << "\n";
    }
    return true;
}

// The window [start, start + max) of the scored (possibly sorted) list, on the GPU.
unsigned __int128 solve_window(int n, int level, const Sorted &st, uint64_t start, uint64_t max_cases,
                               const mf_solve_options &opt) {
    const size_t k = st.order.size();
    const size_t cb = static_cast<size_t>(case_bytes(n));
    const size_t first = static_cast<size_t>(std::min<uint64_t>(start, k));
    const size_t last = max_cases == 0 ? k : static_cast<size_t>(std::min<uint64_t>(k, start + max_cases));
    std::vector<uint8_t> wc;
    std::vector<uint64_t> wm;
    for (size_t i = first; i < last; ++i) {
        wc.insert(wc.end(), st.cyc.begin() + st.order[i] * cb, st.cyc.begin() + (st.order[i] + 1) * cb);
        wm.push_back(st.mul[st.order[i]]);
    }
    mf_result res;
    std::memset(&res, 0, sizeof(res));
    if (!wm.empty()) {
        const int rc = mf_sym_solve(n, level, wc.data(), wm.data(), static_cast<int64_t>(wm.size()), &opt, nullptr,
                                    nullptr, &res);
        if (rc != MF_OK) mfcli::die(rc, "mf_sym_solve");
    }
    const unsigned __int128 cnt = (static_cast<unsigned __int128>(res.count_hi) << 64) | res.count_lo;
    return cnt * static_cast<unsigned>(n * n);
}

// print_count_summary (sym_fold9.cpp:393-424): min, p50/p90/p99 (value at
// index floor(p * (k - 1)) of the sorted list), max, mean (floor), total.
void count_summary(const char *label, std::vector<unsigned __int128> v) {
    if (v.empty()) {
        std::cerr << label << "_count=0\n";
        return;
    }
    std::sort(v.begin(), v.end());
    unsigned __int128 total = 0;
    for (auto x : v) total += x;
    auto pct = [&](double p) { return v[static_cast<size_t>(p * static_cast<double>(v.size() - 1))]; };
    using mfcli::to_decimal;
    std::cerr << label << "_count=" << v.size() << " min=" << to_decimal(v.front()) << " p50=" << to_decimal(pct(0.50))
              << " p90=" << to_decimal(pct(0.90)) << " p99=" << to_decimal(pct(0.99))
              << " max=" << to_decimal(v.back()) << " mean=" << to_decimal(total / v.size())
              << " total=" << to_decimal(total) << "\n";
}

// Center side as choose_center_size (sym_fold9.cpp:968-980), handed to the
// library with mf_sym_set_numbering.
int choose_center(int n) {
    int c = n >= 9 ? 5 : ((n % 2 == 0) ? 4 : 3);
    if (const char *text = std::getenv("MF_CENTER_N")) c = std::atoi(text);
    if (c < 3 || c > n || c * c > 25 || (c % 2) != (n % 2)) {
        std::cerr << "MF_CENTER_N must be between 3 and " << n
                  << ", have the same parity as n, and fit MAX_CENTER=25\n";
        std::exit(2);
    }
    return c;
}

// MF_PROFILE_WORK (profile_cases_work, sym_fold9.cpp:785-848) on the GPU.
unsigned __int128 profile_window(int n, int level, const Slice &sl, uint64_t start, uint64_t max_cases,
                                 const mf_solve_options &opt) {
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<uint8_t> cyc(static_cast<size_t>(max_cases) * case_bytes(n));
    std::vector<uint64_t> mul(static_cast<size_t>(max_cases));
    int64_t got = 0;
    int rc = mf_sym_cases_slice(n, level, sl.start, sl.count, start, max_cases, cyc.data(), mul.data(), &got);
    if (rc != MF_OK) mfcli::die(rc, "mf_sym_cases_slice");
    std::vector<uint64_t> pc(static_cast<size_t>(got)), pn(static_cast<size_t>(got));
    if (got > 0) {
        mf_result res;
        rc = mf_sym_solve(n, level, cyc.data(), mul.data(), got, &opt, pc.data(), pn.data(), &res);
        if (rc != MF_OK) mfcli::die(rc, "mf_sym_solve");
    }
    std::vector<unsigned __int128> nodes, leaves, weighted;
    unsigned __int128 sum = 0;
    const bool verbose = std::getenv("MF_PROFILE_VERBOSE") != nullptr;
    for (int64_t k = 0; k < got; ++k) {
        const unsigned __int128 w = static_cast<unsigned __int128>(mul[k]) * pc[k];
        nodes.push_back(pn[k]);
        leaves.push_back(pc[k]);
        weighted.push_back(w);
        sum += w;
        if (verbose)
            std::cerr << "profile_case index=" << (start + static_cast<uint64_t>(k)) << " multiplier=" << mul[k]
                      << " nodes=" << pn[k] << " leaves=" << pc[k] << " weighted_leaves=" << mfcli::to_decimal(w)
                      << "\n";
    }
    count_summary("profile_nodes", nodes);
    count_summary("profile_leaves", leaves);
    count_summary("profile_weighted_leaves", weighted);
    std::cerr << "profile_start_case=" << start << " profile_max_cases=" << max_cases << " profile_seen_cases=" << got
              << " profile_total_s=" << mfcli::seconds_since(t0) << "\n";
    return sum;
}

}  // namespace

static void usage(const char *exe) {
    std::cerr << "Usage: " << exe << " n [threads] [max_cases] [start_case]\n"
              << "Example: " << exe << " 6 12\n";
}

int main(int argc, char **argv) {
    if (argc < 2 || argc > 5) {
        usage(argv[0]);
        return 1;
    }
    const int n = mfcli::parse_int(argv[1], "n");
    if (n < 1 || n > MF_CLI_MAX_N) {
        std::cerr << "n must be between 1 and " << MF_CLI_MAX_N << "\n";
        return 1;
    }
    int threads = 1;
    if (argc >= 3) threads = mfcli::parse_int(argv[2], "threads");
    if (threads < 1) {
        std::cerr << "threads must be positive\n";
        return 1;
    }
    uint64_t max_cases = 0, start_case = 0;
    if (argc >= 4) max_cases = mfcli::parse_size(argv[3], "max_cases");
    if (argc >= 5) start_case = mfcli::parse_size(argv[4], "start_case");

    const auto t0 = std::chrono::steady_clock::now();
    if (n <= 1) {
        std::cout << "1\n";
        return 0;
    }
    if (n == 2) {
        std::cout << "8\n";
        return 0;
    }
    // sym_fold9's seed center and cell numbering come from its environment
    // (sym_fold9.cpp:160-229,968-980); the 8x8 tool knows neither and keeps
    // the library's defaults (the spiral, c = 4 or 3).
    if (kNine) {
        const int32_t c = n >= 6 ? choose_center(n) : 0;
        int32_t order = MF_ORDER_SPIRAL, lock = 3;
        if (const char *text = std::getenv("MF_ORDER")) {
            const std::string o = text;
            order = o == "reverse" ? MF_ORDER_REVERSE
                    : o == "opposite" ? MF_ORDER_OPPOSITE
                    : o == "axes_first" ? MF_ORDER_AXES_FIRST
                    : o == "corners_first" ? MF_ORDER_CORNERS_FIRST : MF_ORDER_SPIRAL;
            if (const char *lt = std::getenv("MF_LOCK_CENTER_N")) lock = std::max(0, std::atoi(lt));
        }
        const int rc = mf_sym_set_numbering(c, order, lock);
        if (rc != MF_OK) mfcli::die(rc, "mf_sym_set_numbering");
    }
    const int n2 = n * n;
    int expanded = 0;
    int center = 0;
    const bool stream = std::getenv("MF_STREAM_CASES") != nullptr;
    bool profile = false;
    Slice sl;
    Sorted st;
    bool sorted = false;
    if (n >= 6) {
        center = kNine ? choose_center(n) : ((n % 2 == 0) ? 4 : 3);
        const int c2 = center * center;
        int64_t seeds = 0, raw = 0;
        int rc = mf_sym_seeds(n, nullptr, nullptr, 0, &seeds, &raw);
        if (rc != MF_OK) mfcli::die(rc, "mf_sym_seeds");
        if (kNine) {
            std::cerr << "center_size=" << center << " center_cases=" << raw
                      << (std::getenv("MF_CENTER_STREAM_DEDUP") ? " center_stream_dedup=1" : "") << "\n";
        } else {
            std::cerr << "center_cases=" << raw << "\n";
        }
        std::cerr << "reduced_cases=" << seeds << "\n";
        if (kNine && std::getenv("MF_CENTER_HIST") != nullptr) {
            std::vector<uint8_t> cyc(static_cast<size_t>(seeds) * case_bytes(n));
            std::vector<uint64_t> mul(static_cast<size_t>(seeds));
            if ((rc = mf_sym_seeds(n, cyc.data(), mul.data(), seeds, &seeds, &raw)) != MF_OK)
                mfcli::die(rc, "mf_sym_seeds");
            uint64_t hist[17] = {};
            for (uint64_t m : mul)
                if (m < 17) ++hist[m];
            std::cerr << "center_multiplier_hist";
            for (int i = 0; i < 17; ++i)
                if (hist[i] != 0) std::cerr << " " << i << ":" << hist[i];
            std::cerr << "\n";
        }
        uint64_t slice_seeds = static_cast<uint64_t>(seeds);
        if (kNine) {
            if (const char *text = std::getenv("MF_SEED_START")) {
                sl.start = mfcli::parse_size(text, "MF_SEED_START");
                if (const char *ct = std::getenv("MF_SEED_COUNT")) sl.count = mfcli::parse_size(ct, "MF_SEED_COUNT");
                const uint64_t first = std::min<uint64_t>(sl.start, slice_seeds);
                slice_seeds = std::min<uint64_t>(slice_seeds - first, sl.count);
                std::cerr << "seed_slice_start=" << sl.start << " seed_slice_count=" << slice_seeds << "\n";
            }
        }
        expanded = c2;
        uint64_t cases = slice_seeds;
        if (n >= 7) {
            expanded = kNine ? std::max(24, c2) : 24;
            if (const char *text = std::getenv("MF_EXPANDED_N")) expanded = std::atoi(text);
            if (expanded < c2 || expanded > n2 - 1) {
                std::cerr << "MF_EXPANDED_N must be between " << c2 << " and " << (n2 - 1) << "\n";
                return 1;
            }
            if (stream) {
                std::cerr << "stream_seed_cases=" << slice_seeds << " stream_from=" << c2 << " stream_to=" << expanded
                          << "\n";
            } else {
                int64_t k = 0;
                if ((rc = mf_sym_seed_case_counts(n, expanded, sl.start, sl.count, nullptr, 0, &k)) != MF_OK)
                    mfcli::die(rc, "mf_sym_seed_case_counts");
                std::vector<uint64_t> per(static_cast<size_t>(k));
                if ((rc = mf_sym_seed_case_counts(n, expanded, sl.start, sl.count, per.data(), k, &k)) != MF_OK)
                    mfcli::die(rc, "mf_sym_seed_case_counts");
                cases = 0;
                for (uint64_t c : per) cases += c;
                std::cerr << "expanded_cases=" << cases << "\n";
            }
        }
        if (!stream) sorted = score_and_sort(n, expanded, sl, st);
        if (kNine && std::getenv("MF_PROFILE_WORK") != nullptr) {
            if (!stream || expanded <= c2) {
                std::cerr << "MF_PROFILE_WORK requires MF_STREAM_CASES=1 and MF_EXPANDED_N > center depth\n";
                return 2;
            }
            if (max_cases == 0) {
                std::cerr << "MF_PROFILE_WORK requires max_cases > 0\n";
                return 2;
            }
            profile = true;
        }
        if (!profile && std::getenv("MF_SETUP_ONLY") != nullptr) {
            if (kNine && stream && expanded > c2) {
                int64_t k = 0;
                if ((rc = mf_sym_seed_case_counts(n, expanded, sl.start, sl.count, nullptr, 0, &k)) != MF_OK)
                    mfcli::die(rc, "mf_sym_seed_case_counts");
                std::vector<uint64_t> per(static_cast<size_t>(k));
                if ((rc = mf_sym_seed_case_counts(n, expanded, sl.start, sl.count, per.data(), k, &k)) != MF_OK)
                    mfcli::die(rc, "mf_sym_seed_case_counts");
                const bool each = std::getenv("MF_SEED_COUNTS") != nullptr;
                const bool summary = std::getenv("MF_SEED_COUNTS_SUMMARY") != nullptr;
                std::vector<uint64_t> mul;
                if (each) {
                    int64_t ns = 0;
                    std::vector<uint8_t> cyc(static_cast<size_t>(seeds) * case_bytes(n));
                    mul.resize(static_cast<size_t>(seeds));
                    if ((rc = mf_sym_seeds(n, cyc.data(), mul.data(), seeds, &ns, &raw)) != MF_OK)
                        mfcli::die(rc, "mf_sym_seeds");
                }
                const uint64_t base = std::min<uint64_t>(sl.start, static_cast<uint64_t>(seeds));
                unsigned __int128 generated = 0;
                std::vector<unsigned __int128> vals;
                for (size_t s = 0; s < per.size(); ++s) {
                    generated += per[s];
                    if (summary) vals.push_back(per[s]);
                    if (each)
                        std::cerr << "seed_count index=" << (base + s) << " multiplier=" << mul[base + s]
                                  << " expanded_cases=" << per[s] << "\n";
                }
                if (summary) count_summary("seed_expanded", vals);
                std::cout << mfcli::to_decimal(generated) << "\n";
            } else {
                std::cout << cases << "\n";
            }
            std::cerr << "n=" << n << " threads=" << threads;
            if (kNine) std::cerr << " center_n=" << center;
            std::cerr << " max_cases=" << max_cases << " start_case=" << start_case
                      << " setup_only=1 total_s=" << mfcli::seconds_since(t0) << "\n";
            return 0;
        }
    }
    int rc = mf_init(0);
    if (rc != MF_OK) mfcli::die(rc, "mf_init");
    mf_solve_options opt = mfcli::options_from_env();
    if (profile) {
        const unsigned __int128 w = profile_window(n, expanded, sl, start_case, max_cases, opt);
        std::cout << mfcli::to_decimal(w * static_cast<unsigned>(n2)) << "\n";
        std::cerr << "n=" << n << " threads=" << threads << " center_n=" << center << " max_cases=" << max_cases
                  << " start_case=" << start_case << " profile_work=1 total_s=" << mfcli::seconds_since(t0) << "\n";
        mf_shutdown();
        return 0;
    }
    if (sorted) {
        std::cout << mfcli::to_decimal(solve_window(n, expanded, st, start_case, max_cases, opt)) << "\n";
        std::cerr << "n=" << n << " threads=" << threads << " max_cases=" << max_cases << " start_case=" << start_case
                  << " total_s=" << mfcli::seconds_since(t0) << "\n";
        mf_shutdown();
        return 0;
    }
    uint64_t answer[2] = {0, 0};
    mf_result res;
    rc = mf_sym_fold_slice(n, expanded, sl.start, sl.count, start_case, max_cases, &opt, answer, &res);
    if (rc != MF_OK) mfcli::die(rc, "mf_sym_fold_slice");
    const unsigned __int128 a = (static_cast<unsigned __int128>(answer[1]) << 64) | answer[0];
    if (kNine && n <= 5) {
        // recurse_count's own total before the 2 n^2 orientations / starts.
        std::cerr << "raw_count=" << mfcli::to_decimal(a / static_cast<unsigned>(2 * n2)) << "\n";
    }
    std::cout << mfcli::to_decimal(a) << "\n";
    std::cerr << "n=" << n << " threads=" << threads << " max_cases=" << max_cases << " start_case=" << start_case
              << " nodes=" << res.nodes << " rounds=" << res.rounds << " kernel_ms=" << res.kernel_ms
              << " total_s=" << mfcli::seconds_since(t0) << "\n";
    mf_shutdown();
    return 0;
}
