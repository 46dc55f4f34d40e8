// sym_fold n [threads] [max_cases] [start_case] -- drop-in for sym_fold.cpp:642-774.
//
// stdout carries exactly one decimal line (the answer); diagnostics go to
// stderr.  Environment as in the reference: MF_EXPANDED_N (n >= 7, default 24),
// MF_STREAM_CASES, MF_SETUP_ONLY.  `threads` is accepted and validated; the
// counting runs on the GPU.  The case-index window [start_case,
// start_case+max_cases) selects the same cases as the reference in both the
// streaming and the materialised mode (they enumerate in the same order).
// MF_EXPANDED_N = n*n is rejected: the reference would read past its 64-entry
// edge tables there.
//
// Materialised mode (no MF_STREAM_CASES) also honours MF_COST_DEPTH and
// MF_SORT_CASES (sym_fold.cpp:267-314): cases are scored with the lookahead
// proxy and, if asked, reordered before the [start, start+max) window is
// taken -- so, as in the reference, sorting changes which cases a window
// holds.  The order is std::sort with the reference's comparator (score, ties
// by multiplier ascending) over the generation order, the same algorithm and
// comparisons as the reference's host code.
#include <utility>
#include <vector>

#include "../../include/mapfold_cost.h"
#include "cli_common.hpp"

namespace {

// Cases of the materialised list (generation order) for n >= 6.
int all_cases(int n, int level, std::vector<uint8_t> &cyc, std::vector<uint64_t> &mul) {
    int rc;
    if (n == 6) {
        int64_t seeds = 0, raw = 0;
        if ((rc = mf_sym_seeds(n, nullptr, nullptr, 0, &seeds, &raw)) != MF_OK) return rc;
        cyc.assign(static_cast<size_t>(seeds) * 64, 0);
        mul.assign(static_cast<size_t>(seeds), 0);
        return mf_sym_seeds(n, cyc.data(), mul.data(), seeds, &seeds, &raw);
    }
    uint64_t total = 0;
    if ((rc = mf_sym_case_total(n, level, &total)) != MF_OK) return rc;
    cyc.assign(static_cast<size_t>(total) * 64, 0);
    mul.assign(static_cast<size_t>(total), 0);
    int64_t got = 0;
    rc = mf_sym_cases(n, level, 0, total, cyc.data(), mul.data(), &got);
    return rc;
}

struct Sorted {
    std::vector<uint8_t> cyc;
    std::vector<uint64_t> mul;
    std::vector<size_t> order;   // case list order after scoring / sorting
};

// score_and_sort_cases (sym_fold.cpp:267-314): false when MF_COST_DEPTH is unset.
bool score_and_sort(int n, int level, Sorted &st) {
    const char *depth_text = std::getenv("MF_COST_DEPTH");
    const int depth = depth_text == nullptr ? 0 : std::atoi(depth_text);
    if (depth <= 0) return false;
    int rc = all_cases(n, level, st.cyc, st.mul);
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
    const size_t first = static_cast<size_t>(std::min<uint64_t>(start, k));
    const size_t last = max_cases == 0 ? k : static_cast<size_t>(std::min<uint64_t>(k, start + max_cases));
    std::vector<uint8_t> wc;
    std::vector<uint64_t> wm;
    for (size_t i = first; i < last; ++i) {
        wc.insert(wc.end(), st.cyc.begin() + st.order[i] * 64, st.cyc.begin() + st.order[i] * 64 + 64);
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
    if (n < 1 || n > 8) {
        std::cerr << "n must be between 1 and 8\n";
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
    const int n2 = n * n;
    int expanded = 0;
    const bool stream = std::getenv("MF_STREAM_CASES") != nullptr;
    Sorted st;
    bool sorted = false;
    if (n >= 6) {
        const int c2 = (n % 2 == 0) ? 16 : 9;
        int64_t seeds = 0, raw = 0;
        int rc = mf_sym_seeds(n, nullptr, nullptr, 0, &seeds, &raw);
        if (rc != MF_OK) mfcli::die(rc, "mf_sym_seeds");
        std::cerr << "center_cases=" << raw << "\n";
        std::cerr << "reduced_cases=" << seeds << "\n";
        expanded = c2;
        uint64_t cases = static_cast<uint64_t>(seeds);
        if (n >= 7) {
            expanded = 24;
            if (const char *text = std::getenv("MF_EXPANDED_N")) expanded = std::atoi(text);
            if (expanded < c2 || expanded > n2 - 1) {
                std::cerr << "MF_EXPANDED_N must be between " << c2 << " and " << (n2 - 1) << "\n";
                return 1;
            }
            if (stream) {
                std::cerr << "stream_seed_cases=" << seeds << " stream_from=" << c2 << " stream_to=" << expanded
                          << "\n";
            } else {
                rc = mf_sym_case_total(n, expanded, &cases);
                if (rc != MF_OK) mfcli::die(rc, "mf_sym_case_total");
                std::cerr << "expanded_cases=" << cases << "\n";
            }
        }
        if (!stream) sorted = score_and_sort(n, expanded, st);
        if (std::getenv("MF_SETUP_ONLY") != nullptr) {
            std::cout << cases << "\n";
            std::cerr << "n=" << n << " threads=" << threads << " max_cases=" << max_cases
                      << " start_case=" << start_case << " setup_only=1 total_s=" << mfcli::seconds_since(t0) << "\n";
            return 0;
        }
    }
    int rc = mf_init(0);
    if (rc != MF_OK) mfcli::die(rc, "mf_init");
    mf_solve_options opt = mfcli::options_from_env();
    if (sorted) {
        std::cout << mfcli::to_decimal(solve_window(n, expanded, st, start_case, max_cases, opt)) << "\n";
        std::cerr << "n=" << n << " threads=" << threads << " max_cases=" << max_cases << " start_case=" << start_case
                  << " total_s=" << mfcli::seconds_since(t0) << "\n";
        mf_shutdown();
        return 0;
    }
    uint64_t answer[2] = {0, 0};
    mf_result res;
    rc = mf_sym_fold(n, expanded, start_case, max_cases, &opt, answer, &res);
    if (rc != MF_OK) mfcli::die(rc, "mf_sym_fold");
    const unsigned __int128 a = (static_cast<unsigned __int128>(answer[1]) << 64) | answer[0];
    std::cout << mfcli::to_decimal(a) << "\n";
    std::cerr << "n=" << n << " threads=" << threads << " max_cases=" << max_cases << " start_case=" << start_case
              << " nodes=" << res.nodes << " rounds=" << res.rounds << " kernel_ms=" << res.kernel_ms
              << " total_s=" << mfcli::seconds_since(t0) << "\n";
    mf_shutdown();
    return 0;
}
