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
#include <vector>

#include "cli_common.hpp"

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
