// gpu_folding <grid_size> <partition_depth> -- drop-in for gpu_folding.cu:209-311.
//
// Same arguments, checks and two stdout lines as the reference:
//   Number of partial states at depth D: S
//   Total number of ways: T
// Extension (the reference is square-only): MF_DIMS="p,q[,r,s]" counts a
// rectangular (or up to 4-D) map instead of grid x grid.
#include <vector>

#include "cli_common.hpp"

int main(int argc, char **argv) {
    if (argc < 3) {
        std::printf("Usage: ./gpu_folding grid_size partition_depth\n");
        std::printf("Example: ./gpu_folding 20 8\n");
        return 1;
    }
    const int grid = std::atoi(argv[1]);
    const int depth = std::atoi(argv[2]);
    std::vector<int32_t> dims = {grid, grid};
    if (const char *text = std::getenv("MF_DIMS")) {
        dims.clear();
        std::string s(text);
        size_t pos = 0;
        while (pos <= s.size()) {
            const size_t comma = s.find_first_of(",x", pos);
            const std::string tok = s.substr(pos, comma == std::string::npos ? std::string::npos : comma - pos);
            if (!tok.empty()) dims.push_back(mfcli::parse_int(tok.c_str(), "MF_DIMS"));
            if (comma == std::string::npos) break;
            pos = comma + 1;
        }
    }
    long n = 1;
    for (int d : dims) n *= (d > 0 ? d : 0);
    if (n > 64) {
        std::printf("Grid too large, max 8 (64 stamps)\n");
        return 1;
    }
    if (n < 2 || depth < 1 || depth >= n) {
        std::fprintf(stderr, "partition_depth must be between 1 and %ld\n", n - 1);
        return 1;
    }
    int rc = mf_init(0);
    if (rc != MF_OK) mfcli::die(rc, "mf_init");
    int64_t states = 0;
    rc = mf_lunnon_frontier(dims.data(), static_cast<int32_t>(dims.size()), depth, nullptr, 0, &states);
    if (rc != MF_OK) mfcli::die(rc, "mf_lunnon_frontier");
    std::printf("Number of partial states at depth %d: %lld\n", depth, static_cast<long long>(states));
    std::fflush(stdout);
    mf_solve_options opt = mfcli::options_from_env();
    mf_result res;
    rc = mf_gpu_folding(dims.data(), static_cast<int32_t>(dims.size()), depth, &opt, &res, &states);
    if (rc != MF_OK) mfcli::die(rc, "mf_gpu_folding");
    const unsigned __int128 total = (static_cast<unsigned __int128>(res.count_hi) << 64) | res.count_lo;
    std::printf("Total number of ways: %s\n", mfcli::to_decimal(total).c_str());
    std::fprintf(stderr, "states=%lld inserts=%llu rounds=%u kernel_ms=%.3f total_ms=%.3f\n",
                 static_cast<long long>(states), static_cast<unsigned long long>(res.nodes), res.rounds,
                 res.kernel_ms, res.total_ms);
    mf_shutdown();
    return 0;
}
