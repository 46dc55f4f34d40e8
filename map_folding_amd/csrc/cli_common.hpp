// cli_common.hpp -- shared helpers of the drop-in command-line tools.
#pragma once

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "../../include/mapfold.h"

namespace mfcli {

inline std::string to_decimal(unsigned __int128 v) {
    if (v == 0) return "0";
    std::string s;
    while (v > 0) {
        s.push_back(static_cast<char>('0' + static_cast<int>(v % 10)));
        v /= 10;
    }
    std::reverse(s.begin(), s.end());
    return s;
}

// strtol / strtoull with the reference's error convention: exit 2 on a
// malformed integer (sym_fold.cpp:622-640, gpu_folding_v2.cu:544-562).
inline int parse_int(const char *text, const char *name) {
    char *end = nullptr;
    const long v = std::strtol(text, &end, 10);
    if (*text == '\0' || *end != '\0') {
        std::cerr << "Invalid " << name << ": " << text << "\n";
        std::exit(2);
    }
    return static_cast<int>(v);
}

inline uint64_t parse_size(const char *text, const char *name) {
    char *end = nullptr;
    const unsigned long long v = std::strtoull(text, &end, 10);
    if (*text == '\0' || *end != '\0') {
        std::cerr << "Invalid " << name << ": " << text << "\n";
        std::exit(2);
    }
    return static_cast<uint64_t>(v);
}

inline double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Scheduling from the environment: MF_SCHEDULE=static|resplit, MF_BUDGET,
// MF_WAVES_PER_CU, MF_BLOCK.  None of them changes a count.
inline mf_solve_options options_from_env() {
    mf_solve_options o;
    std::memset(&o, 0, sizeof(o));
    o.schedule = MF_SCHED_RESPLIT;
    if (const char *s = std::getenv("MF_SCHEDULE")) {
        if (std::strcmp(s, "static") == 0) o.schedule = MF_SCHED_STATIC;
    }
    if (const char *s = std::getenv("MF_BUDGET")) o.budget = std::atoi(s);
    if (const char *s = std::getenv("MF_WAVES_PER_CU")) o.waves_per_cu = std::atoi(s);
    if (const char *s = std::getenv("MF_BLOCK")) o.block = std::atoi(s);
    o.count_events = 0;
    return o;
}

[[noreturn]] inline void die(int status, const char *what) {
    std::cerr << what << " failed: " << mf_status_string(status) << ": " << mf_last_error() << "\n";
    std::exit(1);
}

}  // namespace mfcli
