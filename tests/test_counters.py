"""Per-item event counters of the DFS kernels are 32-bit and folded into
64-bit lane totals before they can wrap (map_folding_amd/csrc/mf_counters.hpp).
The fold logic is compiled here for the host, exactly as the kernels include
it, and driven like the kernels drive it: items of more than 2^32 events,
folds every `budget` steps, the rest added when the item retires."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r"""
#include <cstdio>
#include <cstdlib>
#include "mf_counters.hpp"
using namespace mf;

static uint64_t rng = 0x9e3779b97f4a7c15ull;
static uint32_t next(uint32_t bound) {   // uniform in [0, bound]
    rng ^= rng << 13; rng ^= rng >> 7; rng ^= rng << 17;
    return static_cast<uint32_t>(rng % (bound + 1ull));
}

int main(int argc, char **argv) {
    const uint64_t target = std::strtoull(argv[1], nullptr, 10);   // events per counter, at least
    const uint32_t budget = static_cast<uint32_t>(std::atoi(argv[2]));
    const int fold = std::atoi(argv[3]);                             // 0: no fold (shows the wrap)
    uint32_t nodes = 0, walks = 0, steps = 0, cands = 0;
    FoldTotals lane;                      // the lane's 64-bit totals
    unsigned long long per_origin = 0;    // the item's per-origin node log
    unsigned long long exact_n = 0, exact_w = 0, exact_s = 0, exact_c = 0;
    uint64_t nsteps = 0;
    bool kept_node = true;
    while (exact_s < target) {
        // one DFS step: at most kMaxStepEvents of each kind, one node at most
        const uint32_t dn = next(1), dw = next(4), ds = next(kMaxStepEvents), dc = next(64);
        nodes += dn; walks += dw; steps += ds; cands += dc;
        exact_n += dn; exact_w += dw; exact_s += ds; exact_c += dc;
        if (fold && ++nsteps % budget == 0) {
            const bool had = nodes != 0;
            per_origin += fold_item_counters(nodes, walks, steps, cands, lane);
            if (had && nodes == 0) kept_node = false;
        }
    }
    // the item retires: the rest of its counters go to the totals
    lane.nodes += nodes; lane.walks += walks; lane.steps += steps; lane.cands += cands;
    per_origin += nodes;
    const bool ok = lane.nodes == exact_n && lane.walks == exact_w && lane.steps == exact_s &&
                    lane.cands == exact_c && per_origin == exact_n && kept_node;
    std::printf("%d %llu %llu %llu\n", ok ? 1 : 0, static_cast<unsigned long long>(lane.steps),
                exact_s, static_cast<unsigned long long>(per_origin));
    return 0;
}
"""


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    d = tmp_path_factory.mktemp("fold")
    src = d / "fold.cpp"
    src.write_text(DRIVER)
    exe = d / "fold"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "map_folding_amd", "csrc"), str(src),
                    "-o", str(exe)], check=True)
    return str(exe)


def _run(exe, target, budget, fold):
    out = subprocess.run([exe, str(target), str(budget), str(fold)], capture_output=True, text=True, check=True)
    ok, got, exact, per = out.stdout.split()
    return ok == "1", int(got), int(exact), int(per)


@pytest.mark.parametrize("budget", [1, 64, 4096])
def test_fold_keeps_counts_exact_past_2_pow_32(driver, budget):
    ok, got, exact, _ = _run(driver, 3 * 2 ** 32, budget, 1)
    assert exact > 2 ** 33
    assert ok and got == exact


def test_without_the_fold_the_counters_wrap(driver):
    """The same item without folds: the 32-bit step counter wraps, which is
    what the fold prevents (the test above would catch a missing fold)."""
    ok, got, exact, _ = _run(driver, 3 * 2 ** 32, 64, 0)
    assert not ok and got == exact % 2 ** 32
