"""GPU parity of the 9x9 symmetry DFS (u128 masks, 96-cell LDS rows, 5x5
center seeds; experiments/9x9/sym_fold9.cpp) through the C ABI and the
sym_fold9 command-line tool.

Goldens: tests/golden/sym9.json (MF_STREAM_CASES windows of the reference
binary) and tests/golden/sym9_cli.json (whole sym_fold9 runs: stdout and the
clock-free stderr lines, incl. MF_PROFILE_WORK summaries and seed slices).
Per-case counts and event counters are checked against the oracle's 9x9
restatement, bit-exactly."""
import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN9 = os.path.join(ROOT, "map_folding_amd", "bin", "sym_fold9")

SCHEDULES = {
    "static": dict(schedule=0),
    "resplit": dict(schedule=1),
    "resplit_tight": dict(schedule=1, budget=4, tail=4),
    "resplit_nograph": dict(schedule=1, graph=2),
    # one page walk at a time (the instance ring orders run on) instead of pairs
    "resplit_unpaired": dict(schedule=1, core=4),
}


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _load(name):
    with open(os.path.join(ROOT, "tests", "golden", name)) as f:
        return json.load(f)


@pytest.mark.parametrize("w", _load("sym9.json")["windows"],
                         ids=lambda w: f"e{w['expanded']}_s{w['start']}_m{w['max']}")
def test_sym9_window(gpu, w):
    _say(f"sym 9x9 d{w['expanded']} [{w['start']},+{w['max']})")
    ans, res = gpu.sym_fold(9, w["expanded"], w["start"], w["max"])
    assert ans == w["answer"]


@pytest.mark.parametrize("sched", list(SCHEDULES))
@pytest.mark.parametrize("e,start,m", [(60, 0, 200), (62, 200000, 400), (44, 0, 50)])
def test_sym9_per_case_and_events(gpu, oracle, sched, e, start, m):
    _say(f"sym 9x9 events d{e} [{start},+{m}) {sched}")
    cyc, mul = gpu.sym_cases(9, e, start, m)
    ocyc, omul = oracle.sym9_cases(9, e, start, m)
    assert np.array_equal(cyc, ocyc) and np.array_equal(mul, omul)
    pc0, pn0, tot, st = oracle.sym9_solve(9, e, cyc, mul)
    res, pc, pn = gpu.sym_solve(9, e, cyc, mul, gpu.SolveOptions(**SCHEDULES[sched]), per_item=True)
    assert np.array_equal(pc, pc0)
    assert np.array_equal(pn, pn0)
    assert res.count == tot
    assert (res.nodes, res.walks, res.steps, res.leaves) == (
        st["nodes"], st["page_walks"], st["walk_steps"], st["leaves"])


@pytest.mark.parametrize("order", ["reverse", "corners_first"])
def test_ring_orders_on_u64_cores(gpu, oracle, monkeypatch, order):
    """7x7 under a ring order (sym_fold9.cpp:160-229): some levels walk no
    page or more than two (4 under reverse, 16 under corners_first), which
    the paired two-page walk of the u64 kernels does not cover; the library
    runs such maps one page walk at a time.  Every schedule and core, and
    the deepening pass, count the oracle's per-case totals and events."""
    monkeypatch.setenv("MF_ORDER", order)   # the oracle reads the reference's environment
    ocyc, omul = oracle.sym9_cases(7, 26, 3000, 300)
    pc0, pn0, tot, st = oracle.sym9_solve(7, 26, ocyc, omul)
    with gpu.sym_numbering(order=order):
        cyc, mul = gpu.sym_cases(7, 26, 3000, 300)
        assert np.array_equal(cyc, ocyc[:, :cyc.shape[1]]) and np.array_equal(mul, omul)
        for opts in (dict(schedule=0), dict(schedule=1), dict(schedule=1, deepen=30),
                     dict(schedule=1, core=1), dict(schedule=1, core=3)):
            _say(f"sym 7x7 {order} {opts}")
            res, pc, pn = gpu.sym_solve(7, 26, cyc, mul, gpu.SolveOptions(**opts), per_item=True)
            assert np.array_equal(pc, pc0) and np.array_equal(pn, pn0)
            assert res.count == tot
            assert (res.nodes, res.walks, res.steps, res.leaves) == (
                st["nodes"], st["page_walks"], st["walk_steps"], st["leaves"])


def test_sym9_seed_slice_windows(gpu):
    """Seed slices (MF_SEED_START / MF_SEED_COUNT): windows are additive, and
    a window that crosses from one seed into the next is the two seeds'
    single-seed windows put together.  Seed 1001 holds 4,870,413 cases at
    depth 58 (whole seeds there reach 10^8 cases, so windows stay bounded)."""
    _say("sym 9x9 seed slices")
    whole, _ = gpu.sym_fold(9, 58, 0, 3000, seed_start=1001, seed_count=1)
    parts = sum(gpu.sym_fold(9, 58, k, 1000, seed_start=1001, seed_count=1)[0] for k in (0, 1000, 2000))
    assert whole == parts
    c0 = int(gpu.sym_seed_case_counts(9, 58, 1001, 1)[0])
    assert c0 == 4870413
    a, _ = gpu.sym_fold(9, 58, c0 - 200, 500, seed_start=1001, seed_count=2)
    b = (gpu.sym_fold(9, 58, c0 - 200, 500, seed_start=1001, seed_count=1)[0] +
         gpu.sym_fold(9, 58, 0, 300, seed_start=1002, seed_count=1)[0])
    assert a == b


def _clean(text):
    out = []
    for line in text.splitlines():
        if "total_s=" in line or line.startswith(("progress=", "stream selected=")):
            continue
        out.append(re.sub(r" \w*_s=\S+", "", line))
    return out


_RUNS = [r for r in _load("sym9_cli.json")["runs"] if r["rc"] == 0 and "MF_SETUP_ONLY" not in r["env"]]


@pytest.mark.parametrize("run", _RUNS, ids=lambda r: " ".join(r["args"]) + " " +
                         " ".join(f"{k}={v}" for k, v in sorted(r["env"].items())))
def test_sym_fold9_cli(gpu, run):
    _say(f"cli sym_fold9 {run['args']} {run['env']}")
    env = {k: v for k, v in os.environ.items() if not k.startswith("MF_")}
    env.update(run["env"])
    r = subprocess.run([BIN9] + run["args"], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stderr
    assert r.stdout == run["stdout"]
    assert _clean(r.stderr) == run["stderr"]


def test_sym9_wasm_abi(gpu, oracle):
    """mf_wasm_solve_cases for n = 9 (sym_kernel.c MAX_SIDE 9, u128 masks):
    112-byte records, count = sum mult x completions, nodes = recurse_count
    calls, the calls that reach stop_i = n^2 included (get_good calls + leaves)."""
    import ctypes as C
    _say("wasm ABI 9x9")
    cyc, mul = oracle.sym9_cases(9, 62, 200000, 120)
    mults = [int(m) * (1 + (i % 5)) for i, m in enumerate(mul)]
    L = gpu._abi.lib()
    base = C.cast(L.mf_wasm_input_ptr(), C.c_void_p).value
    arr = np.frombuffer((C.c_uint8 * (112 * 256)).from_address(base), dtype=np.uint8)
    arr[:] = 0
    for k, (row, m) in enumerate(zip(cyc, mults)):
        arr[k * 112:k * 112 + 81] = row[:81]
        arr[k * 112 + 100:k * 112 + 108] = np.frombuffer(m.to_bytes(8, "little"), dtype=np.uint8)
    pc, _, _, st = oracle.sym9_solve(9, 62, cyc, None)
    want = sum(int(c) * m for c, m in zip(pc, mults))
    assert L.mf_wasm_solve_cases(9, 62, 81, len(mults)) == 0
    out = L.mf_wasm_output_ptr()
    w = [int(out[i]) for i in range(8)]
    assert w[0] | (w[1] << 32) | (w[2] << 64) | (w[3] << 96) == want
    assert w[4] | (w[5] << 32) == st["nodes"] + st["leaves"]
    assert w[6] == 0
