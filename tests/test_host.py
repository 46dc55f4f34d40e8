"""CPU tests of the product's host half: frontier generators, case stream,
library exports and the drop-in CLIs' argument handling (no GPU compute)."""
import ctypes
import json
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "map_folding_amd", "bin")


def test_library_exports_every_declared_symbol(mf):
    lib = mf._abi.lib()
    declared = mf.declared_symbols()
    assert "mf_sym_solve" in declared and "mf_wasm_solve_cases" in declared
    for name in declared:
        assert hasattr(lib, name), name


def test_no_oracle_in_product():
    """The product never loads the oracle (it is only the checker)."""
    for dirpath, _, files in os.walk(os.path.join(ROOT, "map_folding_amd")):
        for fn in files:
            if fn.endswith((".py", ".cpp", ".hpp", ".hip", ".h")) or fn == "Makefile":
                text = open(os.path.join(dirpath, fn), errors="replace").read()
                assert "oracle" not in text.replace("the oracle", ""), fn


@pytest.mark.parametrize("dims,depth", [([5, 5], 15), ([4, 5], 9), ([6, 6], 18), ([3, 7], 10),
                                        ([2, 3, 2], 6), ([1, 6], 3)])
def test_lunnon_frontier_matches_oracle(mf, oracle, dims, depth):
    ours = mf.lunnon_frontier(dims, depth)
    ref, _ = oracle.lunnon_frontier(dims, depth)
    assert ours.shape == ref.shape
    assert np.array_equal(ours, ref)


def test_lunnon_frontier_sizes_golden(mf, golden):
    for key, value in golden["lunnon_frontier_sizes"].items():
        if key.startswith("_"):
            continue
        shape, depth = key.split("@")
        g = int(shape.split("x")[0])
        assert mf.lunnon_frontier([g, g], int(depth)).shape[0] == value, key


@pytest.mark.parametrize("n", [6, 7, 8])
def test_sym_seeds_match_oracle(mf, oracle, n):
    c1, m1, r1 = mf.sym_seeds(n)
    c2, m2, r2 = oracle.sym_seeds(n)
    assert r1 == r2
    assert np.array_equal(c1, c2) and np.array_equal(m1, m2)


@pytest.mark.parametrize("n,e,start,count", [(7, 24, 0, 500), (7, 24, 200000, 300), (8, 24, 414000, 2000),
                                             (8, 28, 3000000, 700), (6, 16, 100, 50), (7, 30, 12345, 400)])
def test_sym_case_stream_matches_oracle(mf, oracle, n, e, start, count):
    c1, m1 = mf.sym_cases(n, e, start, count)
    c2, m2 = oracle.sym_cases(n, e, start, count)
    assert c1.shape == c2.shape
    assert np.array_equal(c1, c2) and np.array_equal(m1, m2)


def test_sym9_seeds_match_oracle(mf, oracle):
    c1, m1, r1 = mf.sym_seeds(9)
    c2, m2, r2 = oracle.sym9_seeds(9)
    assert r1 == r2 == 3721732
    assert c1.shape == (465883, 128)
    assert np.array_equal(c1, c2) and np.array_equal(m1, m2)


@pytest.mark.parametrize("e,start,count,center", [(64, 123456, 300, None), (30, 1000, 500, None),
                                                  (20, 100, 200, 3), (26, 5000, 300, 5)])
def test_sym9_case_stream_matches_oracle(mf, oracle, e, start, count, center):
    with mf.sym_numbering(center=center or 0):
        c1, m1 = mf.sym_cases(9, e, start, count)
    c2, m2 = oracle.sym9_cases(9, e, start, count, c=center)
    assert c1.shape == c2.shape == (count, 128)
    assert np.array_equal(c1, c2) and np.array_equal(m1, m2)


def test_sym9_seed_slice_is_an_offset_window(mf, oracle):
    """Case k of the seed slice starting at seed s is case (cases under seeds
    before s) + k of the whole stream."""
    off = int(mf.sym_seed_case_counts(9, 30, 0, 100).sum())
    c1, m1 = mf.sym_cases(9, 30, 5, 100, seed_start=100, seed_count=50)
    c2, m2 = oracle.sym9_cases(9, 30, off + 5, 100)
    assert np.array_equal(c1, c2) and np.array_equal(m1, m2)
    # a slice ends at its last seed
    per = mf.sym_seed_case_counts(9, 30, 100, 3)
    c3, _ = mf.sym_cases(9, 30, 0, 10 ** 6, seed_start=100, seed_count=3)
    assert c3.shape[0] == int(per.sum())


def test_sym_case_totals_golden(mf, golden):
    for key in ("7@24", "8@24", "8@26", "8@28", "8@30", "8@32"):
        n, e = map(int, key.split("@"))
        assert mf.sym_case_total(n, e) == golden["sym_case_totals"][key], key


def test_sym_case_stream_end(mf):
    total = mf.sym_case_total(7, 24)
    cyc, mul = mf.sym_cases(7, 24, total - 5, 100)
    assert cyc.shape[0] == 5
    cyc, mul = mf.sym_cases(7, 24, total + 10, 100)
    assert cyc.shape[0] == 0


def test_bad_arguments_raise(mf):
    with pytest.raises(mf.MFError):
        mf.lunnon_frontier([9, 9], 10)     # 81 leaves > 64
    with pytest.raises(mf.MFError):
        mf.lunnon_frontier([5, 5], 0)
    with pytest.raises(mf.MFError):
        mf.sym_cases(8, 64, 0, 10)          # expanded = n^2 is rejected
    with pytest.raises(mf.MFError):
        mf.sym_cases(9, 24, 0, 10)


def _run(args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    e["HIP_VISIBLE_DEVICES"] = e.get("HIP_VISIBLE_DEVICES", "")
    return subprocess.run(args, capture_output=True, text=True, env=e, timeout=120)


@pytest.mark.skipif(not os.path.exists(os.path.join(BIN, "sym_fold")), reason="CLIs not built")
def test_cli_argument_contract():
    r = _run([os.path.join(BIN, "gpu_folding")])
    assert r.returncode == 1 and "Usage" in r.stdout
    r = _run([os.path.join(BIN, "gpu_folding"), "9", "10"])
    assert r.returncode == 1 and "Grid too large" in r.stdout
    r = _run([os.path.join(BIN, "sym_fold")])
    assert r.returncode == 1
    r = _run([os.path.join(BIN, "sym_fold"), "9"])
    assert r.returncode == 1 and "between 1 and 8" in r.stderr
    r = _run([os.path.join(BIN, "sym_fold"), "x"])
    assert r.returncode == 2
    r = _run([os.path.join(BIN, "sym_fold"), "2"])
    assert r.returncode == 0 and r.stdout == "8\n"
    r = _run([os.path.join(BIN, "sym_fold"), "1", "4"])
    assert r.stdout == "1\n"
    r = _run([os.path.join(BIN, "sym_fold"), "8", "1"], {"MF_SETUP_ONLY": "1"})
    assert r.returncode == 0 and r.stdout == "414992\n"
    assert "center_cases=9394" in r.stderr and "reduced_cases=1225" in r.stderr
    r = _run([os.path.join(BIN, "sym_fold"), "8", "1"], {"MF_EXPANDED_N": "64"})
    assert r.returncode == 1


def test_sym_expand_matches_oracle(mf, oracle):
    """Children at level+1 in ascending child order (recurse_generate,
    sym_fold.cpp:252-265): per-parent counts and the depth-16 prefixes."""
    import numpy as np
    pre = oracle.sym_prefixes(5, 12)
    ch, par = mf.sym_expand(5, 12, pre)
    assert np.array_equal(np.bincount(par, minlength=len(pre)),
                          [oracle.sym_count_to(5, c, 12, 13)[0] for c in pre])
    assert np.array_equal(ch[:, :13], oracle.sym_prefixes(5, 13)[:, :13])
    with pytest.raises(mf.MFError):
        mf.sym_expand(5, 24, pre)   # no children past level n^2 - 2


def _sorted_windows():
    import json
    with open(os.path.join(os.path.dirname(__file__), "golden", "sym_sorted_windows.json")) as f:
        return json.load(f)["items"]


@pytest.mark.skipif(not os.path.exists(os.path.join(BIN, "sym_fold")), reason="CLIs not built")
@pytest.mark.parametrize("item", _sorted_windows(), ids=lambda it: f"n{it['n']}-d{it['cost_depth']}-{it['sort'] or 'nosort'}")
def test_cli_cost_depth_score_line_matches_reference(item):
    """MF_COST_DEPTH scoring (estimate_work over the materialised case list)
    prints the reference's min/max/total scores (setup-only: no GPU needed)."""
    r = _run([os.path.join(BIN, "sym_fold"), str(item["n"]), "8", str(item["max"]), str(item["start"])],
             {"MF_COST_DEPTH": str(item["cost_depth"]), "MF_SORT_CASES": item["sort"], "MF_SETUP_ONLY": "1"})
    assert r.returncode == 0, r.stderr
    line = next(l for l in r.stderr.splitlines() if l.startswith("cost_depth="))
    assert line.split(" score_s=")[0] == item["cost_line"]
    sorted_line = [l for l in r.stderr.splitlines() if l.startswith("sorted_cases=")]
    want = [] if item["sort"] in ("", "0") else [f"sorted_cases={'asc' if item['sort'] == 'asc' else 'desc'}"]
    assert sorted_line == want


def _clean(text):
    out = []
    for line in text.splitlines():
        if "total_s=" in line or line.startswith(("progress=", "stream selected=")):
            continue
        out.append(re.sub(r" \w*_s=\S+", "", line))
    return out


def _sym9_host_runs():
    with open(os.path.join(ROOT, "tests", "golden", "sym9_cli.json")) as f:
        runs = json.load(f)["runs"]
    return [r for r in runs if r["rc"] != 0 or "MF_SETUP_ONLY" in r["env"]]


@pytest.mark.skipif(not os.path.exists(os.path.join(BIN, "sym_fold9")), reason="CLIs not built")
@pytest.mark.parametrize("run", _sym9_host_runs(), ids=lambda r: " ".join(r["args"]) + " " +
                         " ".join(f"{k}={v}" for k, v in sorted(r["env"].items())))
def test_cli_sym_fold9_host_side_matches_reference(run):
    """sym_fold9 runs that end before any counting (MF_SETUP_ONLY, argument
    errors): stdout, exit status and the clock-free stderr lines of the
    reference binary (tests/golden/make_sym9_cli.py)."""
    env = {k: v for k, v in os.environ.items() if not k.startswith("MF_")}
    env.update(run["env"])
    env["HIP_VISIBLE_DEVICES"] = ""
    r = subprocess.run([os.path.join(BIN, "sym_fold9")] + run["args"], capture_output=True, text=True, env=env,
                       timeout=300)
    assert r.returncode == run["rc"], r.stderr
    assert r.stdout == run["stdout"]
    if run["rc"] == 0:
        assert _clean(r.stderr) == run["stderr"]


@pytest.mark.parametrize("order,lock", [("reverse", None), ("opposite", None), ("axes_first", "5"),
                                        ("corners_first", None), ("corners_first", "1")])
def test_sym_cell_orders_match_oracle(mf, oracle, monkeypatch, order, lock):
    """MF_ORDER / MF_LOCK_CENTER_N renumber odd squares (sym_fold9.cpp:160-229):
    product seeds and case windows equal the oracle's under each numbering
    (7x7 with its 3x3 center, and 9x9 with the 5x5 center whose symmetry maps
    follow the renumbered center)."""
    # The oracle reads the reference's environment variables; the library
    # takes the numbering explicitly.
    monkeypatch.setenv("MF_ORDER", order)
    if lock:
        monkeypatch.setenv("MF_LOCK_CENTER_N", lock)
    cases = [(7, 26, 3000, 400)] + ([(9, 62, 1000, 300)] if order == "reverse" else [])
    for n, e, start, count in cases:
        with mf.sym_numbering(order=order, lock_center_n=int(lock) if lock else 3):
            c1, m1, r1 = mf.sym_seeds(n)
            k1, w1 = mf.sym_cases(n, e, start, count)
        c2, m2, r2 = oracle.sym9_seeds(n)
        assert r1 == r2 and np.array_equal(c1[:, :c2.shape[1]] if n < 9 else c1, c2[:, :c1.shape[1]])
        assert np.array_equal(m1, m2)
        k2, w2 = oracle.sym9_cases(n, e, start, count)
        assert np.array_equal(k1, k2[:, :k1.shape[1]]) and np.array_equal(w1, w2)


def test_library_ignores_numbering_environment(mf, oracle, monkeypatch):
    """The library reads no MF_ORDER / MF_CENTER_N: with them set, the 7x7
    case stream and seeds are still the spiral ones of sym_fold.cpp (only the
    sym_fold9 CLI passes them on, through mf_sym_set_numbering)."""
    want_c, want_m = oracle.sym_cases(7, 26, 3000, 200)
    monkeypatch.setenv("MF_ORDER", "reverse")
    monkeypatch.setenv("MF_CENTER_N", "5")
    c1, m1 = mf.sym_cases(7, 26, 3000, 200)
    assert np.array_equal(c1, want_c) and np.array_equal(m1, want_m)
    assert mf.sym_get_numbering() == (0, 0, 3)


def test_numbering_setting_is_checked(mf):
    with pytest.raises(mf.MFError):
        mf.sym_set_numbering(center=7)
    with pytest.raises(mf.MFError):
        mf.sym_set_numbering(order=9)
    # a center of the wrong parity is refused where n is known
    with mf.sym_numbering(center=4):
        with pytest.raises(mf.MFError):
            mf.sym_seeds(7)
    assert mf.sym_get_numbering() == (0, 0, 3)


@pytest.mark.parametrize("dims,depth", [([2, 3], 4), ([3, 3], 6), ([2, 4], 6), ([3, 4], 7), ([2, 2, 2], 6)])
def test_lunnon_reachable_states_are_the_frontier(mf, dims, depth):
    """mf_lunnon_check_states accepts exactly the states the folding DFS
    reaches: over every circular list of leaves 0..depth with leaf 1 below
    the root, the accepted ones are the frontier at that depth (the GPU
    kernels' gap walks rely on this: they carry no step guard)."""
    import ctypes as C
    import itertools
    L = mf._abi.lib()
    d = np.asarray(dims, dtype=np.int32)
    front = {bytes(r) for r in mf.lunnon_frontier(dims, depth)}
    accepted = set()
    for perm in itertools.permutations(range(2, depth + 1)):
        order = [0, 1, *perm]                      # stack order from the root
        b = np.zeros(depth + 1, dtype=np.uint8)
        for k, leaf in enumerate(order):
            b[leaf] = order[(k + 1) % len(order)]
        bad = C.c_int64()
        rc = L.mf_lunnon_check_states(d.ctypes.data_as(C.POINTER(C.c_int32)), len(dims), depth,
                                      b.ctypes.data_as(C.POINTER(C.c_uint8)), 1, C.byref(bad))
        if rc == 0:
            accepted.add(bytes(b))
        else:
            assert bad.value == 0
    assert accepted == front
