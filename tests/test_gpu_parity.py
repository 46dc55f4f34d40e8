"""GPU parity: the HIP kernels (through the C ABI) against the oracle and the
golden totals.  Integer results must match bit-exactly, including the exact
event counters (nodes, walks, steps, leaves) that make up the roofline's
algorithmic-work figure."""
import collections
import json
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "map_folding_amd", "bin")

SCHEDULES = {
    "static": dict(schedule=0),
    "resplit": dict(schedule=1),
    "resplit_tight": dict(schedule=1, budget=8, tail=8),
    # the wide u64-mask stack (default is the compact one)
    "static_wide": dict(schedule=0, core=1),
    "resplit_wide": dict(schedule=1, core=1),
    "resplit_wide_tight": dict(schedule=1, core=1, budget=4, tail=4),
    "resplit_tiny": dict(schedule=1, budget=2, tail=2),
    # host-driven rounds instead of the captured HIP graph of rounds
    "resplit_nograph": dict(schedule=1, graph=2),
    "resplit_tight_nograph": dict(schedule=1, budget=8, tail=8, graph=2),
    # two lanes per subtree (sym, up to 8x8; re-split only)
    "resplit_pair": dict(schedule=1, core=3),
    "resplit_pair_tight": dict(schedule=1, core=3, budget=4, tail=4),
    # unpaired walks (sym: one page at a time; Lunnon: the two dimensions of a
    # rectangle one after the other instead of in one loop)
    "static_unpaired": dict(schedule=0, core=4),
    "resplit_unpaired": dict(schedule=1, core=4),
}
LUNNON_SCHEDULES = ["static", "resplit", "resplit_tight", "resplit_tight_nograph", "static_unpaired",
                    "resplit_unpaired", "resplit_wide"]


def _opts(mf, name, **kw):
    d = dict(SCHEDULES[name])
    d.update(kw)
    return mf.SolveOptions(**d)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


# ------------------------------------------------------------------------------ Lunnon


@pytest.mark.parametrize("sched", LUNNON_SCHEDULES)
def test_lunnon_5x5_per_state_parity(gpu, oracle, golden, sched):
    _say(f"lunnon 5x5 D15 per-state, schedule={sched}")
    states = gpu.lunnon_frontier([5, 5], 15)
    res, pc, pn = gpu.lunnon_solve([5, 5], 15, states, _opts(gpu, sched), per_item=True)
    cnt, nodes, tot, st = oracle.lunnon_resume([5, 5], 15, states)
    assert res.count == tot == golden["square_totals"]["5"]
    assert np.array_equal(pc * 25, cnt)            # per-state completions x n
    assert np.array_equal(pn, nodes)
    assert res.nodes == st["inserts"]
    assert res.walks == st["gapcomps"]
    assert res.steps == st["walk_steps"]
    assert res.cands == st["candidates"]
    assert res.leaves == st["completions"]


@pytest.mark.parametrize("dims,depth", [([4, 5], 9), ([3, 6], 7), ([2, 2, 3], 5), ([1, 8], 3), ([2, 3], 2),
                                        ([2, 2, 2, 2], 6), ([4, 4], 1), ([3, 3], 7)])
def test_lunnon_small_maps(gpu, oracle, dims, depth):
    _say(f"lunnon {dims} D{depth}")
    total, full = oracle.lunnon_count(dims)
    for sched in ("static", "resplit_tight"):
        count, nstates, res = gpu.gpu_folding(dims, depth, _opts(gpu, sched))
        assert count == total, (dims, sched)
        states, front = oracle.lunnon_frontier(dims, depth)
        assert nstates == states.shape[0]
        assert res.nodes + front["inserts"] == full["inserts"]
        assert res.steps + front["walk_steps"] == full["walk_steps"]


@pytest.mark.parametrize("sched", ["static", "resplit"])
def test_lunnon_6x6_depth18_full(gpu, golden, sched):
    _say(f"lunnon 6x6 D18 full, schedule={sched}")
    count, nstates, res = gpu.gpu_folding([6, 6], 18, _opts(gpu, sched, count_events=False))
    assert nstates == golden["lunnon_frontier_sizes"]["6x6@18"]
    assert count == golden["square_totals"]["6"]


def _rect_draw(seed, k):
    """64 rectangles: p uniform in {2,3,4}, then q uniform with 12 <= p*q <= 30."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(k):
        p = int(rng.integers(2, 5))
        qs = [q for q in range(1, 31) if 12 <= p * q <= 30]
        out.append((p, int(rng.choice(qs))))
    return out


def test_lunnon_batch_small_rectangles(gpu, golden):
    shapes = [s for s in _rect_draw(7, 64) if s[0] * s[1] <= 24][:40]
    _say(f"lunnon batch {len(shapes)} maps")
    depths = [max(1, p * q // 2) for p, q in shapes]
    totals, res = gpu.lunnon_batch(shapes, depths, _opts(gpu, "resplit"))
    for (p, q), t in zip(shapes, totals):
        assert t == golden["rectangles"][f"{p}x{q}"], (p, q)
    assert res.count == sum(totals)


def test_lunnon_batch_bench_draw_full_size(gpu, golden):
    """The benchmark's own draw: 64 rectangles from seed 0 at full size (p in
    {2,3,4}, 12 <= pq <= 30, so 2x15 and 3x10 are in it), each with the
    bench's default per-map depth, in one resident batch frontier.  Every
    per-map total must equal `mf p q`, including those above 2^32
    (2x15 = 39,409,195,740; 3x10 = 17,257,455,960)."""
    import bench
    shapes = bench.rect_shapes(0)
    assert shapes == _rect_draw(0, 64)
    assert (2, 15) in shapes and (3, 10) in shapes
    depths = [bench.rect_depth(p, q) for p, q in shapes]
    _say(f"lunnon batch: the 64-rectangle draw of seed 0, depths {min(depths)}..{max(depths)}")
    front = gpu.LunnonBatchFrontier(shapes, depths)
    for sched in ("resplit", "resplit_nograph", "resplit_unpaired"):
        totals, res = front.solve_maps(_opts(gpu, sched, count_events=False))
        for (p, q), t in zip(shapes, totals):
            assert t == golden["rectangles"][f"{p}x{q}"], (p, q, sched)
        assert res.count == sum(totals)
    assert max(totals) > 2 ** 32
    front.close()


# ------------------------------------------------------------------------------ sym


def _chunk(golden, n, e, start, mx):
    for it in golden["sym_chunks"]["items"]:
        if (it["n"], it["expanded"], it["start"], it["max"]) == (n, e, start, mx):
            return it
    raise KeyError((n, e, start, mx))


@pytest.mark.parametrize("sched", ["resplit", "resplit_tight", "resplit_tiny", "resplit_wide",
                                   "resplit_wide_tight", "resplit_nograph", "resplit_tight_nograph",
                                   "resplit_pair", "resplit_pair_tight", "resplit_unpaired"])
def test_sym_8x8_first_300_events(gpu, golden, sched):
    _say(f"sym 8x8 d36 [0,300) schedule={sched}")
    g = _chunk(golden, 8, 36, 0, 300)
    cyc, mul = gpu.sym_cases(8, 36, 0, 300)
    res, pc, pn = gpu.sym_solve(8, 36, cyc, mul, _opts(gpu, sched), per_item=True)
    assert res.count * 64 == g["answer"]
    assert res.nodes == g["nodes"]
    assert res.walks == g["page_walks"]
    assert res.steps == g["walk_steps"]
    assert res.leaves == g["leaves"]
    assert int(pn.sum()) == g["nodes"]
    assert int((pc * mul).sum()) * 64 == g["answer"]


def _small_cases(gpu, n, e, start, k, max_nodes):
    """Cases of [start, start+k) whose subtree has at most max_nodes nodes (so
    that one lane per case finishes quickly), found with the re-split schedule."""
    cyc, mul = gpu.sym_cases(n, e, start, k)
    _, _, pn = gpu.sym_solve(n, e, cyc, mul, _opts(gpu, "resplit"), per_item=True)
    keep = pn <= max_nodes
    return cyc[keep], mul[keep]


def test_sym_per_case_parity_vs_oracle(gpu, oracle):
    _say("sym per-case parity vs oracle (small 8x8 d36 and 7x7 d24 cases)")
    for n, e, start, k in ((8, 36, 5000, 400), (7, 24, 1000, 400)):
        cyc, mul = _small_cases(gpu, n, e, start, k, 400000)
        assert cyc.shape[0] >= 20
        cnt, nodes, tot, st = oracle.sym_solve(n, e, cyc, mul)
        for sched in ("static", "resplit_tight", "resplit_tiny", "static_wide", "resplit_wide_tight",
                      "resplit_pair_tight"):
            res, pc, pn = gpu.sym_solve(n, e, cyc, mul, _opts(gpu, sched), per_item=True)
            assert np.array_equal(pc, cnt), sched
            assert np.array_equal(pn, nodes), sched
            assert res.count == tot
            assert (res.nodes, res.walks, res.steps, res.leaves) == (
                st["nodes"], st["page_walks"], st["walk_steps"], st["leaves"])


def test_sym_headline_chunk_8x8(gpu, golden):
    """The 50,000-case 8x8 depth-36 chunk at 120,000,000 (full size)."""
    _say("sym 8x8 d36 [120000000,+50000) resplit")
    g = _chunk(golden, 8, 36, 120000000, 50000)
    cyc, mul = gpu.sym_cases(8, 36, 120000000, 50000)
    assert cyc.shape[0] == 50000
    res, _, _ = gpu.sym_solve(8, 36, cyc, mul, _opts(gpu, "resplit", count_events=False))
    assert res.count * 64 == g["answer"]
    g2 = _chunk(golden, 8, 36, 120000000, 200)
    res2, _, _ = gpu.sym_solve(8, 36, cyc[:200], mul[:200], _opts(gpu, "resplit_tight"))
    assert res2.count * 64 == g2["answer"]


def test_sym_7x7_chunk(gpu, golden):
    _say("sym 7x7 d24 [200000,+20000)")
    g = _chunk(golden, 7, 24, 200000, 20000)
    ans, res = gpu.sym_fold(7, 24, 200000, 20000)
    assert ans == g["answer"]


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6])
def test_sym_fold_small_squares(gpu, golden, n):
    _say(f"sym_fold n={n}")
    ans, _ = gpu.sym_fold(n)
    assert ans == golden["square_totals"][str(n)]


def test_sym_chunk_additivity_gpu(gpu):
    _say("sym chunk additivity")
    whole, _ = gpu.sym_fold(8, 36, 77777, 3000)
    parts = sum(gpu.sym_fold(8, 36, 77777 + s, 1000)[0] for s in (0, 1000, 2000))
    assert whole == parts


def test_resident_frontier_repeat(gpu, golden):
    _say("resident frontier solved twice")
    g = _chunk(golden, 8, 36, 0, 2000)
    cyc, mul = gpu.sym_cases(8, 36, 0, 2000)
    f = gpu.SymFrontier(8, 36, cyc, mul)
    r1, _, _ = f.solve(_opts(gpu, "resplit"))
    order = np.arange(2000, dtype=np.uint32)[::-1].copy()
    r2, _, _ = f.solve(_opts(gpu, "resplit_tight", order=order))
    f.close()
    assert r1.count * 64 == g["answer"] == r2.count * 64
    assert r1.nodes == r2.nodes


# ------------------------------------------------------------------------------ CLIs


def _run(args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run(args, capture_output=True, text=True, env=e, timeout=300)


def test_cli_gpu_folding_5x5(gpu):
    _say("cli gpu_folding 5 15")
    r = _run([os.path.join(BIN, "gpu_folding"), "5", "15"])
    assert r.returncode == 0, r.stderr
    assert r.stdout == "Number of partial states at depth 15: 13416\nTotal number of ways: 186086600\n"
    r = _run([os.path.join(BIN, "gpu_folding"), "5", "10"], {"MF_DIMS": "3x6"})
    assert r.stdout.splitlines()[-1] == "Total number of ways: 2016432"


def test_cli_sym_fold_chunk(gpu, golden):
    _say("cli sym_fold 8 chunk")
    g = _chunk(golden, 8, 36, 120000000, 200)
    r = _run([os.path.join(BIN, "sym_fold"), "8", "16", "200", "120000000"],
             {"MF_STREAM_CASES": "1", "MF_EXPANDED_N": "36"})
    assert r.returncode == 0, r.stderr
    assert r.stdout == f"{g['answer']}\n"
    r = _run([os.path.join(BIN, "sym_fold"), "6", "8"])
    assert r.stdout == f"{golden['square_totals']['6']}\n"


def _sorted_windows():
    with open(os.path.join(ROOT, "tests", "golden", "sym_sorted_windows.json")) as f:
        return json.load(f)["items"]


@pytest.mark.parametrize("item", _sorted_windows(), ids=lambda it: f"n{it['n']}-d{it['cost_depth']}-{it['sort'] or 'nosort'}")
def test_cli_sorted_windows_match_reference(gpu, item):
    """sym_fold's materialised mode with MF_COST_DEPTH / MF_SORT_CASES: the
    window of the scored and sorted case list counts what the reference's
    binary printed (tests/golden/make_sorted_windows.sh)."""
    _say(f"cli sym_fold sorted window {item}")
    r = _run([os.path.join(BIN, "sym_fold"), str(item["n"]), "8", str(item["max"]), str(item["start"])],
             {"MF_COST_DEPTH": str(item["cost_depth"]), "MF_SORT_CASES": item["sort"]})
    assert r.returncode == 0, r.stderr
    assert r.stdout == f"{item['answer']}\n"


def test_graph_rounds_match_host_rounds(gpu):
    """The captured graph of rounds and host-driven rounds count the same, and
    a cached graph is replayed correctly on repeated solves."""
    _say("graph vs host rounds, 8x8 d36 [120000000,+2000), tight budget")
    cyc, mul = gpu.sym_cases(8, 36, 120000000, 2000)
    f = gpu.SymFrontier(8, 36, cyc, mul)
    ref, pc0, pn0 = f.solve(_opts(gpu, "resplit_tight_nograph"), per_item=True)
    for rep in range(3):
        r, pc, pn = f.solve(_opts(gpu, "resplit_tight"), per_item=True)
        assert r.count == ref.count and r.nodes == ref.nodes and r.walks == ref.walks and r.steps == ref.steps
        assert np.array_equal(pc, pc0) and np.array_equal(pn, pn0)
        assert r.rounds >= 2
    f.close()


def test_pieces_and_orders_keep_counts(gpu):
    """Mixed-level pieces (split frontiers) add up per origin to the unsplit
    counts under any schedule and order; a non-permutation order is refused."""
    from map_folding_amd import cost
    _say("pieces: 7x7 d24 [1000,+300) split by the lookahead proxy")
    cyc, mul = gpu.sym_cases(7, 24, 1000, 300)
    ref, pc0, pn0 = gpu.sym_solve(7, 24, cyc, mul, _opts(gpu, "resplit"), per_item=True)
    pieces = cost.lookahead_split(7, 24, cyc, mul, 3.0, depth=4, max_extra=3)
    assert len(pieces) > cyc.shape[0] and len(set(pieces.levels.tolist())) > 1
    f = pieces.frontier(7)
    assert f.num == 300 and f.items == len(pieces)
    rng = np.random.default_rng(0)
    for sched, order in (("static", None), ("resplit", pieces.order()), ("resplit_tight", rng.permutation(f.items)),
                         ("static", rng.permutation(f.items))):
        r, pc, pn = f.solve(_opts(gpu, sched, order=order), per_item=True)
        assert r.count == ref.count and np.array_equal(pc, pc0), sched
        assert int(pn.sum()) + pieces.host_nodes == ref.nodes
    bad = np.zeros(f.items, dtype=np.uint32)
    with pytest.raises(gpu.MFError):
        f.solve(_opts(gpu, "resplit", order=bad))
    f.close()


def test_learned_split_7x7_window(gpu, golden):
    """The timed path of the learned schedule: pieces split by the committed
    cost model on the 7x7 depth-24 window [200000, +20000) count the window's
    golden total, and each origin's pieces add up to its unsplit count
    (sym_fold.cpp:491-527 defines the window)."""
    from map_folding_amd import cost
    _say("learned split: 7x7 d24 [200000,+20000) with the committed cost model")
    cyc, mul = gpu.sym_cases(7, 24, 200000, 20000)
    ref, pc0, pn0 = gpu.sym_solve(7, 24, cyc, mul, _opts(gpu, "resplit"), per_item=True)
    model = cost.CostModel(seed=0)
    model.load(os.path.join(ROOT, "map_folding_amd", "cost_default.npy"))
    pieces = cost.learned_split(model, 7, 24, cyc, mul, 12.0)
    assert len(pieces) > cyc.shape[0] and len(set(pieces.levels.tolist())) > 1
    f = pieces.frontier(7)
    for sched in ("resplit", "resplit_tight_nograph"):
        r, pc, pn = f.solve(_opts(gpu, sched), per_item=True)
        assert r.count * 49 == 4968634597136, sched
        assert np.array_equal(pc, pc0), sched
        assert int(pn.sum()) + pieces.host_nodes == ref.nodes, sched
    f.close()
    model.close()


@pytest.mark.parametrize("n,e,start,num,targets", [(8, 36, 120000000, 2000, (40, 47, 50)), (7, 24, 200000, 1000, (28, 30)),
                                                   (8, 28, 1000, 200, (36, 44))])
def test_deepened_frontier_keeps_every_count(gpu, n, e, start, num, targets):
    """The re-split schedule on a frontier deepened on the device to a target
    level (one counted get_good per expanded node) gives the per-case counts,
    per-case nodes and exact event counters of the undeepened solve, at any
    target, budget and round mode."""
    _say(f"deepening: {n}x{n} level {e} [{start},+{num}) targets {targets}")
    cyc, mul = gpu.sym_cases(n, e, start, num)
    f = gpu.SymFrontier(n, e, cyc, mul)
    ref, pc0, pn0 = f.solve(_opts(gpu, "resplit", deepen=-1), per_item=True)
    for target in targets:
        for sched in ("resplit", "resplit_tight", "resplit_tight_nograph"):
            r, pc, pn = f.solve(_opts(gpu, sched, deepen=target), per_item=True)
            assert (r.count, r.nodes, r.walks, r.steps, r.leaves) == (ref.count, ref.nodes, ref.walks, ref.steps,
                                                                      ref.leaves), (target, sched)
            assert np.array_equal(pc, pc0) and np.array_equal(pn, pn0), (target, sched)
    f.close()


def test_deepening_out_of_room_falls_back(gpu):
    """A target whose deepened frontier would not fit the device buffers is
    abandoned and the frontier solved as it is: the counts do not change."""
    _say("deepening fallback: 8x8 level 36 [0,+300) to level 61")
    cyc, mul = gpu.sym_cases(8, 36, 0, 300)
    ref, pc0, pn0 = gpu.sym_solve(8, 36, cyc, mul, _opts(gpu, "resplit", deepen=-1), per_item=True)
    r, pc, pn = gpu.sym_solve(8, 36, cyc, mul, _opts(gpu, "resplit", deepen=61), per_item=True)
    assert (r.count, r.nodes, r.walks, r.steps) == (ref.count, ref.nodes, ref.walks, ref.steps)
    assert np.array_equal(pc, pc0) and np.array_equal(pn, pn0)


def test_per_item_targets_out_of_room_fall_back_to_uniform(gpu):
    """Per-item targets that outgrow the deepening buffers give way to the
    uniform target (not to an undeepened solve): the first round then starts
    from the uniformly deepened frontier, and every count is unchanged."""
    _say("per-item deepening targets out of room: 8x8 level 36 [0,+300)")
    cyc, mul = gpu.sym_cases(8, 36, 0, 300)
    f = gpu.SymFrontier(8, 36, cyc, mul)
    ref, pc0, pn0 = f.solve(_opts(gpu, "resplit", deepen=-1), per_item=True)
    uni, _, _ = f.solve(_opts(gpu, "resplit", deepen=44), per_item=True)
    r, pc, pn = f.solve(_opts(gpu, "resplit", deepen=44, deepen_levels=np.full(300, 61, np.uint8)), per_item=True)
    assert (r.count, r.nodes, r.walks, r.steps) == (ref.count, ref.nodes, ref.walks, ref.steps)
    assert np.array_equal(pc, pc0) and np.array_equal(pn, pn0)
    assert ref.inputs == 300 and uni.inputs > 300 and r.inputs == uni.inputs
    f.close()


def test_per_item_deepening_targets_keep_counts(gpu):
    """Per-item deepening targets (as the cost model picks them), with and
    without an item order, give the counts of the undeepened solve."""
    _say("per-item deepening targets: 8x8 level 36 [120000000,+1500)")
    cyc, mul = gpu.sym_cases(8, 36, 120000000, 1500)
    f = gpu.SymFrontier(8, 36, cyc, mul)
    ref, pc0, pn0 = f.solve(_opts(gpu, "resplit", deepen=-1), per_item=True)
    rng = np.random.default_rng(3)
    lv = rng.integers(36, 52, size=1500).astype(np.uint8)
    for order in (None, rng.permutation(1500)):
        for sched in ("resplit", "resplit_tight"):
            r, pc, pn = f.solve(_opts(gpu, sched, deepen=44, deepen_levels=lv, order=order), per_item=True)
            assert (r.count, r.nodes, r.walks, r.steps) == (ref.count, ref.nodes, ref.walks, ref.steps), sched
            assert np.array_equal(pc, pc0) and np.array_equal(pn, pn0), sched
    with pytest.raises(gpu.MFError):
        f.solve(_opts(gpu, "resplit", deepen_levels=lv[:10]))
    f.close()


def test_two_lanes_per_subtree_keep_every_count(gpu):
    """Core 3 (a lane pair per subtree, the pair splitting each node's page
    walks) gives the per-case counts, per-case nodes and exact event counters
    of the one-lane kernel, deepened or not; it runs the re-split schedule
    only."""
    for n, e, start, num in ((7, 24, 200000, 3000), (8, 36, 120000000, 1500)):
        _say(f"two lanes per subtree: {n}x{n} level {e} [{start},+{num})")
        cyc, mul = gpu.sym_cases(n, e, start, num)
        f = gpu.SymFrontier(n, e, cyc, mul)
        ref, pc0, pn0 = f.solve(_opts(gpu, "resplit", deepen=-1), per_item=True)
        for sched, deepen in (("resplit_pair", -1), ("resplit_pair", 0), ("resplit_pair_tight", -1)):
            r, pc, pn = f.solve(_opts(gpu, sched, deepen=deepen), per_item=True)
            assert (r.count, r.nodes, r.walks, r.steps, r.leaves) == (ref.count, ref.nodes, ref.walks, ref.steps,
                                                                      ref.leaves), (sched, deepen)
            assert np.array_equal(pc, pc0) and np.array_equal(pn, pn0), (sched, deepen)
        with pytest.raises(gpu.MFError):
            f.solve(_opts(gpu, "static", core=3))
        f.close()
