#!/usr/bin/env python3
"""Benchmark: symmetry-reduced map-folding DFS on one MI355X.

Workload: the 8x8 depth-36 expanded-case chunk [start, start+50000) of the
sym_fold case stream (MF_EXPANDED_N=36), default start 120,000,000.  A step is
one pass of the hot path over the chunk: every case resumed and counted to
completion on the GPU, cases resident in HBM before the timed region.  The
chunk total is checked bit-exactly (against the published/reference golden
for the default chunk) on every timed step.

Prints ONE JSON line.  `value` is DFS nodes/s (node = one get_good call) of
the batched persistent re-split schedule fed by the learned cost model: the
model is trained on the GPU (pairs logged by the GPU DFS on case windows
disjoint from the timed chunk), predicts every case, and cases predicted large
are split into subtrees that are queued largest size class first, then by
level (outside the timed region, like the case generation; both times are reported).  The same
schedule without the model, and the plain static-frontier kernel (one lane
per case, one launch) on the chunk's first --plain-cases cases (it is bounded
by its longest case: minutes on the whole chunk), are timed beside it, the
latter also on learned and lookahead-proxy splits.  Other configurations:
--workload sym7 | sym9 | lunnon6 | rect64.

    python bench.py [--steps K] [--warmup W] [--workload W] [--start S] [--cases C]
                    [--no-plain] [--no-cpu] [--cpu-cases M] [--plain-cases P]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

GOLDEN_CHUNK = (120000000, 50000, 26687875796992)   # published/reference chunk total (x n^2)
# Symmetry-DFS workloads: (n, expanded level, default start, cases, golden answer of that window).
SYM_WORKLOADS = {
    "sym8": (8, 36, 120000000, 50000, 26687875796992),
    "sym7": (7, 24, 200000, 20000, 4968634597136),
    # 9x9 (sym_fold9: 5x5 center seeds, u128 masks): MF_EXPANDED_N=44 window [0, 2000)
    "sym9": (9, 44, 0, 2000, None),
}
# Default CPU-baseline samples (cases) per workload: about 10-30 s of host work.
CPU_CASES = {"sym8": 4000, "sym7": 4000, "sym9": 2000, "lunnon6": 4000, "rect64": 4000}
# Default sample of the plain static kernel (bounded by its longest case).
PLAIN_CASES = {"sym8": 300, "sym7": 300, "sym9": 100}


def sym_golden(workload, start, cases):
    """Golden answer of a sym workload window, or None."""
    n, level, start0, cases0, golden = SYM_WORKLOADS[workload]
    if (start, cases) != (start0, cases0):
        return None
    if golden is None:   # the 9x9 windows of the reference binary
        with open(os.path.join(ROOT, "tests", "golden", "sym9.json")) as f:
            for w in json.load(f)["windows"]:
                if (w["n"], w["expanded"], w["start"], w["max"]) == (n, level, start, cases):
                    return w["answer"]
    return golden


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _traffic(args):
    """HBM bytes per launch of the dominant kernel, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes over this same command (scaled by
    the copy-kernel calibration); None for another workload."""
    if args.workload != "sym8" or (args.start, args.cases) != GOLDEN_CHUNK[:2]:
        return None
    path = os.path.join(ROOT, "profiles", "sym_persistent_8x8_d36_chunk_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["hbm_bytes_per_launch"]


# CPU baseline with the reference's own solver (built from its sources into
# oracle/_ref by `make -C oracle ref`): a window of the same case stream, the
# binary's threads = the host cores.  Windows are chosen for 10-30 s of CPU.
REF_WINDOWS = {"sym8": ("sym_fold", 0, 12000), "sym7": ("sym_fold", 200000, 20000), "sym9": ("sym_fold9", 0, 2000)}


def reference_cpu(args, mf, n, level, threads):
    """Time the reference binary on a bounded window; its answer must equal the
    GPU's count of the same window, whose exact node count gives the rate.
    None if the binary is not there."""
    import subprocess
    exe_name, start, count = REF_WINDOWS[args.workload]
    exe = os.path.join(ROOT, "oracle", "_ref", exe_name)
    if not os.path.exists(exe):
        return None
    env = {k: v for k, v in os.environ.items() if not k.startswith("MF_")}
    env.update({"MF_STREAM_CASES": "1", "MF_EXPANDED_N": str(level)})
    t1 = time.perf_counter()
    r = subprocess.run([exe, str(n), str(threads), str(count), str(start)], capture_output=True, text=True, env=env,
                       timeout=600)
    cs = time.perf_counter() - t1
    if r.returncode != 0:
        raise SystemExit(f"reference {exe_name} failed: {r.stderr[-300:]}")
    answer = int(r.stdout.split()[0])
    cyc, mul = mf.sym_cases(n, level, start, count)
    res, _, _ = mf.sym_solve(n, level, cyc, mul, mf.SolveOptions(count_events=True))
    if res.count * n * n != answer:
        raise SystemExit(f"reference {exe_name} answered {answer}, the GPU {res.count * n * n}")
    return {"value": res.nodes / cs, "unit": "nodes/s", "cores": threads, "kind": "reference",
            "sample": f"oracle/_ref/{exe_name} (reference source built here) {n} {threads} {count} {start} with "
                      f"MF_STREAM_CASES=1 MF_EXPANDED_N={level}: {count} cases ({res.nodes} nodes, answer {answer} "
                      f"= the GPU's), {cs:.1f} s wall including its seed and stream setup"}


def _run_ref(exe_name, argv):
    """(stdout, seconds) of a reference binary in oracle/_ref, or None if it is
    not built."""
    import subprocess
    exe = os.path.join(ROOT, "oracle", "_ref", exe_name)
    if not os.path.exists(exe):
        return None
    t1 = time.perf_counter()
    r = subprocess.run([exe] + [str(a) for a in argv], capture_output=True, text=True, timeout=600)
    cs = time.perf_counter() - t1
    if r.returncode != 0:
        raise SystemExit(f"reference {exe_name} failed: {r.stderr[-300:]}")
    return r.stdout, cs


def reference_lunnon6(want, gpu_inserts):
    """The reference's threaded solver on the whole 6x6 map from the same
    depth-18 frontier (`mf_fast 6 18 <threads>`, one state per task).  It runs
    the same DFS, so its inserts below the frontier are the GPU's exact count;
    the frontier collection above it (112,024 states) is not counted."""
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    got = _run_ref("mf_fast", [6, 18, threads])
    if got is None:
        return None
    out, cs = got
    if int(out.split()[0]) != want:
        raise SystemExit(f"reference mf_fast answered {out.split()[0]}, expected {want}")
    return {"value": gpu_inserts / cs, "unit": "inserts/s", "cores": threads, "kind": "reference",
            "sample": f"oracle/_ref/mf_fast 6 18 {threads} (reference source built here): the whole 6x6 map, "
                      f"answer {want} = the GPU's, {gpu_inserts} inserts below the frontier, {cs:.1f} s wall"}


def reference_rects(maps, answers, inserts):
    """The reference's single-threaded mf.c on the maps of the port's sample
    (`mf p q` each); the inserts are the port's exact count of the same DFS."""
    total = 0.0
    for (p, q), want in zip(maps, answers):
        got = _run_ref("mf", [p, q])
        if got is None:
            return None
        out, cs = got
        if int(out.split()[0]) != want:
            raise SystemExit(f"reference mf {p} {q} answered {out.split()[0]}, the oracle {want}")
        total += cs
    return {"value": inserts / total, "unit": "inserts/s", "cores": 1, "kind": "reference",
            "sample": f"oracle/_ref/mf p q (reference source built here) on the maps {maps}, answers = the "
                      f"oracle's, {inserts} inserts, {total:.1f} s wall"}


def rect_shapes(seed, k=64):
    """The batched-instances draw: p uniform in {2,3,4}, then q uniform among
    the q with 12 <= p*q <= 30."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(k):
        p = int(rng.integers(2, 5))
        qs = [q for q in range(1, 31) if 12 <= p * q <= 30]
        out.append((p, int(rng.choice(qs))))
    return out


def rect_depth(p, q):
    """Frontier depth per map: deep enough that every map splits into many
    states, shallow enough that the host collector stays cheap."""
    return max(2, min(p * q - 6, (2 * p * q) // 3))


def roof_peak(rl, cus):
    """The int32 VALU roof in T ops/s: the larger of valu_peak's reading on
    this device (taken after the timed steps) and the best recorded reading,
    so that a box that reads low does not inflate the fraction.  The live
    reading and the spec-sheet figure are reported beside it."""
    live, source = rl.measured_int32_peak(ROOT)
    ops = max(live, rl.INT32_MEASURED_OPS)
    spec = rl.int32_peak_ops(cus or 256) / 1e12
    log(f"int32 VALU peak: {ops:.4e} ops/s (this box: {live:.4e}, {source}); CUs x 64 x 2.4 GHz = {spec*1e12:.4e}")
    return ops / 1e12, {"peak_source": f"max(this box's reading, best recorded {rl.INT32_MEASURED_OPS:.3e}); {source}",
                        "peak_live": live / 1e12, "peak_cus_x64_x2p4ghz": spec,
                        "peak_published_fp32_lane_fma": rl.PUBLISHED_FP32_LANE_OPS / 1e12}


def run_lunnon(args, mf, rl):
    """lunnon6: 6x6 map, frontier at depth 18 (112,024 states), resumed on the
    GPU; rect64: 64 seeded rectangles, each with its own frontier, all in one
    device queue.  Totals checked against the golden table."""
    import numpy as np
    with open(os.path.join(ROOT, "tests", "golden", "golden.json")) as f:
        golden = json.load(f)
    mf.init(0)
    cus = mf.device_cus()
    batched = mf.SolveOptions(schedule=mf.MF_SCHED_RESPLIT, count_events=False, budget=args.budget,
                              tail=args.tail, waves_per_cu=args.waves_per_cu, core=args.core)
    events = mf.SolveOptions(schedule=mf.MF_SCHED_RESPLIT, count_events=True, core=args.core)
    tuned = None
    t0 = time.time()
    if args.workload == "lunnon6":
        dims, depth = [6, 6], 18
        states = mf.lunnon_frontier(dims, depth)
        front = mf.LunnonFrontier(dims, depth, states)
        solve = front.solve
        want = golden["square_totals"]["6"]
        label = f"6x6 Lunnon DFS, partition depth 18 ({states.shape[0]} states)"
        items = states.shape[0]
        if not args.no_tune and not (args.budget or args.tail or args.waves_per_cu):
            from map_folding_amd import autotune
            t1 = time.time()
            cfg = autotune.tune_resplit(front, budgets=(32, 64), tail_mult=(1,), waves=(16, 32), log=log)
            batched = mf.SolveOptions(schedule=mf.MF_SCHED_RESPLIT, count_events=False, budget=cfg["budget"],
                                      tail=cfg["tail"], waves_per_cu=cfg["waves_per_cu"], core=args.core)
            tuned = {k: cfg[k] for k in ("budget", "tail", "waves_per_cu", "kernel_ms")}
            tuned.update({"trials": len(cfg["trials"]), "tune_s": time.time() - t1})
            log(f"autotune: budget {cfg['budget']} tail {cfg['tail']} waves {cfg['waves_per_cu']} "
                f"({cfg['kernel_ms']:.1f} ms) in {tuned['tune_s']:.1f}s")
    else:
        shapes = rect_shapes(args.seed)
        depths = [rect_depth(p, q) for p, q in shapes]
        want_maps = [golden["rectangles"][f"{p}x{q}"] for p, q in shapes]
        want = sum(want_maps)
        if not args.no_tune:
            from map_folding_amd import autotune
            t1 = time.time()
            cfg = autotune.tune_lunnon_batch(shapes, targets=(256, 1024, 4096, 16384), budgets=(32, 64, 128),
                                             waves=(16, 32), log=log, extra_depths=[depths])
            depths = cfg["depths"]
            batched = mf.SolveOptions(schedule=mf.MF_SCHED_RESPLIT, count_events=False, budget=cfg["budget"],
                                      waves_per_cu=cfg["waves_per_cu"], core=args.core)
            tuned = {"depths": depths, "budget": cfg["budget"], "waves_per_cu": cfg["waves_per_cu"],
                     "kernel_ms": cfg["kernel_ms"], "trials": len(cfg["trials"]), "tune_s": time.time() - t1,
                     "default_rule_depths": [rect_depth(p, q) for p, q in shapes]}
            log(f"autotune: depths {min(depths)}..{max(depths)} budget {cfg['budget']} waves "
                f"{cfg['waves_per_cu']} ({cfg['kernel_ms']:.1f} ms) in {tuned['tune_s']:.1f}s")
        t0 = time.time()
        batch = mf.LunnonBatchFrontier(shapes, depths)

        def solve(opt):
            totals, res = batch.solve_maps(opt)
            if totals != want_maps:
                raise SystemExit("a map total disagreed with the golden table")
            return res, None, None
        label = f"64 rectangles (seed {args.seed}), per-map frontier depths {min(depths)}..{max(depths)}"
        items = len(shapes)
    gen_s = time.time() - t0
    res, _, _ = solve(events)
    if res.count != want:
        raise SystemExit(f"total {res.count} != golden {want}")
    log(f"{label}: counting run total={res.count} inserts={res.nodes} gapcomps={res.walks} steps={res.steps}")
    for _ in range(args.warmup):
        solve(batched)
    wall, kern = [], []
    for k in range(args.steps):
        t1 = time.perf_counter()
        r, _, _ = solve(batched)
        wall.append(time.perf_counter() - t1)
        kern.append(r.kernel_ms / 1e3)
        if r.count != want:
            raise SystemExit("a timed step disagreed")
        log(f"step {k}: wall {wall[-1]*1e3:.1f} ms, kernels {r.kernel_ms:.1f} ms, rounds {r.rounds}")
    ms = 1e3 * sum(wall) / len(wall)
    ops = rl.lunnon_ops(res.walks, res.steps, res.cands, res.nodes, res.leaves)
    achieved = ops / (sum(kern) / len(kern)) / 1e12
    peak, peak_notes = roof_peak(rl, cus)
    plain = None
    if not args.no_plain and args.workload == "lunnon6":
        t1 = time.perf_counter()
        rp, _, _ = solve(mf.SolveOptions(schedule=mf.MF_SCHED_STATIC, count_events=False, core=args.core))
        pw = time.perf_counter() - t1
        if rp.count != want:
            raise SystemExit("static schedule disagreed")
        plain = {"nodes_per_s": res.nodes / pw, "ms": pw * 1e3, "speedup_batched_vs_plain": pw * 1e3 / ms}
        log(f"plain static: {pw*1e3:.1f} ms")
    cpu = cpu_port = None
    if not args.no_cpu:
        from oracle import oracle
        t1 = time.perf_counter()
        if args.workload == "lunnon6":
            m = min(args.cpu_cases, states.shape[0])
            _, _, _, cst = oracle.lunnon_resume([6, 6], 18, states[:m])
            sample = f"first {m} frontier states, oracle C restatement of mf.c / solve_from_state, 1 thread"
        else:
            # Maps of the draw in order (those of at most 24 leaves, so that no
            # single map runs for minutes), until about 10 s of CPU work.
            small, answers, cst = [], [], {"inserts": 0}
            for p, q in shapes:
                if p * q > 24:
                    continue
                tot, st = oracle.lunnon_count([p, q])
                cst["inserts"] += st["inserts"]
                small.append((p, q))
                answers.append(tot)
                if time.perf_counter() - t1 > 10.0:
                    break
            sample = f"full counts of the draw's maps {small} (oracle C restatement of mf.c), 1 thread"
        cs = time.perf_counter() - t1
        cpu_port = {"value": cst["inserts"] / cs, "unit": "inserts/s", "cores": 1, "kind": "port",
                    "sample": f"{sample}, {cs:.1f} s"}
        log(f"cpu port: {cpu_port['value']:.3e} inserts/s")
        if args.workload == "lunnon6":
            cpu = reference_lunnon6(want, res.nodes)
        else:
            cpu = reference_rects(small, answers, cst["inserts"])
        if cpu is not None:
            log(f"cpu baseline (reference binary): {cpu['value']:.3e} inserts/s on {cpu['cores']} threads")
        else:
            cpu = cpu_port
    return {
        "metric": "DFS nodes/s & cases/s on 1 MI355X, batched+learned vs plain, bit-exact",
        "value": res.nodes / (ms / 1e3), "unit": "inserts/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic: deterministic Lunnon frontiers",
        "config": {"workload": label, "items": items, "total": str(res.count), "bit_exact_vs_golden": True},
        "autotune": tuned,
        "items_per_s": items / (ms / 1e3), "nodes": res.nodes, "gapcomps": res.walks, "walk_steps": res.steps,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TOP/s", "frac": achieved / peak,
                     "traffic": None, "ops_per_step": ops, **peak_notes,
                     "ops_model": "10*gapcomps + 8*walk_steps + 6*candidates + 8*inserts + 2*completions"},
        "plain": plain, "cpu_baseline": cpu, "cpu_port": cpu_port, "frontier_gen_s": gen_s,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="sym8", choices=["sym8", "sym7", "sym9", "lunnon6", "rect64"],
                    help="sym8 (default): 8x8 depth-36 50k-case chunk; sym7: 7x7 depth-24 20k cases; "
                         "sym9: 9x9 depth-44 2000 cases (no learned model: its features cover n <= 8); "
                         "lunnon6: 6x6 Lunnon frontier at depth 18; rect64: 64 seeded rectangles in one queue")
    ap.add_argument("--start", type=int, default=None)
    ap.add_argument("--cases", type=int, default=None)
    ap.add_argument("--no-plain", action="store_true", help="skip timing the plain static-frontier kernel")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-cases", type=int, default=None)
    ap.add_argument("--plain-cases", type=int, default=None)
    ap.add_argument("--seed", type=int, default=0, help="rect64: seed of the shape draw")
    ap.add_argument("--budget", type=int, default=0)
    ap.add_argument("--tail", type=int, default=0)
    ap.add_argument("--waves-per-cu", type=int, default=0)
    ap.add_argument("--core", type=int, default=0,
                    help="sym DFS stack: 0 default, 1 one u64 mask per level, 2 compact 16-bit entries")
    ap.add_argument("--no-learned", action="store_true",
                    help="sym: time the persistent schedule without the learned pre-split")
    ap.add_argument("--cost-params", default="",
                    help="sym: load cost-model parameters (.npy) instead of training them on the GPU")
    ap.add_argument("--split-threshold", type=float, default=15.0,
                    help="sym: split cases whose predicted log2(1+nodes) exceeds this")
    ap.add_argument("--no-compare", action="store_true",
                    help="sym: time only the batched+learned schedule (no no-model step, no plain kernel); "
                         "used for profiling runs")
    ap.add_argument("--no-graph", action="store_true",
                    help="host-driven re-split rounds instead of the captured HIP graph of rounds")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (re-split budget/waves; rect64: per-map depths too)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on exactly one GPU: this workload has no multi-GPU path. A chunk would "
              "shard as independent case ranges, one per GPU, whose exact counts add (one 16-byte "
              "u128 sum per rank).", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import map_folding_amd as mf
    from map_folding_amd import roofline as rl

    if args.cpu_cases is None:
        args.cpu_cases = CPU_CASES[args.workload]
    if args.workload in ("lunnon6", "rect64"):
        print(json.dumps(run_lunnon(args, mf, rl)), flush=True)
        return
    n, level, start0, cases0, _ = SYM_WORKLOADS[args.workload]
    if args.start is None:
        args.start = start0
    if args.cases is None:
        args.cases = cases0
    if args.plain_cases is None:
        args.plain_cases = PLAIN_CASES[args.workload]
    if n > 8 and not args.no_learned:
        log("9x9: the cost model's features cover n <= 8; timing the re-split schedule without it")
        args.no_learned = True
    mf.init(0)
    cus = mf.device_cus()
    t0 = time.time()
    cyc, mul = mf.sym_cases(n, level, args.start, args.cases)
    gen_s = time.time() - t0
    log(f"frontier: {cyc.shape[0]} cases from {args.start} generated in {gen_s:.2f}s (host)")
    front = mf.SymFrontier(n, level, cyc, mul)

    graph = 2 if args.no_graph else 0
    batched = mf.SolveOptions(schedule=mf.MF_SCHED_RESPLIT, count_events=False, budget=args.budget,
                              tail=args.tail, waves_per_cu=args.waves_per_cu, core=args.core, graph=graph)
    # Exact event counts (schedule-independent) for the algorithmic-work figure.
    res, _, case_nodes = front.solve(mf.SolveOptions(schedule=mf.MF_SCHED_RESPLIT, count_events=True,
                                                     core=args.core, graph=graph), per_item=True)
    nodes, walks, steps, leaves = res.nodes, res.walks, res.steps, res.leaves
    total = res.count * n * n
    golden = sym_golden(args.workload, args.start, args.cases)
    if golden is not None and total != golden:
        raise SystemExit(f"chunk total {total} != golden {golden}")
    log(f"counting run: total={total} nodes={nodes} walks={walks} steps={steps} rounds={res.rounds} "
        f"kernel_ms={res.kernel_ms:.1f}")

    # The persistent schedule without the learned part (one step, for reference).
    no_model = None
    if not args.no_compare:
        r0, _, _ = front.solve(batched)
        no_model = {"ms": r0.total_ms, "kernel_ms": r0.kernel_ms, "rounds": r0.rounds,
                    "nodes_per_s": nodes / (r0.total_ms / 1e3)}
        log(f"persistent, no model: {r0.total_ms:.1f} ms (kernels {r0.kernel_ms:.1f} ms, {r0.rounds} rounds)")

    # Learned part: train the cost model on the GPU (pairs logged by the GPU DFS
    # on windows disjoint from this chunk), predict every case, split the ones
    # predicted large into subtrees, lay the pieces out largest size class first.
    learned = None
    timed, timed_opts, gpu_nodes = front, batched, nodes
    n_inputs = cyc.shape[0]
    if not args.no_learned:
        from map_folding_amd import cost
        t1 = time.time()
        if args.cost_params:
            model = cost.CostModel(seed=0)
            model.load(args.cost_params)
            pairs, losses = None, []
        else:
            model, losses, pairs = cost.train_default(seed=0)
        train_s = time.time() - t1
        pred = model.predict(cost.features_sym(n, level, cyc))
        y = cost.node_target(case_nodes)
        proxy = cost.estimate_work(n, level, cyc, 8)
        t1 = time.time()
        pieces = cost.learned_split(model, n, level, cyc, mul, args.split_threshold)
        split_s = time.time() - t1
        timed = pieces.frontier(n)
        timed_opts = mf.SolveOptions(schedule=mf.MF_SCHED_RESPLIT, count_events=False, budget=args.budget,
                                     tail=args.tail, waves_per_cu=args.waves_per_cu, core=args.core,
                                     graph=2 if args.no_graph else 0)
        rl_ev, pc_l, _ = timed.solve(mf.SolveOptions(schedule=mf.MF_SCHED_RESPLIT, count_events=True,
                                                     core=args.core, graph=2 if args.no_graph else 0),
                                     per_item=True)
        if rl_ev.count * n * n != total or rl_ev.nodes + pieces.host_nodes != nodes:
            raise SystemExit("learned pieces disagree with the case frontier")
        gpu_nodes = rl_ev.nodes
        walks, steps = rl_ev.walks, rl_ev.steps
        n_inputs = len(pieces)
        learned = {"pairs": pairs, "train_s": train_s, "final_loss": float(losses[-1]) if len(losses) else None,
                   "rmse_log2_nodes": float(np.sqrt(np.mean((pred - y) ** 2))),
                   "spearman": cost.spearman(pred, case_nodes),
                   "spearman_lookahead8_proxy": cost.spearman(proxy, case_nodes),
                   "split_threshold": args.split_threshold, "pieces": len(pieces), "host_nodes": pieces.host_nodes,
                   "split_s": split_s}
        log(f"cost model: {train_s:.1f}s to train, rmse {learned['rmse_log2_nodes']:.2f} (log2 nodes), "
            f"spearman {learned['spearman']:.3f} (lookahead-8 proxy {learned['spearman_lookahead8_proxy']:.3f}); "
            f"{len(pieces)} pieces in {split_s:.2f}s")

    tuned = None
    if not args.no_tune and not (args.budget or args.tail or args.waves_per_cu):
        from map_folding_amd import autotune
        t1 = time.time()
        cfg = autotune.tune_resplit(timed, budgets=(32, 64, 128), tail_mult=(1,), waves=(20, 32), log=log)
        timed_opts = mf.SolveOptions(schedule=mf.MF_SCHED_RESPLIT, count_events=False, budget=cfg["budget"],
                                     tail=cfg["tail"], waves_per_cu=cfg["waves_per_cu"], core=args.core,
                                     graph=graph)
        tuned = {k: cfg[k] for k in ("budget", "tail", "waves_per_cu", "kernel_ms")}
        tuned.update({"trials": len(cfg["trials"]), "tune_s": time.time() - t1})
        log(f"autotune: budget {cfg['budget']} tail {cfg['tail']} waves {cfg['waves_per_cu']} "
            f"({cfg['kernel_ms']:.1f} ms) in {tuned['tune_s']:.1f}s")

    for w in range(args.warmup):
        r, _, _ = timed.solve(timed_opts)
        log(f"warmup {w}: {r.total_ms:.1f} ms")

    wall, kern, span, rounds, items = [], [], [], [], []
    for k in range(args.steps):
        t1 = time.perf_counter()
        r, _, _ = timed.solve(timed_opts)
        wall.append(time.perf_counter() - t1)
        if r.count * n * n != total or r.nodes != gpu_nodes:
            raise SystemExit("a timed step disagreed with the counting run")
        kern.append(r.kernel_ms / 1e3)
        span.append(r.span_ms / 1e3)
        rounds.append(r.rounds)
        items.append(r.items)
        log(f"step {k}: wall {wall[-1]*1e3:.1f} ms, kernels {r.kernel_ms:.1f} ms, rounds {r.rounds}, "
            f"items {r.items}")

    ms_per_step = 1e3 * sum(wall) / len(wall)
    item_bytes = 80 if n * n <= 64 else 128
    node_rate = gpu_nodes / (ms_per_step / 1e3)
    case_rate = cyc.shape[0] / (ms_per_step / 1e3)
    kernel_s = sum(kern) / len(kern)
    ops = rl.sym_ops(gpu_nodes, walks, steps)
    achieved = ops / kernel_s / 1e12
    peak, peak_notes = roof_peak(rl, cus)

    # The plain static-frontier kernel (one lane per case, one launch) is
    # bounded by its single longest case; on the whole chunk it runs for
    # minutes, so both schedules are timed on the chunk's first --plain-cases
    # cases, resident the same way.
    plain = None
    if not args.no_plain and not args.no_compare:
        m = min(args.plain_cases, cyc.shape[0])
        sub = mf.SymFrontier(n, level, cyc[:m], mul[:m])
        rs, _, _ = sub.solve(mf.SolveOptions(schedule=mf.MF_SCHED_RESPLIT, count_events=True, core=args.core))
        sub_ops = rl.sym_ops(rs.nodes, rs.walks, rs.steps)
        t1 = time.perf_counter()
        rb, _, _ = sub.solve(batched)
        bw = time.perf_counter() - t1
        log(f"plain static-frontier kernel on the first {m} cases ({rs.nodes} nodes): running")
        t1 = time.perf_counter()
        rp, _, _ = sub.solve(mf.SolveOptions(schedule=mf.MF_SCHED_STATIC, count_events=False, core=args.core))
        pw = time.perf_counter() - t1
        if rp.count != rs.count or rb.count != rs.count:
            raise SystemExit("static and re-split schedules disagreed")
        plain = {"cases": m, "nodes": rs.nodes, "nodes_per_s": rs.nodes / pw, "ms": pw * 1e3,
                 "frac_of_int32_peak": sub_ops / (rp.kernel_ms / 1e3) / 1e12 / peak,
                 "batched_nodes_per_s_same_cases": rs.nodes / bw, "batched_ms_same_cases": bw * 1e3,
                 "speedup_batched_vs_plain": pw / bw}
        log(f"plain static: {pw*1e3:.1f} ms; batched on the same cases: {bw*1e3:.1f} ms")
        sub.close()
        if learned is not None:
            # The same static one-lane-per-item kernel on split pieces: learned
            # split versus the reference's lookahead proxy (estimate_work, depth 8).
            for name, pcs in (("learned_static", cost.learned_split(model, n, level, cyc[:m], mul[:m], 12.0)),
                              ("lookahead_static", cost.lookahead_split(n, level, cyc[:m], mul[:m], 5.0, depth=8))):
                fr = pcs.frontier(n)
                t1 = time.perf_counter()
                rq, _, _ = fr.solve(mf.SolveOptions(schedule=mf.MF_SCHED_STATIC, count_events=False,
                                                    core=args.core))
                qw = time.perf_counter() - t1
                if rq.count != rs.count:
                    raise SystemExit(f"{name} disagreed with the re-split schedule")
                plain[name] = {"pieces": len(pcs), "ms": qw * 1e3, "speedup_vs_plain": pw / qw}
                log(f"{name}: {len(pcs)} pieces, {qw*1e3:.1f} ms")
                fr.close()

    cpu = cpu_port = None
    if not args.no_cpu:
        from oracle import oracle
        m = min(args.cpu_cases, cyc.shape[0])
        threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        t1 = time.perf_counter()
        if n <= 8:
            _, _, csum, cst = oracle.sym_solve(n, level, cyc[:m], mul[:m], threads=threads)
            src = "sym_fold.cpp recurse_count"
        else:
            _, _, csum, cst = oracle.sym9_solve(n, level, cyc[:m], mul[:m], threads=threads)
            src = "sym_fold9.cpp recurse_count (u128 masks)"
        cs = time.perf_counter() - t1
        cpu_port = {"value": cst["nodes"] / cs, "unit": "nodes/s", "cores": threads, "kind": "port",
                    "sample": f"first {m} cases of the chunk ({cst['nodes']} nodes), oracle C restatement of "
                              f"{src} on {threads} threads, {cs:.1f} s"}
        log(f"cpu port: {cpu_port['value']:.3e} nodes/s on {threads} threads ({cs:.1f}s)")
        cpu = reference_cpu(args, mf, n, level, threads)
        if cpu is not None:
            log(f"cpu baseline (reference binary): {cpu['value']:.3e} nodes/s on {threads} threads")
        else:
            cpu = cpu_port

    out = {
        "metric": "DFS nodes/s & cases/s on 1 MI355X, batched+learned vs plain, bit-exact",
        "value": node_rate,
        "unit": "nodes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic: deterministic sym_fold case stream (no dataset)",
        "config": {"workload": f"{n}x{n} sym_fold MF_EXPANDED_N={level} chunk [{args.start},{args.start + cyc.shape[0]})",
                   "cases": int(cyc.shape[0]),
                   "schedule": "persistent re-split" + ("" if learned is None else " + learned pre-split/order"),
                   "chunk_total": str(total),
                   "bit_exact_vs_golden": golden is not None},
        "cases_per_s": case_rate,
        "nodes": nodes, "page_walks": walks, "walk_steps": steps, "leaves": leaves,
        "rounds": rounds[-1] if rounds else None,
        "kernel_ms_per_step": kernel_s * 1e3,
        # Every work item (80 B; 128 B for 9x9) is read once; every item but
        # the inputs was written once by the round that split it.
        "items_per_step": items[-1] if items else None,
        "algorithmic_bytes_per_launch": (item_bytes * (2 * items[-1] - n_inputs) / rounds[-1]) if items else None,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TOP/s",
                     "frac": achieved / peak, "traffic": _traffic(args),
                     "ops_per_step": ops,
                     "ops_model": "16*nodes + 6*page_walks + 8*walk_steps (int32), GPU events of the timed step",
                     **peak_notes},
        "no_model": no_model,
        "learned": learned,
        "autotune": tuned,
        "plain": plain,
        "speedup_vs_plain": plain["speedup_batched_vs_plain"] if plain else None,
        "cpu_baseline": cpu,
        "cpu_port": cpu_port,
        "frontier_gen_s": gen_s,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
