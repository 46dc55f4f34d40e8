#!/usr/bin/env python3
"""Benchmark: map-folding subtree DFS on one MI355X, every configuration of
the workload table in one run.

Headline workload: the 8x8 depth-36 expanded-case chunk [start, start+50000)
of the sym_fold case stream (MF_EXPANDED_N=36), default start 120,000,000.  A
step is one pass of the hot path over the chunk: every case resumed and
counted to completion on the GPU, cases resident in HBM before the timed
region, the chunk total checked bit-exactly against the golden on every step.

`value` is DFS nodes/s (node = one get_good call) of the batched persistent
re-split schedule with the learned schedule: the cost model (trained on the
GPU by the library's own kernels, on pairs logged by the GPU DFS from windows
disjoint from the chunk) predicts every case's subtree size, and the predicted
large cases are deepened on the device a few levels further than the rest
before the rounds.  The tuner picks the deepening target (partition depth),
the learned threshold and the re-split budget.  Beside it, in the same JSON
line:
  * no_model      the same schedule with the uniform deepening target;
  * end_to_end_ms generation + (learned: features and inference) + upload +
                  solve of one chunk, for both;
  * plain         the plain static-frontier kernel (one lane per case, one
                  launch) on the whole chunk, timed while the CPU baselines run
                  on the host (it is bounded by its largest case, minutes);
  * cpu_baseline  the reference's own CPU command built from its sources
                  (oracle/_ref) on the same chunk, on the host's cores;
  * configs       5x5 Lunnon depth 15, 6x6 Lunnon depth 18, the 64 seeded
                  rectangles in one queue, and a 20,000-case 7x7 window, each
                  with its time, rate, roofline fraction, bit-exactness and CPU
                  baseline.

    python bench.py [--steps K] [--warmup W] [--headline-only] [--no-plain] [--no-cpu]
    python bench.py --workload sym8|sym7|sym9|lunnon5|lunnon6|rect64 [...]   (one configuration)
"""
import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Symmetry-DFS workloads: (n, expanded level, default start, cases, golden answer of that window).
SYM_WORKLOADS = {
    "sym8": (8, 36, 120000000, 50000, 26687875796992),
    "sym7": (7, 24, 200000, 20000, 4968634597136),
    # 9x9 (sym_fold9: 5x5 center seeds, u128 masks): MF_EXPANDED_N=44 window [0, 2000)
    "sym9": (9, 44, 0, 2000, None),
}
# Deepening targets the tuner tries per square (the cases' level must be below).
DEEPENS = {8: (45, 46, 47, 48), 7: (28, 30, 32, 34), 9: ()}
# The reference CPU solver of each square for the CPU baselines, run on the
# very window the GPU times (the reference's own command for it).  The default
# run times the plain static kernel (minutes) while the host runs them.
REF_EXE = {7: "sym_fold", 8: "sym_fold", 9: "sym_fold9"}
# Oracle-port samples (cases): the first cases of the timed window.
PORT_CASES = {"sym8": 6000, "sym7": 6000, "sym9": 500}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_threads():
    return len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)


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


def roof_peak(rl, cus):
    """The int32 VALU roof in T ops/s: the larger of valu_peak's reading on
    this device and the best recorded reading (a box that reads low does not
    inflate the fraction).  The live reading and the spec-sheet figure are
    reported beside it."""
    live, source = rl.measured_int32_peak(ROOT)
    ops = max(live, rl.INT32_MEASURED_OPS)
    spec = rl.int32_peak_ops(cus or 256) / 1e12
    log(f"int32 VALU peak: {ops:.4e} ops/s (this box: {live:.4e}, {source}); CUs x 64 x 2.4 GHz = {spec*1e12:.4e}")
    return ops / 1e12, {"peak_source": f"max(this box's reading, best recorded {rl.INT32_MEASURED_OPS:.3e}); {source}",
                        "peak_live": live / 1e12, "peak_cus_x64_x2p4ghz": spec,
                        "peak_published_fp32_lane_fma": rl.PUBLISHED_FP32_LANE_OPS / 1e12}


def _traffic(workload, start, cases):
    """HBM bytes per launch of the dominant kernel, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes over the headline chunk (scaled
    by the copy-kernel calibration); None for another workload."""
    if workload != "sym8" or (start, cases) != SYM_WORKLOADS["sym8"][2:4]:
        return None
    path = os.path.join(ROOT, "profiles", "sym_persistent_8x8_d36_chunk_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["hbm_bytes_per_launch"]


def _run_ref(exe_name, argv, env_extra=None):
    """(stdout, seconds) of a reference binary in oracle/_ref, or None if it is
    not built."""
    import subprocess
    exe = os.path.join(ROOT, "oracle", "_ref", exe_name)
    if not os.path.exists(exe):
        return None
    env = {k: v for k, v in os.environ.items() if not k.startswith("MF_")}
    env.update(env_extra or {})
    t1 = time.perf_counter()
    r = subprocess.run([exe] + [str(a) for a in argv], capture_output=True, text=True, env=env, timeout=900)
    cs = time.perf_counter() - t1
    if r.returncode != 0:
        raise SystemExit(f"reference {exe_name} failed: {r.stderr[-300:]}")
    return r.stdout, cs


# ----------------------------------------------------------------------------- CPU baselines
# Each job runs on the host only (reference binaries or the oracle), so the
# jobs can run while the plain static kernel occupies the GPU.  The GPU's own
# answers for the same windows are computed beforehand.


def sym_ref_job(exe_name, n, level, start, count, gpu_answer, gpu_nodes, threads):
    """The reference's own command on the window [start, start+count):
    `MF_STREAM_CASES=1 MF_EXPANDED_N=level sym_fold n threads count start`.
    Its producer enumerates every case before `start` on one thread first, so
    the same command with one case (setup + skip) is timed as well and the
    rate is taken over the difference."""
    env = {"MF_STREAM_CASES": "1", "MF_EXPANDED_N": str(level)}

    def job():
        got = _run_ref(exe_name, [n, threads, 1, start], env)
        if got is None:
            return None
        skip_s = got[1]
        out, cs = _run_ref(exe_name, [n, threads, count, start], env)
        answer = int(out.split()[0])
        if answer != gpu_answer:
            raise SystemExit(f"reference {exe_name} answered {answer}, the GPU {gpu_answer}")
        solve_s = max(cs - skip_s, 1e-3)
        return {"value": gpu_nodes / solve_s, "unit": "nodes/s", "cores": threads, "kind": "reference",
                "wall_s": cs, "setup_and_skip_s": skip_s, "nodes_per_s_incl_skip": gpu_nodes / cs,
                "sample": f"oracle/_ref/{exe_name} (reference source built here) {n} {threads} {count} {start} "
                          f"with MF_STREAM_CASES=1 MF_EXPANDED_N={level}: the same {count} cases as the GPU "
                          f"({gpu_nodes} nodes, answer {answer} = the GPU's), {cs:.1f} s wall, of which "
                          f"{skip_s:.1f} s is its seed setup and one-thread enumeration of the {start} cases "
                          f"before the window (timed alone with max_cases 1); rate over the remaining "
                          f"{solve_s:.1f} s"}
    return job


def sym_port_job(n, level, cyc, mul, threads):
    def job():
        from oracle import oracle
        t1 = time.perf_counter()
        if n <= 8:
            _, _, _, st = oracle.sym_solve(n, level, cyc, mul, threads=threads)
            src = "sym_fold.cpp recurse_count"
        else:
            _, _, _, st = oracle.sym9_solve(n, level, cyc, mul, threads=threads)
            src = "sym_fold9.cpp recurse_count (u128 masks)"
        cs = time.perf_counter() - t1
        return {"value": st["nodes"] / cs, "unit": "nodes/s", "cores": threads, "kind": "port",
                "cases": int(cyc.shape[0]), "nodes": int(st["nodes"]), "seconds": cs, "source": src,
                "sample": f"first {cyc.shape[0]} cases of the window ({st['nodes']} nodes), oracle C restatement of "
                          f"{src} on {threads} threads, {cs:.1f} s"}
    return job


def extend_port(rec, more):
    """Fold the next block of the window into an oracle-port sample."""
    rec["cases"] += more["cases"]
    rec["nodes"] += more["nodes"]
    rec["seconds"] += more["seconds"]
    rec["value"] = rec["nodes"] / rec["seconds"]
    rec["sample"] = (f"first {rec['cases']} cases of the window ({rec['nodes']} nodes), oracle C restatement of "
                     f"{rec['source']} on {rec['cores']} threads, {rec['seconds']:.1f} s")
    return rec


# ----------------------------------------------------------------------------- sym workloads


def sym_workload(args, mf, rl, cost, name, steps, warmup, model=None, pairs=None, losses=None, train_s=None,
                 headline=False):
    """One symmetry-DFS window: learned + deepened re-split schedule timed over
    `steps`, the same without the model, end-to-end times, and (headline) the
    pieces for the plain kernel and the CPU baselines.  Returns (record,
    extras) where extras holds what the caller needs later."""
    import numpy as np
    from map_folding_amd import autotune
    n, level, start0, cases0, _ = SYM_WORKLOADS[name]
    start = args.start if (headline and args.start is not None) else start0
    cases = args.cases if (headline and args.cases is not None) else cases0
    golden = sym_golden(name, start, cases)
    t0 = time.perf_counter()
    cyc, mul = mf.sym_cases(n, level, start, cases)
    gen_cold = time.perf_counter() - t0
    t0 = time.perf_counter()
    cyc, mul = mf.sym_cases(n, level, start, cases)
    gen_warm = time.perf_counter() - t0
    log(f"{name}: {cyc.shape[0]} cases from {start} generated in {gen_cold:.2f} s (index cold), "
        f"{gen_warm*1e3:.0f} ms (warm)")
    front = mf.SymFrontier(n, level, cyc, mul)
    # Exact event counts of the window (schedule-independent) and per-case nodes.
    res, _, case_nodes = front.solve(mf.SolveOptions(count_events=True), per_item=True)
    total = res.count * n * n
    if golden is not None and total != golden:
        raise SystemExit(f"{name}: total {total} != golden {golden}")
    log(f"{name}: counting run total={total} nodes={res.nodes} walks={res.walks} steps={res.steps} "
        f"rounds={res.rounds} kernel_ms={res.kernel_ms:.1f}")
    learned = None
    pred = None
    if model is not None and n <= 8:
        t1 = time.perf_counter()
        pred = model.predict_frontier(front)   # features and inference on the device
        pred_s = time.perf_counter() - t1
        y = cost.node_target(case_nodes)
        learned = {"pairs": pairs, "train_s": train_s, "final_loss": float(losses[-1]) if losses is not None and len(losses) else None,
                   "rmse_log2_nodes": float(np.sqrt(np.mean((pred - y) ** 2))),
                   "spearman": cost.spearman(pred, case_nodes),
                   "spearman_lookahead8_proxy": cost.spearman(cost.estimate_work(n, level, cyc, 8), case_nodes),
                   "predict_s": pred_s}
        log(f"{name}: cost model rmse {learned['rmse_log2_nodes']:.2f} (log2 nodes), spearman "
            f"{learned['spearman']:.3f} (lookahead-8 proxy {learned['spearman_lookahead8_proxy']:.3f}), "
            f"features + inference on the device {pred_s*1e3:.1f} ms")
    # Tuner: deepening target, learned threshold, budget.
    t1 = time.perf_counter()
    if args.no_tune or n > 8:
        cfg = {"deepen": args.deepen if args.deepen is not None else (DEEPENS[n][1] if len(DEEPENS[n]) > 1 else -1),
               "theta": args.theta if pred is not None else None, "cap": args.cap, "budget": args.budget,
               "lanes": 2 if (args.lanes == 2 and n <= 8) else 1, "trials": []}
    else:
        cfg = autotune.tune_sym(front, pred=pred, level=level, deepens=DEEPENS[n], log=log)
    cfg["tune_s"] = time.perf_counter() - t1
    levels = None if cfg["theta"] is None else cost.learned_levels(pred, cfg["deepen"], cfg["theta"], cfg["cap"])
    core = 3 if cfg.get("lanes", 1) == 2 else 0
    log(f"{name}: tuned deepen {cfg['deepen']} theta {cfg['theta']} cap {cfg['cap']} budget {cfg['budget']} "
        f"lanes per subtree {cfg.get('lanes', 1)} ({cfg['tune_s']:.1f} s)")
    tail = args.tail if (args.no_tune and args.tail > 0) else cfg["budget"]
    opts = mf.SolveOptions(count_events=False, budget=cfg["budget"], tail=tail, waves_per_cu=32,
                           deepen=cfg["deepen"], deepen_levels=levels, core=core)
    for w in range(warmup):
        r, _, _ = front.solve(opts)
        log(f"{name}: warmup {w}: {r.total_ms:.1f} ms")
    wall, kern, rounds, items, inputs = [], [], [], [], []
    for k in range(steps):
        t1 = time.perf_counter()
        r, _, _ = front.solve(opts)
        wall.append(time.perf_counter() - t1)
        if r.count * n * n != total or r.nodes != res.nodes:
            raise SystemExit(f"{name}: a timed step disagreed with the counting run")
        kern.append(r.kernel_ms / 1e3)
        rounds.append(r.rounds)
        items.append(r.items)
        inputs.append(r.inputs)
        log(f"{name}: step {k}: wall {wall[-1]*1e3:.1f} ms, kernels {r.kernel_ms:.1f} ms, rounds {r.rounds}, "
            f"items {r.items}")
    ms = 1e3 * sum(wall) / len(wall)
    kernel_s = sum(kern) / len(kern)
    ops = rl.sym_ops(res.nodes, res.walks, res.steps)
    # The same schedule without the model (uniform deepening) and without deepening.
    comp = {}
    variants = (("no_model", mf.SolveOptions(count_events=False, budget=cfg["budget"], tail=cfg["budget"],
                                             waves_per_cu=32, deepen=cfg["deepen"], core=core)),
                ("undeepened", mf.SolveOptions(count_events=False, budget=cfg["budget"], tail=cfg["budget"],
                                               waves_per_cu=32, deepen=-1, core=core)))
    for label, o in [] if args.no_compare else variants:
        if label == "no_model" and levels is None:
            comp[label] = {"ms": ms, "kernel_ms": kernel_s * 1e3, "rounds": rounds[-1], "same_as_learned": True}
            continue
        front.solve(o)
        tt = []
        for _ in range(2 if steps > 1 else 1):
            t1 = time.perf_counter()
            r0, _, _ = front.solve(o)
            tt.append((time.perf_counter() - t1, r0))
            if r0.count * n * n != total:
                raise SystemExit(f"{name}: {label} disagreed")
        w0, r0 = min(tt, key=lambda x: x[0])
        comp[label] = {"ms": w0 * 1e3, "kernel_ms": r0.kernel_ms, "rounds": r0.rounds,
                       "nodes_per_s": res.nodes / w0}
        log(f"{name}: {label}: {w0*1e3:.1f} ms (kernels {r0.kernel_ms:.1f} ms, {r0.rounds} rounds)")
    # End to end for one chunk: generation (warm index) + [features + inference]
    # + upload + solve.
    e2e = {}
    for label in [] if args.no_compare else (["learned", "no_model"] if levels is not None else ["no_model"]):
        t1 = time.perf_counter()
        c2, m2 = mf.sym_cases(n, level, start, cases)
        t_gen = time.perf_counter() - t1
        t2 = time.perf_counter()
        f2 = mf.SymFrontier(n, level, c2, m2)
        t_up = time.perf_counter() - t2
        lv = None
        t_model = 0.0
        if label == "learned":
            t2 = time.perf_counter()
            p2 = model.predict_frontier(f2)
            lv = cost.learned_levels(p2, cfg["deepen"], cfg["theta"], cfg["cap"])
            t_model = time.perf_counter() - t2
        t2 = time.perf_counter()
        r2, _, _ = f2.solve(mf.SolveOptions(count_events=False, budget=cfg["budget"], tail=cfg["budget"],
                                            waves_per_cu=32, deepen=cfg["deepen"], deepen_levels=lv, core=core))
        t_solve = time.perf_counter() - t2
        f2.close()
        if r2.count * n * n != total:
            raise SystemExit(f"{name}: end-to-end {label} disagreed")
        tot = time.perf_counter() - t1
        e2e[label] = {"ms": tot * 1e3, "gen_ms": t_gen * 1e3, "model_ms": t_model * 1e3, "upload_ms": t_up * 1e3,
                      "solve_ms": t_solve * 1e3}
        log(f"{name}: end to end ({label}): {tot*1e3:.1f} ms (generation {t_gen*1e3:.0f}, model {t_model*1e3:.0f}, "
            f"upload {t_up*1e3:.0f}, solve {t_solve*1e3:.1f})")
    rec = {"workload": f"{n}x{n} sym_fold MF_EXPANDED_N={level} window [{start},{start + cyc.shape[0]})",
           "cases": int(cyc.shape[0]), "total": str(total), "bit_exact_vs_golden": golden is not None,
           "ms_per_step": ms, "kernel_ms_per_step": kernel_s * 1e3, "nodes": res.nodes, "page_walks": res.walks,
           "walk_steps": res.steps, "leaves": res.leaves, "nodes_per_s": res.nodes / (ms / 1e3),
           "cases_per_s": cyc.shape[0] / (ms / 1e3), "rounds": rounds[-1] if rounds else None,
           "items_per_step": items[-1] if items else None,
           "first_round_items": inputs[-1] if inputs else None, "ops_per_step": ops,
           "achieved_tops": ops / kernel_s / 1e12,
           "schedule": "persistent re-split, device deepening" + ("" if levels is None else " + learned per-case depth"),
           "tuned": {**{k: cfg.get(k, 1) for k in ("deepen", "theta", "cap", "budget", "lanes", "tune_s")},
                     "trials": [[t["trial"], round(t["kernel_ms"], 2)] for t in cfg.get("trials", [])]},
           "learned": learned, **comp, "end_to_end_ms": e2e}
    return rec, {"n": n, "level": level, "start": start, "cyc": cyc, "mul": mul, "front": front, "total": total,
                 "nodes": res.nodes, "walks": res.walks, "steps": res.steps, "case_nodes": case_nodes,
                 "n_inputs": cyc.shape[0], "item_bytes": 80 if n * n <= 64 else 128}


# ----------------------------------------------------------------------------- Lunnon workloads


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


def lunnon_workload(args, mf, rl, name, steps, warmup):
    """lunnon5 / lunnon6: the square's frontier at depth 15 / 18 resumed on the
    GPU; rect64: the 64 seeded rectangles, each with its own frontier, in one
    device queue.  Returns (record, CPU jobs)."""
    with open(os.path.join(ROOT, "tests", "golden", "golden.json")) as f:
        golden = json.load(f)
    events = mf.SolveOptions(schedule=mf.MF_SCHED_RESPLIT, count_events=True)
    tuned = None
    jobs = []
    threads = host_threads()
    t0 = time.time()
    if name in ("lunnon5", "lunnon6"):
        side, depth = (5, 15) if name == "lunnon5" else (6, 18)
        dims = [side, side]
        states = mf.lunnon_frontier(dims, depth)
        front = mf.LunnonFrontier(dims, depth, states)
        want = golden["square_totals"][str(side)]
        label = f"{side}x{side} Lunnon DFS, partition depth {depth} ({states.shape[0]} states)"
        items = states.shape[0]
        batched = mf.SolveOptions(count_events=False, budget=args.budget, tail=args.budget, waves_per_cu=args.waves)
        if not args.no_tune:
            from map_folding_amd import autotune
            t1 = time.time()
            cfg = autotune.tune_resplit(front, budgets=(32, 64), tail_mult=(1,), waves=(8, 16, 32), log=log)
            batched = mf.SolveOptions(count_events=False, budget=cfg["budget"], tail=cfg["tail"],
                                      waves_per_cu=cfg["waves_per_cu"], core=cfg["core"])
            tuned = {k: cfg[k] for k in ("budget", "tail", "waves_per_cu", "core", "kernel_ms")}
            tuned.update({"trials": len(cfg["trials"]), "tune_s": time.time() - t1})

        def solve(opt):
            r, _, _ = front.solve(opt)
            return r
    else:
        shapes = rect_shapes(args.seed)
        depths = [rect_depth(p, q) for p, q in shapes]
        want_maps = [golden["rectangles"][f"{p}x{q}"] for p, q in shapes]
        want = sum(want_maps)
        batched = mf.SolveOptions(count_events=False, budget=args.budget, waves_per_cu=args.waves)
        if not args.no_tune:
            from map_folding_amd import autotune
            t1 = time.time()
            cfg = autotune.tune_lunnon_batch(shapes, targets=(256, 1024, 4096, 16384), budgets=(32, 64, 128),
                                             waves=(16, 32), log=log, extra_depths=[depths])
            depths = cfg["depths"]
            batched = mf.SolveOptions(count_events=False, budget=cfg["budget"], waves_per_cu=cfg["waves_per_cu"],
                                      core=cfg["core"])
            tuned = {"depths": depths, "budget": cfg["budget"], "waves_per_cu": cfg["waves_per_cu"],
                     "core": cfg["core"], "kernel_ms": cfg["kernel_ms"], "trials": len(cfg["trials"]),
                     "tune_s": time.time() - t1}
        batch = mf.LunnonBatchFrontier(shapes, depths)
        label = f"64 rectangles (seed {args.seed}), per-map frontier depths {min(depths)}..{max(depths)}"
        items = len(shapes)

        def solve(opt):
            totals, r = batch.solve_maps(opt)
            if totals != want_maps:
                raise SystemExit("a map total disagreed with the golden table")
            return r
    gen_s = time.time() - t0
    res = solve(events)
    if res.count != want:
        raise SystemExit(f"{name}: total {res.count} != golden {want}")
    log(f"{label}: counting run total={res.count} inserts={res.nodes} gapcomps={res.walks} steps={res.steps}")
    for _ in range(warmup):
        solve(batched)
    wall, kern = [], []
    for k in range(steps):
        t1 = time.perf_counter()
        r = solve(batched)
        wall.append(time.perf_counter() - t1)
        kern.append(r.kernel_ms / 1e3)
        if r.count != want:
            raise SystemExit(f"{name}: a timed step disagreed")
        log(f"{name}: step {k}: wall {wall[-1]*1e3:.1f} ms, kernels {r.kernel_ms:.1f} ms, rounds {r.rounds}")
    ms = 1e3 * sum(wall) / len(wall)
    ops = rl.lunnon_ops(res.walks, res.steps, res.cands, res.nodes, res.leaves)
    kernel_s = sum(kern) / len(kern)
    rec = {"workload": label, "items": items, "total": str(res.count), "bit_exact_vs_golden": True,
           "ms_per_step": ms, "kernel_ms_per_step": kernel_s * 1e3, "inserts": res.nodes, "gapcomps": res.walks,
           "walk_steps": res.steps, "inserts_per_s": res.nodes / (ms / 1e3), "items_per_s": items / (ms / 1e3),
           "ops_per_step": ops, "achieved_tops": ops / kernel_s / 1e12, "autotune": tuned, "frontier_gen_s": gen_s}
    # CPU baselines (host only).
    if name == "lunnon5":
        def ref_job():
            got = _run_ref("mf", [5, 5])
            if got is None:
                return None
            out, cs = got
            if int(out.split()[0]) != want:
                raise SystemExit("reference mf 5 5 disagreed")
            from oracle import oracle
            _, full = oracle.lunnon_count([5, 5])
            return {"value": full["inserts"] / cs, "unit": "inserts/s", "cores": 1, "kind": "reference",
                    "sample": f"oracle/_ref/mf 5 5 (reference source built here): the whole 5x5 map, answer {want} "
                              f"= the GPU's, {full['inserts']} inserts of the full tree (the oracle's exact count of "
                              f"the same DFS), {cs:.2f} s wall"}
        jobs.append(ref_job)
    elif name == "lunnon6":
        gpu_inserts = res.nodes

        def ref_job():
            got = _run_ref("mf_fast", [6, 18, threads])
            if got is None:
                return None
            out, cs = got
            if int(out.split()[0]) != want:
                raise SystemExit("reference mf_fast 6 18 disagreed")
            return {"value": gpu_inserts / cs, "unit": "inserts/s", "cores": threads, "kind": "reference",
                    "sample": f"oracle/_ref/mf_fast 6 18 {threads} (reference source built here): the whole 6x6 map "
                              f"from the same depth-18 frontier, answer {want} = the GPU's, {gpu_inserts} inserts "
                              f"below the frontier, {cs:.1f} s wall"}
        jobs.append(ref_job)

        def port_job():
            from oracle import oracle
            m = 800
            t1 = time.perf_counter()
            _, _, _, st = oracle.lunnon_resume(dims, depth, states[:m])
            cs = time.perf_counter() - t1
            return {"value": st["inserts"] / cs, "unit": "inserts/s", "cores": 1, "kind": "port",
                    "sample": f"first {m} frontier states, oracle C restatement of mf.c / solve_from_state, "
                              f"1 thread, {st['inserts']} inserts, {cs:.1f} s"}
        jobs.append(port_job)
    else:
        def ref_job():
            # `mf p q` for every map of the draw, `threads` processes at a
            # time, largest first; the insert counts of each full tree are the
            # oracle's exact event counts (golden table).
            from concurrent.futures import ThreadPoolExecutor
            if not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "mf")):
                return None
            with open(os.path.join(ROOT, "tests", "golden", "rect_inserts.json")) as f:
                ins = json.load(f)["inserts"]
            order = sorted(range(len(shapes)), key=lambda i: -ins["%dx%d" % shapes[i]])
            t1 = time.perf_counter()

            def one(i):
                p, q = shapes[i]
                out, cs = _run_ref("mf", [p, q])
                if int(out.split()[0]) != want_maps[i]:
                    raise SystemExit(f"reference mf {p} {q} disagreed with the golden table")
                if cs > 5.0:
                    log(f"cpu baseline rect64: mf {p} {q} done in {cs:.1f} s "
                        f"({time.perf_counter() - t1:.0f} s into the batch)")
                return cs
            with ThreadPoolExecutor(max_workers=threads) as pool:
                secs = list(pool.map(one, order))
            wall = time.perf_counter() - t1
            inserts = sum(ins["%dx%d" % s] for s in shapes)
            return {"value": inserts / wall, "unit": "inserts/s", "cores": threads, "kind": "reference",
                    "sample": f"oracle/_ref/mf p q (reference source built here) for all {len(shapes)} maps of the "
                              f"draw, {threads} processes at a time, largest first; every answer = the golden "
                              f"table = the GPU's; {inserts} inserts of the full trees (the oracle's exact "
                              f"counts); {wall:.1f} s wall, {sum(secs):.1f} process-seconds, longest map "
                              f"{max(secs):.1f} s"}
        jobs.append(ref_job)
    return rec, jobs


# ----------------------------------------------------------------------------- main


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="all",
                    choices=["all", "sym8", "sym7", "sym9", "lunnon5", "lunnon6", "rect64"],
                    help="all (default): the 8x8 headline chunk plus every other configuration in `configs`")
    ap.add_argument("--headline-only", action="store_true", help="only the 8x8 chunk (no other configurations)")
    ap.add_argument("--start", type=int, default=None, help="headline chunk start (default 120,000,000)")
    ap.add_argument("--cases", type=int, default=None, help="headline chunk size (default 50,000)")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain static-frontier kernel")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-tune", action="store_true",
                    help="fixed knobs instead of the tuner (sym: --deepen/--theta/--cap/--budget)")
    ap.add_argument("--deepen", type=int, default=None, help="sym with --no-tune: deepening target (-1 off)")
    ap.add_argument("--theta", type=float, default=None,
                    help="sym with --no-tune: learned threshold (log2 predicted nodes); unset = no learned targets")
    ap.add_argument("--cap", type=int, default=4, help="sym with --no-tune: most levels the model adds")
    ap.add_argument("--budget", type=int, default=64, help="with --no-tune: re-split budget")
    ap.add_argument("--tail", type=int, default=0,
                    help="sym with --no-tune: steps a lane runs after the input is drained (0: the budget)")
    ap.add_argument("--waves", type=int, default=0,
                    help="Lunnon workloads with --no-tune: resident waves per CU (0: the library default)")
    ap.add_argument("--lanes", type=int, default=1, choices=[1, 2],
                    help="sym with --no-tune: lanes per subtree (2: a lane pair splits each node's page walks)")
    ap.add_argument("--no-compare", action="store_true",
                    help="sym: skip the no-model / undeepened / end-to-end runs (profiling)")
    ap.add_argument("--cost-params", default="",
                    help="load cost-model parameters (.npy) instead of training them on the GPU")
    ap.add_argument("--seed", type=int, default=0, help="rect64: seed of the shape draw")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on exactly one GPU: this workload has no multi-GPU path. A chunk would "
              "shard as independent case ranges, one per GPU, whose exact counts add (one 16-byte "
              "u128 sum per rank).", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import map_folding_amd as mf
    from map_folding_amd import cost
    from map_folding_amd import roofline as rl

    mf.init(0)
    cus = mf.device_cus()
    threads = host_threads()
    wl = args.workload
    head = "sym8" if wl == "all" else wl
    others = [] if (wl != "all" or args.headline_only) else ["lunnon5", "lunnon6", "rect64", "sym7"]
    cpu_jobs = []   # (config name, job)

    # The learned part: the cost model, trained on the GPU (pairs logged by
    # the GPU DFS on windows disjoint from the timed chunks).
    model = pairs = losses = train_s = None
    if head in ("sym8", "sym7") or "sym7" in others:
        t1 = time.perf_counter()
        if args.cost_params:
            model = cost.CostModel(seed=0)
            model.load(args.cost_params)
        else:
            model, losses, pairs = cost.train_default(seed=0)
        train_s = time.perf_counter() - t1
        log(f"cost model: {train_s:.1f} s to {'load' if args.cost_params else f'collect {pairs} pairs and train'}")

    configs = []
    out = None
    if head in SYM_WORKLOADS:
        rec, ex = sym_workload(args, mf, rl, cost, head, args.steps, args.warmup, model, pairs, losses, train_s,
                               headline=True)
        out = {"headline": rec, "ex": ex}
    else:
        rec, jobs = lunnon_workload(args, mf, rl, head, args.steps, args.warmup)
        out = {"headline": rec, "ex": None}
        cpu_jobs += [(head if k == 0 else head + ":port", j) for k, j in enumerate(jobs)]
    for name in others:
        st = min(args.steps, 5)
        if name in SYM_WORKLOADS:
            rec, ex = sym_workload(args, mf, rl, cost, name, st, 1, model, pairs, losses, train_s)
            if not args.no_cpu:
                cpu_jobs.append((name, sym_ref_job(REF_EXE[ex["n"]], ex["n"], ex["level"], ex["start"],
                                                   ex["n_inputs"], ex["total"], ex["nodes"], threads)))
                k = PORT_CASES[name]
                cpu_jobs.append((name + ":port", sym_port_job(ex["n"], ex["level"], ex["cyc"][:k], ex["mul"][:k],
                                                              threads)))
            ex["front"].close()
        else:
            rec, jobs = lunnon_workload(args, mf, rl, name, st, 1)
            cpu_jobs += [(name if k == 0 else name + ":port", j) for k, j in enumerate(jobs)]
        rec["name"] = name
        configs.append(rec)

    peak, peak_notes = roof_peak(rl, cus)
    h = out["headline"]
    ex = out["ex"]
    if ex is not None and not args.no_cpu:
        cpu_jobs.insert(0, (head, sym_ref_job(REF_EXE[ex["n"]], ex["n"], ex["level"], ex["start"], ex["n_inputs"],
                                              ex["total"], ex["nodes"], threads)))
        k = PORT_CASES[head]
        cpu_jobs.insert(1, (head + ":port", sym_port_job(ex["n"], ex["level"], ex["cyc"][:k], ex["mul"][:k],
                                                         threads)))

    # The plain static-frontier kernel on the whole headline window (one lane
    # per case, one launch): it runs while the host runs the CPU baselines.
    plain = None
    plain_box = {}
    th = None
    if ex is not None and not args.no_plain:
        front = ex["front"]

        def run_plain():
            t1 = time.perf_counter()
            rp, _, _ = front.solve(mf.SolveOptions(schedule=mf.MF_SCHED_STATIC, count_events=False))
            plain_box["wall"] = time.perf_counter() - t1
            plain_box["res"] = rp
        log(f"plain static-frontier kernel on all {ex['n_inputs']} cases (largest case "
            f"{int(ex['case_nodes'].max())} nodes): running while the CPU baselines run")
        th = threading.Thread(target=run_plain)
        th.start()
    cpu = {}
    for cname, job in cpu_jobs:
        if args.no_cpu:
            break
        t1 = time.perf_counter()
        r = job()
        cpu[cname] = r
        if r is not None:
            log(f"cpu baseline {cname}: {r['value']:.3e} {r['unit']} on {r['cores']} threads "
                f"({time.perf_counter() - t1:.1f} s)")
    # While the plain kernel still runs, the oracle-port sample of the
    # headline window grows by further blocks of cases.
    hp = cpu.get(head + ":port")
    while th is not None and th.is_alive() and hp is not None and hp["cases"] < ex["n_inputs"]:
        k, blk = hp["cases"], PORT_CASES[head]
        extend_port(hp, sym_port_job(ex["n"], ex["level"], ex["cyc"][k:k + blk], ex["mul"][k:k + blk], threads)())
        log(f"cpu baseline {head}:port: {hp['value']:.3e} nodes/s over the first {hp['cases']} cases")
    if th is not None:
        th.join()
        rp, pw = plain_box["res"], plain_box["wall"]
        if rp.count * ex["n"] ** 2 != ex["total"]:
            raise SystemExit("the plain static kernel disagreed")
        plain = {"cases": ex["n_inputs"], "ms": pw * 1e3, "kernel_ms": rp.kernel_ms, "nodes_per_s": ex["nodes"] / pw,
                 "largest_case_nodes": int(ex["case_nodes"].max()),
                 "frac_of_int32_peak": rl.sym_ops(ex["nodes"], ex["walks"], ex["steps"]) / (rp.kernel_ms / 1e3)
                 / 1e12 / peak,
                 "note": "one lane per case, one launch: bounded by the lane of the largest case; timed while "
                         "the CPU baselines ran on the host"}
        log(f"plain static: {pw*1e3:.1f} ms on the whole window")

    # Roofline per record.
    for rec in [h] + configs:
        rec["roofline_frac"] = rec["achieved_tops"] / peak
    if ex is not None:
        # Every item a round handles is read once; all but the first round's
        # (written by the deepening pass or the upload) were written once by
        # the round that split them.
        bytes_per_launch = (ex["item_bytes"] * (2 * h["items_per_step"] - h["first_round_items"]) / h["rounds"]
                            if h["items_per_step"] else None)
        result = {
            "metric": "DFS nodes/s & cases/s on 1 MI355X, batched+learned vs plain, bit-exact",
            "value": h["nodes_per_s"], "unit": "nodes/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": h["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "int32", "data": "synthetic: deterministic sym_fold case stream (no dataset)",
            "config": {"workload": h["workload"], "cases": h["cases"], "schedule": h["schedule"],
                       "chunk_total": h["total"], "bit_exact_vs_golden": h["bit_exact_vs_golden"],
                       "tuned": h["tuned"]},
            "cases_per_s": h["cases_per_s"], "nodes": h["nodes"], "page_walks": h["page_walks"],
            "walk_steps": h["walk_steps"], "leaves": h["leaves"], "rounds": h["rounds"],
            "kernel_ms_per_step": h["kernel_ms_per_step"], "items_per_step": h["items_per_step"],
            "first_round_items": h["first_round_items"],
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "roofline": {"bound": "valu", "achieved": h["achieved_tops"], "peak": peak, "unit": "TOP/s",
                         "frac": h["roofline_frac"], "traffic": _traffic(head, *(SYM_WORKLOADS[head][2:4]))
                         if (args.start, args.cases) == (None, None) else None,
                         "ops_per_step": h["ops_per_step"],
                         "ops_model": "16*nodes + 6*page_walks + 8*walk_steps (int32), exact events of the window",
                         **peak_notes},
            "no_model": h.get("no_model"), "undeepened": h.get("undeepened"), "learned": h["learned"],
            "end_to_end_ms": h["end_to_end_ms"],
            "plain": plain, "speedup_vs_plain": (plain["ms"] / h["ms_per_step"]) if plain else None,
            "cpu_baseline": cpu.get(head) or cpu.get(head + ":port"), "cpu_port": cpu.get(head + ":port"),
            "configs": [],
        }
    else:
        result = {
            "metric": "DFS nodes/s & cases/s on 1 MI355X, batched+learned vs plain, bit-exact",
            "value": h["inserts_per_s"], "unit": "inserts/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": h["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "int32", "data": "synthetic: deterministic Lunnon frontiers",
            "config": {"workload": h["workload"], "total": h["total"], "bit_exact_vs_golden": True},
            "roofline": {"bound": "valu", "achieved": h["achieved_tops"], "peak": peak, "unit": "TOP/s",
                         "frac": h["roofline_frac"], "traffic": None, "ops_per_step": h["ops_per_step"],
                         "ops_model": "10*gapcomps + 8*walk_steps + 6*candidates + 8*inserts + 2*completions",
                         **peak_notes},
            "autotune": h["autotune"], "cpu_baseline": cpu.get(head), "configs": [],
        }
    for rec in configs:
        name = rec["name"]
        ent = {"name": name, "workload": rec["workload"], "ms_per_step": rec["ms_per_step"],
               "kernel_ms_per_step": rec["kernel_ms_per_step"], "bit_exact_vs_golden": rec["bit_exact_vs_golden"],
               "total": rec["total"], "roofline_frac": rec["roofline_frac"], "achieved_tops": rec["achieved_tops"]}
        if "nodes_per_s" in rec:
            ent.update({"rate": rec["nodes_per_s"], "unit": "nodes/s", "cases_per_s": rec["cases_per_s"],
                        "schedule": rec["schedule"], "tuned": rec["tuned"], "learned": rec["learned"],
                        "no_model": rec.get("no_model"), "undeepened": rec.get("undeepened"),
                        "end_to_end_ms": rec["end_to_end_ms"]})
        else:
            ent.update({"rate": rec["inserts_per_s"], "unit": "inserts/s", "items_per_s": rec["items_per_s"],
                        "autotune": rec["autotune"]})
        ent["cpu_baseline"] = cpu.get(name) or cpu.get(name + ":port")
        if cpu.get(name + ":port") and cpu.get(name):
            ent["cpu_port"] = cpu[name + ":port"]
        result["configs"].append(ent)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
