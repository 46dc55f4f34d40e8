"""Chunk sweeps of the symmetry-reduced count on the local GPU, in the JSONL
format of the reference's Modal launcher (modal_sym.py:36-77, 152-250).

    python -m map_folding_amd.sweep run      --n 8 --expanded-n 36 --start-case 120000000 --max-cases 50000
    python -m map_folding_amd.sweep sweep    --n 8 --expanded-n 36 --chunk-size 50000 --chunk-count 4 --out r.jsonl
    python -m map_folding_amd.sweep sweep    ... --out r.jsonl --resume-existing
    python -m map_folding_amd.sweep validate --out r.jsonl --chunk-size 50000 [--expected N]

Rows carry the reference's keys (n, threads, expanded_n, start_case,
max_cases, count, wall_s, completed_hint, stderr_tail) so that files from
either launcher can be merged and validated together; `threads` is recorded
as given (the count runs on the GPU).  `count` is the decimal answer of
`sym_fold n threads max_cases start_case` with MF_STREAM_CASES=1, i.e. the
chunk's contribution to the total.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from datetime import datetime, timezone
from pathlib import Path

TOTAL_8X8_DEPTH36_CASES = 240_192_829


def run_chunk(n: int, threads: int, expanded_n: int, start_case: int, max_cases: int) -> dict:
    """One chunk on the GPU (modal_sym.py:88-141 run_chunk, without the subprocess)."""
    import map_folding_amd as mf
    t0 = time.perf_counter()
    count, res = mf.sym_fold(n, expanded=expanded_n, start=start_case, max_cases=max_cases,
                             options=mf.SolveOptions(count_events=False))
    wall_s = time.perf_counter() - t0
    return {
        "n": n, "threads": threads, "expanded_n": expanded_n, "start_case": start_case, "max_cases": max_cases,
        "count": str(count), "wall_s": wall_s, "completed_hint": max_cases,
        "stderr_tail": f"kernel_ms={res.kernel_ms:.1f} rounds={res.rounds} nodes={res.nodes}",
    }


def gen_chunk(n: int, expanded_n: int, start_case: int, max_cases: int):
    """Host side of a chunk: its cases from the stream index (cycles, mults, seconds)."""
    import map_folding_amd as mf
    t0 = time.perf_counter()
    cyc, mul = mf.sym_cases(n, expanded_n, start_case, max_cases)
    return cyc, mul, time.perf_counter() - t0


def solve_chunk(n: int, threads: int, expanded_n: int, start_case: int, max_cases: int, cases,
                schedule: dict | None = None) -> dict:
    """GPU side of a chunk whose cases `gen_chunk` made: upload, (learned:
    features + inference on the device, per-case deepening targets), solve.
    `schedule` holds deepen / model / theta / cap (all optional)."""
    import map_folding_amd as mf
    cyc, mul, gen_s = cases
    schedule = schedule or {}
    t0 = time.perf_counter()
    front = mf.SymFrontier(n, expanded_n, cyc, mul)
    try:
        levels = None
        deepen = schedule.get("deepen")
        if schedule.get("model") is not None:
            from map_folding_amd import cost
            levels = cost.learned_levels(schedule["model"].predict_frontier(front), deepen,
                                         schedule["theta"], schedule["cap"])
        res, _, _ = front.solve(mf.SolveOptions(count_events=False, deepen=deepen if deepen is not None else 0,
                                                deepen_levels=levels))
    finally:
        front.close()
    wall_s = time.perf_counter() - t0
    count = res.count * n * n
    return {
        "n": n, "threads": threads, "expanded_n": expanded_n, "start_case": start_case, "max_cases": max_cases,
        "count": str(count), "wall_s": wall_s, "completed_hint": max_cases,
        "stderr_tail": f"kernel_ms={res.kernel_ms:.1f} rounds={res.rounds} nodes={res.nodes} "
                       f"host_generation_s={gen_s:.4f} (overlapped with the previous chunk) "
                       f"schedule={'learned' if levels is not None else 'uniform'}",
    }


def pipelined_chunks(inputs, schedule: dict | None = None):
    """Rows of the chunks in order, the host generating chunk k+1 on a worker
    thread while the GPU counts chunk k (the generator and the solve both run
    in the native library, outside the interpreter lock)."""
    import concurrent.futures as cf
    if not inputs:
        return
    with cf.ThreadPoolExecutor(max_workers=1) as pool:
        n, _, e, s, m = inputs[0]
        pending = pool.submit(gen_chunk, n, e, s, m)
        for k, item in enumerate(inputs):
            cases = pending.result()
            if k + 1 < len(inputs):
                n1, _, e1, s1, m1 = inputs[k + 1]
                pending = pool.submit(gen_chunk, n1, e1, s1, m1)
            yield solve_chunk(*item, cases, schedule)


def summarize_jsonl(path: Path, total_cases: int, chunk_size: int) -> dict:
    """Coverage and sums of a results file (modal_sym.py:36-77)."""
    rows = []
    if path.exists():
        with path.open("r", encoding="utf-8") as f:
            for line_no, line in enumerate(f, 1):
                text = line.strip()
                if not text:
                    continue
                row = json.loads(text)
                row["_line_no"] = line_no
                rows.append(row)
    expected_starts = list(range(0, total_cases, chunk_size))
    expected_set = set(expected_starts)
    seen: dict[int, dict] = {}
    duplicate_starts = []
    for row in rows:
        start = int(row["start_case"])
        if start in seen:
            duplicate_starts.append(start)
        else:
            seen[start] = row
    missing_starts = [s for s in expected_starts if s not in seen]
    extra_starts = sorted(s for s in seen if s not in expected_set)
    unique_sum = sum(int(r["count"]) for r in seen.values())
    all_rows_sum = sum(int(r["count"]) for r in rows)
    return {
        "path": str(path), "rows": len(rows), "unique_chunks": len(seen), "expected_chunks": len(expected_starts),
        "duplicate_count": len(duplicate_starts), "duplicate_starts": duplicate_starts[:20],
        "missing_count": len(missing_starts), "missing_starts": missing_starts[:20],
        "extra_count": len(extra_starts), "extra_starts": extra_starts[:20],
        "unique_sum": str(unique_sum), "all_rows_sum": str(all_rows_sum),
        "complete": len(seen) == len(expected_starts) and not duplicate_starts and not extra_starts,
    }


def sweep(n: int = 8, threads: int = 16, expanded_n: int = 36, total_cases: int = TOTAL_8X8_DEPTH36_CASES,
          chunk_size: int = 100000, start_case: int = 0, chunk_count: int = 0, out: str = "",
          resume_existing: bool = False, log=print, pipeline: bool = True, schedule: dict | None = None) -> dict:
    """Run chunks in order and append one row per chunk (modal_sym.py:152-232).

    For n >= 7 (chunks are windows of the expanded-case stream) each chunk is
    generated (gen_chunk) and counted (solve_chunk, `schedule` as there); with
    `pipeline` the next chunk's generation overlaps the current solve,
    otherwise it follows it.  For n < 7 each chunk is one run_chunk call."""
    if chunk_size <= 0:
        raise ValueError("chunk_size must be positive")
    starts = list(range(start_case, total_cases, chunk_size))
    if chunk_count > 0:
        starts = starts[:chunk_count]
    inputs = [(n, threads, expanded_n, s, min(chunk_size, total_cases - s)) for s in starts]
    if not inputs:
        raise ValueError("No chunks selected")
    if not out:
        stamp = datetime.now(timezone.utc).strftime("%Y%m%dT%H%M%SZ")
        out = f"results/{n}x{n}-depth{expanded_n}-{stamp}.jsonl"
    out_path = Path(out)
    out_path.parent.mkdir(parents=True, exist_ok=True)
    if resume_existing:
        summary = summarize_jsonl(out_path, total_cases, chunk_size)
        existing = set()
        if out_path.exists():
            with out_path.open("r", encoding="utf-8") as f:
                for line in f:
                    text = line.strip()
                    if text:
                        existing.add(int(json.loads(text)["start_case"]))
        before = len(inputs)
        inputs = [it for it in inputs if it[3] not in existing]
        log(f"resume_existing=1 skipped={before - len(inputs)} remaining={len(inputs)} "
            f"existing_rows={summary['rows']} unique_chunks={summary['unique_chunks']}")
        if not inputs:
            log(json.dumps(summary, indent=2, sort_keys=True))
            return summary
    total = 0
    started = time.perf_counter()
    log(f"chunks={len(inputs)} chunk_size={chunk_size} out={out_path}")
    if n < 7:
        rows = (run_chunk(*item) for item in inputs)
    elif pipeline:
        rows = pipelined_chunks(inputs, schedule)
    else:
        rows = (solve_chunk(*item, gen_chunk(item[0], item[2], item[3], item[4]), schedule) for item in inputs)
    with out_path.open("a", encoding="utf-8") as f:
        for k, row in enumerate(rows, 1):
            total += int(row["count"])
            f.write(json.dumps(row, sort_keys=True) + "\n")
            f.flush()
            log(f"completed={k}/{len(inputs)} start={row['start_case']} wall_s={row['wall_s']:.2f} "
                f"elapsed_s={time.perf_counter() - started:.1f}")
    summary = {
        "chunks": len(inputs), "chunk_size": chunk_size, "start_case": start_case, "total_cases": total_cases,
        "expanded_n": expanded_n, "partial_sum": str(total), "elapsed_s": time.perf_counter() - started,
        "out": str(out_path), "pipelined": bool(pipeline and n >= 7),
    }
    out_path.with_suffix(out_path.suffix + ".summary.json").write_text(
        json.dumps(summary, indent=2, sort_keys=True) + "\n", encoding="utf-8")
    log(json.dumps(summary, indent=2, sort_keys=True))
    return summary


def validate(out: str, total_cases: int = TOTAL_8X8_DEPTH36_CASES, chunk_size: int = 50000,
             expected: str = "") -> dict:
    """modal_sym.py:235-250."""
    summary = summarize_jsonl(Path(out), total_cases, chunk_size)
    if expected:
        summary["expected"] = expected
        summary["matches_expected"] = summary["unique_sum"] == expected
    return summary


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m map_folding_amd.sweep")
    sub = ap.add_subparsers(dest="cmd", required=True)
    common = dict(n=8, threads=16, expanded_n=36)
    r = sub.add_parser("run")
    s = sub.add_parser("sweep")
    v = sub.add_parser("validate")
    for p in (r, s):
        for k, d in common.items():
            p.add_argument("--" + k.replace("_", "-"), type=int, default=d)
    r.add_argument("--start-case", type=int, default=0)
    r.add_argument("--max-cases", type=int, default=100000)
    s.add_argument("--total-cases", type=int, default=TOTAL_8X8_DEPTH36_CASES)
    s.add_argument("--chunk-size", type=int, default=100000)
    s.add_argument("--start-case", type=int, default=0)
    s.add_argument("--chunk-count", type=int, default=0)
    s.add_argument("--out", default="")
    s.add_argument("--resume-existing", action="store_true")
    s.add_argument("--no-pipeline", action="store_true", help="generate each chunk only after the previous solve")
    s.add_argument("--deepen", type=int, default=None, help="deepening target (default: the library's)")
    s.add_argument("--cost-params", default="", help="cost-model parameters (.npy): learned per-case targets")
    s.add_argument("--theta", type=float, default=21.0, help="learned threshold (log2 predicted nodes)")
    s.add_argument("--cap", type=int, default=4, help="most extra levels the model may add")
    v.add_argument("--out", required=True)
    v.add_argument("--total-cases", type=int, default=TOTAL_8X8_DEPTH36_CASES)
    v.add_argument("--chunk-size", type=int, default=50000)
    v.add_argument("--expected", default="")
    a = ap.parse_args(argv)
    if a.cmd == "run":
        print(json.dumps(run_chunk(a.n, a.threads, a.expanded_n, a.start_case, a.max_cases), indent=2,
                         sort_keys=True))
    elif a.cmd == "sweep":
        sched = {"deepen": a.deepen}
        if a.cost_params:
            from map_folding_amd import cost
            model = cost.CostModel(seed=0)
            model.load(a.cost_params)
            if a.deepen is None:
                raise SystemExit("--cost-params needs --deepen (the base the learned targets add to)")
            sched.update(model=model, theta=a.theta, cap=a.cap)
        sweep(a.n, a.threads, a.expanded_n, a.total_cases, a.chunk_size, a.start_case, a.chunk_count, a.out,
              a.resume_existing, log=lambda m: print(m, flush=True), pipeline=not a.no_pipeline, schedule=sched)
    else:
        print(json.dumps(validate(a.out, a.total_cases, a.chunk_size, a.expected), indent=2, sort_keys=True))
    return 0


if __name__ == "__main__":
    sys.exit(main())
