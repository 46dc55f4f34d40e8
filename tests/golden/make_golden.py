#!/usr/bin/env python3
"""Regenerate or verify the `reference_binaries` entries of tests/golden/golden.json.

Every value is printed by the reference's own CPU solvers, compiled from
their sources where they lie (`make -C oracle ref` -> oracle/_ref/: mf,
mf_fast, sym_fold, sym_fold9).  Nothing here imports or copies reference
source.

    python tests/golden/make_golden.py            # verify (exit 1 on a mismatch)
    python tests/golden/make_golden.py --write    # rewrite the regenerated entries
    python tests/golden/make_golden.py --quick    # skip the runs of minutes (2x15, 3x10, 8x8 50k chunk)

Entries and the binary that prints them:
  rectangles                 `mf p q` for p in {2,3,4}, 12 <= pq <= 30 (bare total on stdout)
  strips                     `mf 1 5`
  lunnon_frontier_sizes      `MF_COLLECT_ONLY=1 mf_fast g D` (first stdout line); 7x7 rows are published
  sym_seeds                  `sym_fold n` stderr center_cases= / reduced_cases=
  sym_case_totals 7@24, 8@24 `MF_SETUP_ONLY=1 sym_fold n` (materialised) stderr expanded_cases=
  sym_chunks answers         `MF_STREAM_CASES=1 MF_EXPANDED_N=E sym_fold n 8 max start` (stdout)
  sym_chunks nodes / leaves  `MF_PROFILE_WORK=1 ... sym_fold9` profile_nodes / profile_leaves totals
                             (a node is one recurse_count_work call, sym_fold9.cpp:342-357)

Not regenerated here (pinned otherwise, see tests/test_oracle.py): the published
totals, the multiplier histograms (Sigma mult = 2 x raw is checked), the 5x5
event counts and the page_walks / walk_steps of the [0,300) windows, which no
reference binary prints.
About 6 minutes on 8 cores in full (2x15 alone is ~2 minutes single-threaded).
"""
import argparse
import concurrent.futures as cf
import json
import os
import re
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF = os.path.join(ROOT, "oracle", "_ref")
GOLDEN = os.path.join(ROOT, "tests", "golden", "golden.json")
SLOW = {"2x15", "3x10", "2x14", (8, 36, 120000000, 50000)}


def run(exe, args, env=None):
    e = {k: v for k, v in os.environ.items() if not k.startswith("MF_")}
    e.update(env or {})
    r = subprocess.run([os.path.join(REF, exe)] + [str(a) for a in args], capture_output=True, text=True, env=e,
                       check=True)
    return r.stdout, r.stderr


def first_int(text):
    for line in text.splitlines():
        if line.strip().isdigit():
            return int(line.strip())
    raise ValueError(f"no integer line in {text!r}")


def summary_total(stderr, label):
    m = re.search(rf"^{label}_count=\S+ .* total=(\d+)", stderr, re.M)
    if not m:
        raise ValueError(f"no {label} summary")
    return int(m.group(1))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--write", action="store_true")
    ap.add_argument("--quick", action="store_true")
    a = ap.parse_args()
    if not os.path.exists(os.path.join(REF, "sym_fold9")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    with open(GOLDEN) as f:
        g = json.load(f)
    jobs = {}
    with cf.ThreadPoolExecutor(max_workers=os.cpu_count() or 4) as pool:
        for key in g["rectangles"]:
            if key.startswith("_") or (a.quick and key in SLOW):
                continue
            p, q = map(int, key.split("x"))
            jobs[("rectangles", key)] = pool.submit(lambda p=p, q=q: first_int(run("mf", [p, q])[0]))
        jobs[("strips", "1x5")] = pool.submit(lambda: first_int(run("mf", [1, 5])[0]))
        for key in ("5x5@15", "6x6@15", "6x6@18"):
            grid, depth = key.split("@")
            jobs[("lunnon_frontier_sizes", key)] = pool.submit(
                lambda s=int(grid[0]), d=int(depth): first_int(run("mf_fast", [s, d, 4], {"MF_COLLECT_ONLY": "1"})[0]))
        for n in ("7", "8"):
            def seeds(n=int(n)):
                err = run("sym_fold", [n, 4], {"MF_SETUP_ONLY": "1", "MF_STREAM_CASES": "1",
                                              "MF_EXPANDED_N": "24"})[1]
                return (int(re.search(r"center_cases=(\d+)", err).group(1)),
                        int(re.search(r"reduced_cases=(\d+)", err).group(1)))
            jobs[("sym_seeds", n)] = pool.submit(seeds)
        for key in ("7@24", "8@24"):
            n, e = map(int, key.split("@"))
            jobs[("sym_case_totals", key)] = pool.submit(
                lambda n=n, e=e: int(re.search(r"expanded_cases=(\d+)", run(
                    "sym_fold", [n, 4], {"MF_SETUP_ONLY": "1", "MF_EXPANDED_N": str(e)})[1]).group(1)))
        for k, it in enumerate(g["sym_chunks"]["items"]):
            if it.get("source") == "published":
                continue
            w = (it["n"], it["expanded"], it["start"], it["max"])
            if a.quick and w in SLOW:
                continue
            env = {"MF_STREAM_CASES": "1", "MF_EXPANDED_N": str(w[1])}
            jobs[("sym_chunks", k, "answer")] = pool.submit(
                lambda w=w, env=env: first_int(run("sym_fold", [w[0], 8, w[3], w[2]], env)[0]))
            if "nodes" in it:
                def prof(w=w, env=env):
                    err = run("sym_fold9", [w[0], 8, w[3], w[2]], dict(env, MF_PROFILE_WORK="1"))[1]
                    return summary_total(err, "profile_nodes"), summary_total(err, "profile_leaves")
                jobs[("sym_chunks", k, "profile")] = pool.submit(prof)
        bad = 0
        for key, fut in jobs.items():
            got = fut.result()
            sec = key[0]
            if sec == "sym_seeds":
                want = (g[sec][key[1]]["raw"], g[sec][key[1]]["seeds"])
                g[sec][key[1]]["raw"], g[sec][key[1]]["seeds"] = got
            elif sec == "sym_chunks":
                it = g[sec]["items"][key[1]]
                if key[2] == "answer":
                    want = it["answer"]
                    it["answer"] = got
                else:
                    # profile_leaves of sym_fold9 counts leaves as recurse_count_work does
                    # (popcount at the last level), the golden `leaves`.
                    want = (it["nodes"], it["leaves"]) if "leaves" in it else (it["nodes"],)
                    got = got[:len(want)]
                    it["nodes"] = got[0]
                    if "leaves" in it:
                        it["leaves"] = got[1]
            else:
                want = g[sec][key[1]]
                g[sec][key[1]] = got
            status = "ok" if got == want else "MISMATCH"
            bad += status != "ok"
            print(f"{status:8s} {key}: {got}" + ("" if status == "ok" else f" (golden {want})"), flush=True)
    if a.write:
        with open(GOLDEN, "w") as f:
            json.dump(g, f, indent=1)
            f.write("\n")
        print(f"wrote {GOLDEN}")
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main())
