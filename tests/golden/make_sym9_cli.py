"""Golden runs of the 9x9 command-line tool (experiments/9x9/sym_fold9.cpp).

The reference binary is built from its own source by `make -C oracle ref`
(oracle/_ref/sym_fold9) and run on a handful of invocations that exercise
what only the 9x9 tool has: seed slices (MF_SEED_START / MF_SEED_COUNT),
MF_CENTER_N, MF_ORDER / MF_LOCK_CENTER_N cell numberings, per-case work
profiles (MF_PROFILE_WORK), the cost-sorted
materialised mode with a 5x5 center, and raw_count for n <= 5.  stdout is kept
whole; stderr is kept without the lines and fields that carry wall-clock
times.  Writes tests/golden/sym9_cli.json; about 20 minutes on 8 cores (most
of it the 3x3-center window).
"""
import json
import os
import re
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
BIN = os.path.join(ROOT, "oracle", "_ref", "sym_fold9")

RUNS = [
    (["9", "8", "50", "0"], {"MF_STREAM_CASES": "1", "MF_EXPANDED_N": "44", "MF_PROFILE_WORK": "1"}),
    (["9", "8", "100", "3"], {"MF_STREAM_CASES": "1", "MF_EXPANDED_N": "60", "MF_SEED_START": "5",
                              "MF_SEED_COUNT": "10"}),
    (["9", "8", "400", "0"], {"MF_STREAM_CASES": "1", "MF_EXPANDED_N": "62", "MF_SEED_START": "465000",
                              "MF_PROFILE_WORK": "1"}),
    (["7", "8", "50", "100"], {"MF_CENTER_N": "5", "MF_COST_DEPTH": "3", "MF_SORT_CASES": "desc",
                               "MF_EXPANDED_N": "30"}),
    (["9", "8", "2000", "10"], {"MF_CENTER_N": "3", "MF_STREAM_CASES": "1", "MF_EXPANDED_N": "40"}),
    (["5", "8"], {}),
    # MF_ORDER / MF_LOCK_CENTER_N: ring-wise cell numberings of odd squares
    (["9", "8", "300", "1000"], {"MF_STREAM_CASES": "1", "MF_EXPANDED_N": "62", "MF_ORDER": "reverse"}),
    (["9", "8", "300", "1000"], {"MF_STREAM_CASES": "1", "MF_EXPANDED_N": "62", "MF_ORDER": "axes_first",
                                 "MF_LOCK_CENTER_N": "5"}),
    (["7", "8", "1000", "0"], {"MF_STREAM_CASES": "1", "MF_EXPANDED_N": "24", "MF_ORDER": "corners_first"}),
    (["7", "8", "200", "1000"], {"MF_STREAM_CASES": "1", "MF_EXPANDED_N": "28", "MF_ORDER": "corners_first"}),
    (["7", "8", "500", "2000"], {"MF_STREAM_CASES": "1", "MF_EXPANDED_N": "30", "MF_ORDER": "opposite"}),
    (["7", "8", "50", "0"], {"MF_ORDER": "corners_first", "MF_COST_DEPTH": "2", "MF_SORT_CASES": "desc",
                             "MF_EXPANDED_N": "22"}),
    (["5", "8"], {"MF_ORDER": "corners_first"}),
    (["9", "8"], {"MF_SETUP_ONLY": "1", "MF_CENTER_HIST": "1"}),
    # setup-only runs (no counting: the host side of the tool)
    (["9", "8"], {"MF_SETUP_ONLY": "1", "MF_STREAM_CASES": "1", "MF_EXPANDED_N": "30", "MF_SEED_START": "100",
                  "MF_SEED_COUNT": "50", "MF_SEED_COUNTS": "1", "MF_SEED_COUNTS_SUMMARY": "1"}),
    (["9", "8"], {"MF_SETUP_ONLY": "1", "MF_STREAM_CASES": "1", "MF_EXPANDED_N": "28", "MF_SEED_START": "465880",
                  "MF_SEED_COUNTS_SUMMARY": "1"}),
    (["9", "8"], {"MF_SETUP_ONLY": "1", "MF_EXPANDED_N": "26", "MF_SEED_START": "465800"}),
    (["9", "8"], {"MF_SETUP_ONLY": "1", "MF_COST_DEPTH": "2", "MF_SORT_CASES": "asc", "MF_EXPANDED_N": "27",
                  "MF_SEED_START": "1000", "MF_SEED_COUNT": "20000"}),
    (["7", "4"], {"MF_SETUP_ONLY": "1", "MF_CENTER_N": "5", "MF_STREAM_CASES": "1", "MF_EXPANDED_N": "26",
                  "MF_CENTER_HIST": "1"}),
    (["7", "4"], {"MF_SETUP_ONLY": "1", "MF_CENTER_N": "3", "MF_STREAM_CASES": "1", "MF_EXPANDED_N": "20",
                  "MF_SEED_START": "10", "MF_SEED_COUNT": "3", "MF_SEED_COUNTS": "1"}),
    (["9", "4"], {"MF_SETUP_ONLY": "1", "MF_CENTER_N": "3", "MF_STREAM_CASES": "1", "MF_EXPANDED_N": "20"}),
    (["6", "4"], {"MF_SETUP_ONLY": "1", "MF_SEED_START": "3"}),
    (["9", "8"], {"MF_SETUP_ONLY": "1", "MF_STREAM_CASES": "1", "MF_EXPANDED_N": "28", "MF_ORDER": "opposite",
                  "MF_SEED_START": "400000", "MF_SEED_COUNT": "30", "MF_SEED_COUNTS": "1"}),
    (["9", "4"], {"MF_CENTER_N": "4"}),
    (["9", "4"], {"MF_STREAM_CASES": "1", "MF_EXPANDED_N": "44", "MF_PROFILE_WORK": "1"}),
]


def clean_stderr(text):
    """Drop what depends on the clock: `total_s` lines, progress lines, `*_s=` fields."""
    out = []
    for line in text.splitlines():
        if "total_s=" in line or line.startswith(("progress=", "stream selected=")):
            continue
        out.append(re.sub(r" \w*_s=\S+", "", line))
    return out


def main():
    """`--fresh` runs every invocation again; otherwise entries already in the
    JSON with the same arguments and environment are kept."""
    if not os.path.exists(BIN):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    path = os.path.join(ROOT, "tests", "golden", "sym9_cli.json")
    old = {}
    if os.path.exists(path) and "--fresh" not in sys.argv:
        with open(path) as f:
            for row in json.load(f)["runs"]:
                old[json.dumps([row["args"], row["env"]], sort_keys=True)] = row
    rows = []
    for args, env in RUNS:
        key = json.dumps([args, env], sort_keys=True)
        if key in old and "rc" in old[key]:
            rows.append(old[key])
            continue
        e = dict(os.environ)
        for k in list(e):
            if k.startswith("MF_"):
                del e[k]
        e.update(env)
        r = subprocess.run([BIN] + args, capture_output=True, text=True, env=e)
        rows.append({"args": args, "env": env, "rc": r.returncode, "stdout": r.stdout,
                     "stderr": clean_stderr(r.stderr)})
        print(args, env, r.returncode, r.stdout.strip(), file=sys.stderr)
    doc = {"_source": "oracle/_ref/sym_fold9 (reference experiments/9x9/sym_fold9.cpp built from source); "
                      "stderr without timing lines/fields", "runs": rows}
    with open(path, "w") as f:
        json.dump(doc, f, indent=1)
        f.write("\n")
    print("wrote", path, file=sys.stderr)


if __name__ == "__main__":
    main()
