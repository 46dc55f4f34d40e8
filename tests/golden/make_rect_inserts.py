#!/usr/bin/env python3
"""Full-tree insert counts of every rectangle of the batched-instances draw
(p in {2,3,4}, 12 <= p*q <= 30), from the oracle's exact event counters (the
DFS of mf.c, whose totals equal `mf p q`).  The CPU baseline of the
64-rectangle configuration divides these by the wall time of the reference's
`mf p q` runs.  Reads the totals from golden.json and writes
rect_inserts.json.  About 2-3 minutes on 8 cores (2x15 alone is ~2 minutes)."""
import concurrent.futures as cf
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)


def count(shape):
    from oracle import oracle
    p, q = shape
    tot, st = oracle.lunnon_count([p, q])
    return shape, tot, st["inserts"]


def main():
    with open(os.path.join(ROOT, "tests", "golden", "golden.json")) as f:
        g = json.load(f)
    shapes = [tuple(map(int, k.split("x"))) for k in g["rectangles"] if not k.startswith("_")]
    out = {}
    with cf.ProcessPoolExecutor(max_workers=os.cpu_count() or 4) as pool:
        for (p, q), tot, ins in pool.map(count, sorted(shapes, key=lambda s: -s[0] * s[1])):
            if tot != g["rectangles"][f"{p}x{q}"]:
                raise SystemExit(f"{p}x{q}: oracle total {tot} != golden {g['rectangles'][f'{p}x{q}']}")
            out[f"{p}x{q}"] = ins
            print(p, q, ins, flush=True)
    doc = {"_source": "oracle event counters of the full DFS tree of mf p q (tests/golden/make_rect_inserts.py); "
                      "each map's total equals golden.json's rectangles entry",
           "inserts": {k: out[k] for k in sorted(out)}}
    with open(os.path.join(ROOT, "tests", "golden", "rect_inserts.json"), "w") as f:
        json.dump(doc, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
