"""Algorithmic work model of the DFS kernels and the MI355X int32 VALU roof.

The figure of merit is exact and independent of scheduling: each event the
DFS performs (counted exactly by the kernels, checked against the oracle) is
priced at a fixed number of int32 operations -- the minimal scalar sequence
for that event.  The weights are frozen here and mirrored as MF_W_* in
include/mapfold.h.

symmetry DFS (per get_good node / page walk / walk step):
    node 16  mask init, 4 page tests, ctz, 2 stores + 1 load to place a cell,
             push/pop of the remaining-children mask
    walk  6  anchor lookup, compare, first successor load, mask merge
    step  8  partner load, compare, select, bit set into a 64-bit mask (3),
             successor load, loop compare
Lunnon DFS (per gap computation / walk step / candidate / insert / completion):
    10 / 8 / 6 / 8 / 2

The roof is the int32 VALU rate the device sustains, measured by the
`valu_peak` kernel (16 independent add/xor chains per lane, 32 waves per CU):
6.1-6.2e13 lane-ops/s on one MI355X.  That is 1.57x the CUs x 64 lanes x
2.4 GHz figure, so the latter is reported beside it but never divided by.
MI355X's published vector FP32 rate, 157.3 TFLOP/s, is 78.6e12 lane-FMAs/s
through packed FP32 instructions that integer code has no counterpart of.
"""
from __future__ import annotations

import os
import re
import subprocess

W_SYM_NODE, W_SYM_WALK, W_SYM_STEP = 16, 6, 8
W_LUN_GAPCOMP, W_LUN_STEP, W_LUN_CAND, W_LUN_INSERT, W_LUN_COMPLETION = 10, 8, 6, 8, 2

MI355X_CLOCK_HZ = 2.4e9          # peak engine clock
LANES_PER_CU_PER_CLK = 64        # 4 SIMD x 16 lanes, one int32 op each per clock
HBM_PEAK_GBS = 8000.0
INT32_MEASURED_OPS = 6.19e13     # valu_peak's best reading on one MI355X (256 CUs); readings span 5.5-6.3e13
PUBLISHED_FP32_LANE_OPS = 157.3e12 / 2


def int32_peak_ops(cus: int = 256) -> float:
    """CUs x 64 lanes x 2.4 GHz (39.3 T at 256 CUs): a spec-sheet figure that
    the device exceeds for int32 VALU work (see measured_int32_peak)."""
    return cus * LANES_PER_CU_PER_CLK * MI355X_CLOCK_HZ


def measured_int32_peak(root: str, timeout: float = 60.0):
    """(ops/s, source): runs the built `valu_peak` program on the current
    device and returns its reading; the recorded one-MI355X reading if the
    program is not built or does not finish."""
    exe = os.path.join(root, "map_folding_amd", "bin", "valu_peak")
    if os.path.exists(exe):
        try:
            out = subprocess.run([exe], capture_output=True, text=True, timeout=timeout, check=True).stdout
            m = re.search(r"^peak_ops_per_s (\S+)$", out, re.M)
            if m:
                return float(m.group(1)), "measured now by valu_peak (16 add/xor chains, 32 waves/CU)"
        except (OSError, subprocess.SubprocessError):
            pass
    return INT32_MEASURED_OPS, "valu_peak reading recorded on one MI355X (program not run)"


def sym_ops(nodes: int, walks: int, steps: int) -> int:
    return W_SYM_NODE * nodes + W_SYM_WALK * walks + W_SYM_STEP * steps


def lunnon_ops(gapcomps: int, steps: int, cands: int, inserts: int, completions: int) -> int:
    return (W_LUN_GAPCOMP * gapcomps + W_LUN_STEP * steps + W_LUN_CAND * cands + W_LUN_INSERT * inserts
            + W_LUN_COMPLETION * completions)
