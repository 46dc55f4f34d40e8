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
"""
from __future__ import annotations

W_SYM_NODE, W_SYM_WALK, W_SYM_STEP = 16, 6, 8
W_LUN_GAPCOMP, W_LUN_STEP, W_LUN_CAND, W_LUN_INSERT, W_LUN_COMPLETION = 10, 8, 6, 8, 2

MI355X_CLOCK_HZ = 2.4e9          # peak engine clock
LANES_PER_CU_PER_CLK = 64        # 4 SIMD x 16 lanes, one int32 op each per clock
HBM_PEAK_GBS = 8000.0


def int32_peak_ops(cus: int = 256) -> float:
    """int32 VALU lane-ops per second: CUs x 64 lanes x 2.4 GHz (39.3 T at 256 CUs)."""
    return cus * LANES_PER_CU_PER_CLK * MI355X_CLOCK_HZ


def sym_ops(nodes: int, walks: int, steps: int) -> int:
    return W_SYM_NODE * nodes + W_SYM_WALK * walks + W_SYM_STEP * steps


def lunnon_ops(gapcomps: int, steps: int, cands: int, inserts: int, completions: int) -> int:
    return (W_LUN_GAPCOMP * gapcomps + W_LUN_STEP * steps + W_LUN_CAND * cands + W_LUN_INSERT * inserts
            + W_LUN_COMPLETION * completions)
