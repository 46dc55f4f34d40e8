"""The roofline's work model and peak (map_folding_amd/roofline.py).

The per-event weights are mirrored in include/mapfold.h (MF_W_*) for C
callers; both copies must agree.  The peak is the measured int32 VALU rate;
without a GPU the `valu_peak` program cannot run and the recorded reading is
returned, never the CUs x 64 x clock figure.
"""
import os
import re

from map_folding_amd import roofline as rl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_weights_match_the_c_header():
    with open(os.path.join(ROOT, "include", "mapfold.h")) as f:
        text = f.read()
    got = {k: int(v) for k, v in re.findall(r"MF_W_(\w+) = (\d+)", text)}
    want = {"SYM_NODE": rl.W_SYM_NODE, "SYM_WALK": rl.W_SYM_WALK, "SYM_STEP": rl.W_SYM_STEP,
            "LUN_GAPCOMP": rl.W_LUN_GAPCOMP, "LUN_STEP": rl.W_LUN_STEP, "LUN_CAND": rl.W_LUN_CAND,
            "LUN_INSERT": rl.W_LUN_INSERT, "LUN_COMPLETION": rl.W_LUN_COMPLETION}
    assert got == want


def test_ops_are_the_weighted_sums():
    assert rl.sym_ops(1, 0, 0) == 16 and rl.sym_ops(0, 1, 0) == 6 and rl.sym_ops(0, 0, 1) == 8
    assert rl.sym_ops(3, 5, 7) == 3 * 16 + 5 * 6 + 7 * 8
    assert rl.lunnon_ops(1, 2, 3, 4, 5) == 10 + 16 + 18 + 32 + 10


def test_peak_falls_back_to_the_recorded_reading(tmp_path):
    ops, source = rl.measured_int32_peak(str(tmp_path))
    assert ops == rl.INT32_MEASURED_OPS and "not run" in source
    # The recorded reading is above the spec-sheet figure it replaces.
    assert rl.INT32_MEASURED_OPS > 1.5 * rl.int32_peak_ops(256)
    assert abs(rl.int32_peak_ops(256) - 39.3216e12) < 1e6
