"""mf_sym_solve_noalloc (no device allocation; scratch in a static device
array) against mf_sym_solve, its allocating counterpart: the same counts,
per-case counts and nodes and event counters, bit-exact, on two calls in a
row (the second smaller, so a stale tail of the shared array would show),
for 8x8 depth-36, 7x7 depth-24 and 9x9 cases."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(mf, n, level, cyc, mul, opts=None):
    L = mf._abi.lib()
    cyc = np.ascontiguousarray(cyc, dtype=np.uint8)
    mul = np.ascontiguousarray(mul, dtype=np.uint64)
    num = cyc.shape[0]
    pc = np.zeros(num, dtype=np.uint64)
    pn = np.zeros(num, dtype=np.uint64)
    res = mf.MFResult()
    o, keep = (opts or mf.SolveOptions()).to_c()
    rc = L.mf_sym_solve_noalloc(n, level, cyc.ctypes.data_as(C.POINTER(C.c_uint8)),
                                mul.ctypes.data_as(C.POINTER(C.c_uint64)), num, C.byref(o),
                                pc.ctypes.data_as(C.POINTER(C.c_uint64)), pn.ctypes.data_as(C.POINTER(C.c_uint64)),
                                C.byref(res))
    return rc, res, pc, pn


def _same(a, b):
    return (a.count == b.count and a.nodes == b.nodes and a.walks == b.walks and a.steps == b.steps
            and a.leaves == b.leaves)


@pytest.mark.parametrize("n,level,start,counts", [(8, 36, 120000000, (3000, 700)), (7, 24, 200000, (2000, 300)),
                                                  (9, 44, 0, (200, 50))])
def test_noalloc_matches_allocating_solve(gpu, n, level, start, counts):
    for k, num in enumerate(counts):
        print(f"noalloc {n}x{n} level {level} [{start},+{num}) call {k}", flush=True)
        sys.stdout.flush()
        cyc, mul = gpu.sym_cases(n, level, start, num)
        ref, pc0, pn0 = gpu.sym_solve(n, level, cyc, mul, gpu.SolveOptions(), per_item=True)
        rc, res, pc, pn = _noalloc(gpu, n, level, cyc, mul)
        assert rc == 0, gpu._abi.lib().mf_last_error()
        assert _same(res, ref)
        assert np.array_equal(pc, pc0) and np.array_equal(pn, pn0)


def test_noalloc_tight_budget_and_limits(gpu):
    """A tight budget splits constantly (many spill rounds through the
    static queues); too many cases is refused with MF_ERR_CAPACITY."""
    cyc, mul = gpu.sym_cases(8, 36, 5000, 1500)
    ref, pc0, _ = gpu.sym_solve(8, 36, cyc, mul, gpu.SolveOptions(), per_item=True)
    rc, res, pc, _ = _noalloc(gpu, 8, 36, cyc, mul, gpu.SolveOptions(budget=4, tail=4))
    assert rc == 0 and _same(res, ref) and np.array_equal(pc, pc0) and res.rounds > 2
    big = np.zeros((65537, 64), dtype=np.uint8)
    rc, _, _, _ = _noalloc(gpu, 8, 36, big, np.ones(big.shape[0], dtype=np.uint64))
    assert rc == gpu._abi.MF_ERR_CAPACITY
