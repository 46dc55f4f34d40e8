"""map_folding_amd -- MI355X-native map-folding DFS (Lunnon and symmetry-reduced).

Python front-end over libmapfold.so (include/mapfold.h).  Every count is
computed by the HIP kernels on the GPU; the host side only generates frontiers
(the reference's CPU half) and moves buffers.  There is no CPU fallback: if
the library or a GPU is missing, calls raise.

Reference operations mirrored (johnowhitaker/map_folding):
    gpu_folding(dims, depth)        gpu_folding.cu:209-311   (`gpu_folding <n> <depth>`)
    lunnon_frontier(dims, depth)    mf_fast.cpp:185-221 / gpu_folding_v2.cu:220-256
    lunnon_solve(dims, depth, S)    gpu_folding_v2.cu:381-453 (solve_kernel)
    sym_seeds(n)                    sym_fold.cpp:689-710
    sym_cases(n, E, start, count)   sym_fold.cpp:491-527 (MF_STREAM_CASES window)
    sym_solve(n, level, cases)      sym_fold.cpp:219-233,529-620
    sym_fold(n, ...)                sym_fold.cpp:649-774 (`sym_fold n threads max start`)
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _abi
from ._abi import (MF_SCHED_RESPLIT, MF_SCHED_STATIC, MFError, MFResult, MFSolveOptions, check,
                   declared_symbols, lib)

__all__ = [
    "MF_SCHED_STATIC", "MF_SCHED_RESPLIT", "MFError", "MFResult", "SolveOptions", "init",
    "device_cus", "lunnon_frontier", "lunnon_solve", "gpu_folding", "lunnon_batch", "sym_seeds",
    "sym_cases", "sym_seed_case_counts", "ALL_SEEDS", "sym_case_total", "sym_solve", "sym_fold",
    "SymFrontier", "LunnonFrontier", "LunnonBatchFrontier", "SymPieces", "sym_expand", "lunnon_frontier_size", "case_bytes",
    "declared_symbols", "sym_set_numbering", "sym_get_numbering", "sym_numbering", "SYM_ORDERS",
]

_P = C.POINTER


def _p(arr, ctype):
    if arr is None:
        return None
    return arr.ctypes.data_as(_P(ctype))


@dataclass
class SolveOptions:
    """Scheduling knobs (never change a count).  Zeros select defaults."""
    schedule: int = MF_SCHED_RESPLIT
    block: int = 0
    waves_per_cu: int = 0
    budget: int = 0
    tail: int = 0
    count_events: bool = True
    core: int = 0                     # stack and walk: 0 default, 1 wide u64 masks, 2 compact 16-bit,
                                      # 3 compact with two lanes per subtree (sym, re-split, up to 8x8),
                                      # 4 compact with unpaired walks (one page / one dimension at a time)
    queue_capacity: int = 0
    order: np.ndarray | None = None   # permutation of item indices (learned order)
    graph: int = 0                    # re-split rounds as a HIP graph: 0 default (on), 1 on, 2 off
    deepen: int = 0                   # sym re-split: device deepening target level: 0 default, -1 off
    deepen_levels: np.ndarray | None = None   # per-item deepening targets (at least `deepen`)

    def to_c(self):
        o = MFSolveOptions()
        o.schedule = int(self.schedule)
        o.block = int(self.block)
        o.waves_per_cu = int(self.waves_per_cu)
        o.budget = int(self.budget)
        o.tail = int(self.tail)
        o.count_events = 1 if self.count_events else 0
        o.core = int(self.core)
        o.queue_capacity = int(self.queue_capacity)
        o.graph = int(self.graph)
        o.deepen = int(self.deepen)
        keep = []
        if self.order is not None:
            a = np.ascontiguousarray(self.order, dtype=np.uint32).reshape(-1)
            o.order = _p(a, C.c_uint32)
            o.order_len = a.shape[0]
            keep.append(a)
        if self.deepen_levels is not None:
            d = np.ascontiguousarray(self.deepen_levels, dtype=np.uint8).reshape(-1)
            o.deepen_levels = _p(d, C.c_uint8)
            o.deepen_levels_len = d.shape[0]
            keep.append(d)
        return o, keep


def _opts(options):
    if options is None:
        options = SolveOptions()
    return options.to_c()


def init(device: int = 0) -> None:
    check(lib().mf_init(device), "mf_init")


def device_cus() -> int:
    return int(lib().mf_device_cus())


def case_bytes(n: int) -> int:
    """Bytes per sym case row: 64 up to 8x8, 128 for 9x9 (u128-mask kernels)."""
    return 64 if n * n <= 64 else 128


def _dims(dims):
    d = np.ascontiguousarray(np.asarray(dims, dtype=np.int32))
    return d, _p(d, C.c_int32)


# ----------------------------------------------------------------------------- Lunnon


def lunnon_frontier(dims, depth: int) -> np.ndarray:
    """Canonical frontier at partition depth D: uint8 [S, D+1] rows b[0..D], DFS order."""
    d, dp = _dims(dims)
    num = C.c_int64()
    check(lib().mf_lunnon_frontier(dp, len(d), depth, None, 0, C.byref(num)), "mf_lunnon_frontier")
    out = np.zeros((num.value, depth + 1), dtype=np.uint8)
    check(lib().mf_lunnon_frontier(dp, len(d), depth, _p(out, C.c_uint8), num.value, C.byref(num)),
          "mf_lunnon_frontier")
    return out


def lunnon_frontier_size(dims, depth: int) -> int:
    """Number of canonical frontier states at partition depth D (nothing stored)."""
    d, dp = _dims(dims)
    num = C.c_int64()
    check(lib().mf_lunnon_frontier(dp, len(d), depth, None, 0, C.byref(num)), "mf_lunnon_frontier")
    return int(num.value)


def lunnon_solve(dims, depth: int, states: np.ndarray, options: SolveOptions | None = None,
                 per_item: bool = False):
    """Resume frontier states on the GPU.  Returns (MFResult, per-state counts, per-state inserts);
    the per-state arrays are None unless per_item."""
    d, dp = _dims(dims)
    states = np.ascontiguousarray(states, dtype=np.uint8)
    num = states.shape[0]
    pc = np.zeros(num, dtype=np.uint64) if per_item else None
    pn = np.zeros(num, dtype=np.uint64) if per_item else None
    res = MFResult()
    o, keep = _opts(options)
    check(lib().mf_lunnon_solve(dp, len(d), depth, _p(states, C.c_uint8), num, C.byref(o),
                                _p(pc, C.c_uint64), _p(pn, C.c_uint64), C.byref(res)), "mf_lunnon_solve")
    del keep
    return res, pc, pn


def gpu_folding(dims, depth: int, options: SolveOptions | None = None):
    """`gpu_folding` for any dims (square in the reference): (total, #states, MFResult)."""
    d, dp = _dims(dims)
    res = MFResult()
    ns = C.c_int64()
    o, keep = _opts(options)
    check(lib().mf_gpu_folding(dp, len(d), depth, C.byref(o), C.byref(res), C.byref(ns)), "mf_gpu_folding")
    return res.count, int(ns.value), res


def lunnon_batch(shapes, depths, options: SolveOptions | None = None):
    """Many 2-D maps in one queue: (per-map totals, MFResult)."""
    shapes = np.ascontiguousarray(np.asarray(shapes, dtype=np.int32).reshape(-1, 2))
    depths = np.ascontiguousarray(np.asarray(depths, dtype=np.int32))
    nm = shapes.shape[0]
    out = np.zeros(nm, dtype=np.uint64)
    res = MFResult()
    o, keep = _opts(options)
    check(lib().mf_lunnon_batch(_p(shapes, C.c_int32), _p(depths, C.c_int32), nm, C.byref(o),
                                _p(out, C.c_uint64), C.byref(res)), "mf_lunnon_batch")
    return [int(x) for x in out], res


# ----------------------------------------------------------------------------- sym

SYM_ORDERS = {"spiral": 0, "reverse": 1, "opposite": 2, "axes_first": 3, "corners_first": 4}


def sym_set_numbering(center: int = 0, order: str | int = "spiral", lock_center_n: int = 3) -> None:
    """Process-wide seed center (0 = the reference's default for n) and cell
    numbering of the sym entry points (sym_fold9's MF_CENTER_N, MF_ORDER,
    MF_LOCK_CENTER_N; experiments/9x9/sym_fold9.cpp:160-229,968-980).  An
    order name the reference does not know means the spiral, as there."""
    o = SYM_ORDERS.get(order, 0) if isinstance(order, str) else int(order)
    check(lib().mf_sym_set_numbering(int(center), o, int(lock_center_n)), "mf_sym_set_numbering")


def sym_get_numbering():
    """(center, order, lock_center_n) of the process-wide numbering."""
    c, o, lk = C.c_int32(), C.c_int32(), C.c_int32()
    lib().mf_sym_get_numbering(C.byref(c), C.byref(o), C.byref(lk))
    return int(c.value), int(o.value), int(lk.value)


class sym_numbering:
    """Context manager: the given numbering inside, the previous one after."""

    def __init__(self, center: int = 0, order: str | int = "spiral", lock_center_n: int = 3):
        self.want = (center, order, lock_center_n)

    def __enter__(self):
        self.prev = sym_get_numbering()
        sym_set_numbering(*self.want)
        return self

    def __exit__(self, *exc):
        sym_set_numbering(*self.prev)
        return False


def sym_seeds(n: int):
    """(cycles uint8 [S, case_bytes(n)], multipliers uint64 [S], raw center count)."""
    ns, raw = C.c_int64(), C.c_int64()
    check(lib().mf_sym_seeds(n, None, None, 0, C.byref(ns), C.byref(raw)), "mf_sym_seeds")
    cyc = np.zeros((ns.value, case_bytes(n)), dtype=np.uint8)
    mul = np.zeros(ns.value, dtype=np.uint64)
    check(lib().mf_sym_seeds(n, _p(cyc, C.c_uint8), _p(mul, C.c_uint64), ns.value, C.byref(ns), C.byref(raw)),
          "mf_sym_seeds")
    return cyc, mul, int(raw.value)


ALL_SEEDS = (1 << 64) - 1


def sym_cases(n: int, expanded: int, start: int, count: int, seed_start: int = 0, seed_count: int | None = None):
    """Cases [start, start+count) of the stream at level `expanded`: (cycles [k, case_bytes(n)], mults [k]).

    With seed_start / seed_count the stream is that of the seed slice
    (sym_fold9 MF_SEED_START / MF_SEED_COUNT): case 0 is seed seed_start's first."""
    cyc = np.zeros((count, case_bytes(n)), dtype=np.uint8)
    mul = np.zeros(count, dtype=np.uint64)
    got = C.c_int64()
    sc = ALL_SEEDS if seed_count is None else seed_count
    check(lib().mf_sym_cases_slice(n, expanded, seed_start, sc, start, count, _p(cyc, C.c_uint8),
                                   _p(mul, C.c_uint64), C.byref(got)), "mf_sym_cases_slice")
    return cyc[:got.value], mul[:got.value]


def sym_seed_case_counts(n: int, expanded: int, seed_start: int = 0, seed_count: int | None = None) -> np.ndarray:
    """Cases at level `expanded` under each seed of the slice (uint64 [seeds])."""
    sc = ALL_SEEDS if seed_count is None else seed_count
    num = C.c_int64()
    check(lib().mf_sym_seed_case_counts(n, expanded, seed_start, sc, None, 0, C.byref(num)),
          "mf_sym_seed_case_counts")
    out = np.zeros(num.value, dtype=np.uint64)
    check(lib().mf_sym_seed_case_counts(n, expanded, seed_start, sc, _p(out, C.c_uint64), num.value,
                                        C.byref(num)), "mf_sym_seed_case_counts")
    return out


def sym_case_total(n: int, expanded: int) -> int:
    t = C.c_uint64()
    check(lib().mf_sym_case_total(n, expanded, C.byref(t)), "mf_sym_case_total")
    return int(t.value)


def sym_solve(n: int, level: int, cycles: np.ndarray, mults=None, options: SolveOptions | None = None,
              per_item: bool = False):
    """Count cases on the GPU: (MFResult with count = sum mult*count, per-case counts, per-case nodes)."""
    cycles = np.ascontiguousarray(cycles, dtype=np.uint8)
    num = cycles.shape[0]
    if cycles.ndim != 2 or cycles.shape[1] != case_bytes(n):
        raise ValueError(f"cycles must be [num, {case_bytes(n)}] for n={n}, got {cycles.shape}")
    if mults is not None:
        mults = np.ascontiguousarray(mults, dtype=np.uint64)
        if mults.shape != (num,):
            raise ValueError("mults must have one entry per case")
    pc = np.zeros(num, dtype=np.uint64) if per_item else None
    pn = np.zeros(num, dtype=np.uint64) if per_item else None
    res = MFResult()
    o, keep = _opts(options)
    check(lib().mf_sym_solve(n, level, _p(cycles, C.c_uint8), _p(mults, C.c_uint64), num, C.byref(o),
                             _p(pc, C.c_uint64), _p(pn, C.c_uint64), C.byref(res)), "mf_sym_solve")
    del keep
    return res, pc, pn


def sym_fold(n: int, expanded: int = 24, start: int = 0, max_cases: int = 0,
             options: SolveOptions | None = None, seed_start: int = 0, seed_count: int | None = None):
    """The decimal answer of `sym_fold n threads max_cases start_case` (MF_STREAM_CASES);
    n = 9 and seed slices as `sym_fold9` (MF_SEED_START / MF_SEED_COUNT)."""
    ans = np.zeros(2, dtype=np.uint64)
    res = MFResult()
    o, keep = _opts(options)
    sc = ALL_SEEDS if seed_count is None else seed_count
    check(lib().mf_sym_fold_slice(n, expanded, seed_start, sc, start, max_cases, C.byref(o),
                                  _p(ans, C.c_uint64), C.byref(res)), "mf_sym_fold_slice")
    return int(ans[0]) + (int(ans[1]) << 64), res


# ----------------------------------------------------------------------------- resident frontiers


class _Frontier:
    def __init__(self, handle, num, items=None):
        self._h = handle
        self.num = num
        self.items = num if items is None else items   # resident work items (what an order permutes)

    def solve(self, options: SolveOptions | None = None, per_item: bool = False):
        if options is not None and options.order is not None and self.items is not None \
                and np.asarray(options.order).size != self.items:
            raise ValueError(f"order has {np.asarray(options.order).size} entries, the frontier {self.items} items")
        pc = np.zeros(self.num, dtype=np.uint64) if per_item else None
        pn = np.zeros(self.num, dtype=np.uint64) if per_item else None
        res = MFResult()
        o, keep = _opts(options)
        check(lib().mf_frontier_solve(self._h, C.byref(o), _p(pc, C.c_uint64), _p(pn, C.c_uint64),
                                      C.byref(res)), "mf_frontier_solve")
        del keep
        return res, pc, pn

    def close(self):
        if self._h:
            lib().mf_frontier_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SymFrontier(_Frontier):
    """Cases resident in HBM (solve repeatedly without re-uploading)."""

    def __init__(self, n: int, level: int, cycles: np.ndarray, mults=None):
        cycles = np.ascontiguousarray(cycles, dtype=np.uint8)
        if cycles.ndim != 2 or cycles.shape[1] != case_bytes(n):
            raise ValueError(f"cycles must be [num, {case_bytes(n)}] for n={n}, got {cycles.shape}")
        if mults is not None:
            mults = np.ascontiguousarray(mults, dtype=np.uint64)
            if mults.shape != (cycles.shape[0],):
                raise ValueError("mults must have one entry per case")
        h = C.c_void_p()
        check(lib().mf_sym_frontier_create(n, level, _p(cycles, C.c_uint8), _p(mults, C.c_uint64),
                                           cycles.shape[0], C.byref(h)), "mf_sym_frontier_create")
        super().__init__(h, cycles.shape[0])


class LunnonBatchFrontier(_Frontier):
    """Many 2-D maps, each with its own frontier depth, resident as one queue."""

    def __init__(self, shapes, depths):
        shapes = np.ascontiguousarray(np.asarray(shapes, dtype=np.int32).reshape(-1, 2))
        depths = np.ascontiguousarray(np.asarray(depths, dtype=np.int32))
        h = C.c_void_p()
        check(lib().mf_lunnon_batch_frontier_create(_p(shapes, C.c_int32), _p(depths, C.c_int32),
                                                    shapes.shape[0], C.byref(h)), "mf_lunnon_batch_frontier_create")
        super().__init__(h, 0, items=None)
        self.nmaps = shapes.shape[0]

    def solve_maps(self, options: SolveOptions | None = None):
        """(per-map exact totals, MFResult)."""
        out = np.zeros(self.nmaps, dtype=np.uint64)
        res = MFResult()
        o, keep = _opts(options)
        check(lib().mf_frontier_solve_maps(self._h, C.byref(o), _p(out, C.c_uint64), C.byref(res)),
              "mf_frontier_solve_maps")
        del keep
        return [int(x) for x in out], res


class LunnonFrontier(_Frontier):
    def __init__(self, dims, depth: int, states: np.ndarray):
        d, dp = _dims(dims)
        states = np.ascontiguousarray(states, dtype=np.uint8)
        h = C.c_void_p()
        check(lib().mf_lunnon_frontier_create(dp, len(d), depth, _p(states, C.c_uint8), states.shape[0],
                                              C.byref(h)), "mf_lunnon_frontier_create")
        super().__init__(h, states.shape[0])


def sym_expand(n: int, level: int, cycles: np.ndarray):
    """Children at level+1 of each case (ascending child order): (cycles, parent index)."""
    cycles = np.ascontiguousarray(cycles, dtype=np.uint8)
    written = C.c_int64()
    cap = max(1, cycles.shape[0] * 4)
    while True:
        out = np.zeros((cap, 64), dtype=np.uint8)
        par = np.zeros(cap, dtype=np.uint32)
        rc = lib().mf_sym_expand(n, level, _p(cycles, C.c_uint8), cycles.shape[0], _p(out, C.c_uint8),
                                 _p(par, C.c_uint32), cap, C.byref(written))
        if rc == _abi.MF_ERR_CAPACITY and written.value > cap:
            cap = written.value
            continue
        check(rc, "mf_sym_expand")
        return out[:written.value], par[:written.value]


class SymPieces(_Frontier):
    """Subtrees at mixed levels, each adding its count to one origin case.

    `num` (the length of per-item results) is the number of origins; `items`
    is the number of resident pieces (what an order permutes)."""

    def __init__(self, n: int, cycles: np.ndarray, levels, origins, mults, num_origins: int):
        cycles = np.ascontiguousarray(cycles, dtype=np.uint8)
        levels = np.ascontiguousarray(levels, dtype=np.uint8)
        origins = np.ascontiguousarray(origins, dtype=np.uint32)
        mults = np.ascontiguousarray(mults, dtype=np.uint64)
        h = C.c_void_p()
        check(lib().mf_sym_pieces_create(n, _p(cycles, C.c_uint8), _p(levels, C.c_uint8), _p(origins, C.c_uint32),
                                         _p(mults, C.c_uint64), cycles.shape[0], int(num_origins), C.byref(h)),
              "mf_sym_pieces_create")
        super().__init__(h, int(num_origins), items=cycles.shape[0])
