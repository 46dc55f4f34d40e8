"""ctypes front-end for the CPU oracle (oracle/mf_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, and there only as the checker.  The product
package (map_folding_amd) never imports this module.

Each wrapper names the reference function it restates; the C file carries the
file:line citations.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


class LunnonStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in
                ("inserts", "gapcomps", "walk_steps", "candidates", "completions", "peak_gap")]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class SymStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("nodes", "page_walks", "walk_steps", "leaves")]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


def build():
    """Compile liboracle.so (gcc; seconds)."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "liboracle.so"], check=True)


def build_ref():
    """Compile the reference's own CPU solvers into oracle/_ref (needs the reference tree)."""
    subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.POINTER
        L.orc_lunnon_table.argtypes = [P(C.c_int), C.c_int, P(C.c_int16), C.c_int]
        L.orc_lunnon_count.argtypes = [P(C.c_int), C.c_int, P(C.c_uint64), P(LunnonStats)]
        L.orc_lunnon_frontier.argtypes = [P(C.c_int), C.c_int, C.c_int, P(C.c_uint8), C.c_int64,
                                          P(LunnonStats)]
        L.orc_lunnon_frontier.restype = C.c_int64
        L.orc_lunnon_resume.argtypes = [P(C.c_int), C.c_int, C.c_int, P(C.c_uint8), C.c_int64,
                                        P(C.c_uint64), P(C.c_uint64), P(C.c_uint64), P(LunnonStats)]
        L.orc_sym_edges.argtypes = [C.c_int, P(C.c_uint8), P(C.c_int)]
        L.orc_sym_count_case.argtypes = [C.c_int, P(C.c_uint8), C.c_int, P(C.c_uint64), P(SymStats)]
        L.orc_sym_seeds.argtypes = [C.c_int, P(C.c_uint8), P(C.c_uint64), C.c_int64, P(C.c_int64)]
        L.orc_sym_seeds.restype = C.c_int64
        L.orc_sym_cases.argtypes = [C.c_int, C.c_int, C.c_uint64, C.c_uint64, P(C.c_uint8),
                                    P(C.c_uint64), P(C.c_uint64)]
        L.orc_sym_cases.restype = C.c_int64
        L.orc_sym_solve.argtypes = [C.c_int, C.c_int, P(C.c_uint8), P(C.c_uint64), C.c_int64, C.c_int,
                                    P(C.c_uint64), P(C.c_uint64), P(C.c_uint64), P(SymStats)]
        L.orc_sym_answer.argtypes = [C.c_int, C.c_int, C.c_uint64, C.c_uint64, C.c_int,
                                     P(C.c_uint64), P(SymStats)]
        L.orc9_default_center.argtypes = [C.c_int]
        L.orc9_seeds.argtypes = [C.c_int, C.c_int, P(C.c_uint8), P(C.c_uint64), C.c_int64, P(C.c_int64)]
        L.orc9_seeds.restype = C.c_int64
        L.orc9_cases.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_uint64, P(C.c_uint8),
                                 P(C.c_uint64)]
        L.orc9_cases.restype = C.c_int64
        L.orc9_solve.argtypes = [C.c_int, C.c_int, P(C.c_uint8), P(C.c_uint64), C.c_int64, C.c_int,
                                 P(C.c_uint64), P(C.c_uint64), C.c_uint64 * 2, P(Sym9Stats)]
        _lib = L
    return _lib


def _ptr(arr, ctype):
    if arr is None:
        return None
    return arr.ctypes.data_as(C.POINTER(ctype))


def _dims(dims):
    d = np.ascontiguousarray(np.asarray(dims, dtype=np.int32))
    return d, _ptr(d, C.c_int)


# ----------------------------------------------------------------------------- Lunnon


def lunnon_table(dims):
    """d[i][l][m] of mf.c:29-75 as an int16 array [ndim, n+1, n+1] (m = 0 -> 0)."""
    d, dp = _dims(dims)
    n = int(np.prod(d))
    out = np.zeros((len(d), n + 1, n + 1), dtype=np.int16)
    rc = lib().orc_lunnon_table(dp, len(d), _ptr(out, C.c_int16), out.size)
    if rc:
        raise ValueError(f"bad dims {dims}: rc={rc}")
    return out


def lunnon_count(dims):
    """Total of `mf p q ...` (mf.c:165-205) and the full-tree event counters."""
    d, dp = _dims(dims)
    tot = np.zeros(2, dtype=np.uint64)
    st = LunnonStats()
    rc = lib().orc_lunnon_count(dp, len(d), _ptr(tot, C.c_uint64), C.byref(st))
    if rc:
        raise ValueError(f"bad dims {dims}: rc={rc}")
    return int(tot[0]) + (int(tot[1]) << 64), st.as_dict()


def lunnon_frontier(dims, depth, max_states=None):
    """Canonical frontier at partition depth D (mf_fast.cpp:185-221): uint8 [S, D+1] rows
    of b[0..D] in DFS order, plus the counters of the levels <= D."""
    d, dp = _dims(dims)
    st = LunnonStats()
    count = lib().orc_lunnon_frontier(dp, len(d), depth, None, 0, C.byref(st))
    if count < 0:
        raise ValueError(f"bad frontier request dims={dims} depth={depth}: {count}")
    if max_states is not None:
        count = min(count, max_states)
    out = np.zeros((count, depth + 1), dtype=np.uint8)
    lib().orc_lunnon_frontier(dp, len(d), depth, _ptr(out, C.c_uint8), count, C.byref(st))
    return out, st.as_dict()


def lunnon_resume(dims, depth, states):
    """solve_from_state (mf_fast.cpp:142-183) for each row of `states`.
    Returns (per-state counts, per-state inserts, total, summed counters)."""
    d, dp = _dims(dims)
    states = np.ascontiguousarray(states, dtype=np.uint8)
    num = states.shape[0]
    cnt = np.zeros(num, dtype=np.uint64)
    nodes = np.zeros(num, dtype=np.uint64)
    tot = np.zeros(2, dtype=np.uint64)
    st = LunnonStats()
    rc = lib().orc_lunnon_resume(dp, len(d), depth, _ptr(states, C.c_uint8), num, _ptr(cnt, C.c_uint64),
                                 _ptr(nodes, C.c_uint64), _ptr(tot, C.c_uint64), C.byref(st))
    if rc:
        raise ValueError(f"resume failed rc={rc}")
    return cnt, nodes, int(tot[0]) + (int(tot[1]) << 64), st.as_dict()


# ----------------------------------------------------------------------------- sym


def sym_edges(n):
    """4 edge pages (make_edges, sym_fold.cpp:147-173) as uint8 [4, 64] and the spiral map [n, n]."""
    out = np.zeros((4, 64), dtype=np.uint8)
    mp = np.zeros(n * n, dtype=np.int32)
    rc = lib().orc_sym_edges(n, _ptr(out, C.c_uint8), _ptr(mp, C.c_int))
    if rc:
        raise ValueError(f"bad n {n}")
    return out, mp.reshape(n, n)


def sym_count_case(n, cycle, i):
    """recurse_count (sym_fold.cpp:219-233) of one case, plus node / walk counters."""
    cyc = np.zeros(64, dtype=np.uint8)
    cyc[:i] = np.asarray(cycle, dtype=np.uint8)[:i]
    c = C.c_uint64()
    st = SymStats()
    rc = lib().orc_sym_count_case(n, _ptr(cyc, C.c_uint8), i, C.byref(c), C.byref(st))
    if rc:
        raise ValueError(f"bad case rc={rc}")
    return int(c.value), st.as_dict()


def sym_seeds(n):
    """Deduplicated center cycles (sym_fold.cpp:689-710): (cycles [S,64] uint8, multipliers, raw count)."""
    raw = C.c_int64()
    ns = lib().orc_sym_seeds(n, None, None, 0, C.byref(raw))
    if ns < 0:
        raise ValueError(f"bad n {n}")
    cyc = np.zeros((ns, 64), dtype=np.uint8)
    mul = np.zeros(ns, dtype=np.uint64)
    lib().orc_sym_seeds(n, _ptr(cyc, C.c_uint8), _ptr(mul, C.c_uint64), ns, C.byref(raw))
    return cyc, mul, int(raw.value)


def sym_cases(n, expanded, start, max_cases):
    """Expanded cases [start, start+max_cases) in stream order (sym_fold.cpp:491-527)."""
    cyc = np.zeros((max_cases, 64), dtype=np.uint8)
    mul = np.zeros(max_cases, dtype=np.uint64)
    got = lib().orc_sym_cases(n, expanded, start, max_cases, _ptr(cyc, C.c_uint8), _ptr(mul, C.c_uint64), None)
    if got < 0:
        raise ValueError(f"bad request n={n} expanded={expanded}: {got}")
    return cyc[:got], mul[:got]


def sym_case_total(n, expanded):
    """Number of expanded cases at depth `expanded` (whole stream)."""
    tot = C.c_uint64()
    got = lib().orc_sym_cases(n, expanded, 0, 0, None, None, C.byref(tot))
    if got < 0:
        raise ValueError("bad request")
    return int(tot.value)


def sym_solve(n, i, cycles, mults=None, threads=None):
    """Count cases at level i (solve_cases_parallel, sym_fold.cpp:410-489).
    Returns (per-case counts, per-case nodes, sum mult*count (not times n2), counters)."""
    cycles = np.ascontiguousarray(cycles, dtype=np.uint8)
    num = cycles.shape[0]
    if mults is not None:
        mults = np.ascontiguousarray(mults, dtype=np.uint64)
    if threads is None:
        threads = os.cpu_count() or 1
    cnt = np.zeros(num, dtype=np.uint64)
    nodes = np.zeros(num, dtype=np.uint64)
    tot = np.zeros(2, dtype=np.uint64)
    st = SymStats()
    rc = lib().orc_sym_solve(n, i, _ptr(cycles, C.c_uint8), _ptr(mults, C.c_uint64), num, threads,
                             _ptr(cnt, C.c_uint64), _ptr(nodes, C.c_uint64), _ptr(tot, C.c_uint64),
                             C.byref(st))
    if rc:
        raise ValueError(f"solve failed rc={rc}")
    return cnt, nodes, int(tot[0]) + (int(tot[1]) << 64), st.as_dict()


def sym_answer(n, expanded=24, start=0, max_cases=0, threads=None):
    """The decimal answer `sym_fold n threads max start` prints (sym_fold.cpp:649-774)."""
    if threads is None:
        threads = os.cpu_count() or 1
    tot = np.zeros(2, dtype=np.uint64)
    st = SymStats()
    rc = lib().orc_sym_answer(n, expanded, start, max_cases, threads, _ptr(tot, C.c_uint64), C.byref(st))
    if rc:
        raise ValueError(f"answer failed rc={rc}")
    return int(tot[0]) + (int(tot[1]) << 64), st.as_dict()


def sym_count_to(n, cycle, start, stop):
    """sym_kernel.c recurse_count (mfah/wasm/sym_kernel.c:212-230): (paths start -> stop, nodes)."""
    L = lib()
    L.orc_sym_count_to.argtypes = [C.c_int, C.POINTER(C.c_uint8), C.c_int, C.c_int,
                                   C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    cyc = np.zeros(64, dtype=np.uint8)
    cyc[:start] = np.asarray(cycle, dtype=np.uint8)[:start]
    c, nd = C.c_uint64(), C.c_uint64()
    rc = L.orc_sym_count_to(n, _ptr(cyc, C.c_uint8), start, stop, C.byref(c), C.byref(nd))
    if rc:
        raise ValueError(f"bad request rc={rc}")
    return int(c.value), int(nd.value)


def sym_prefixes(n, depth):
    """All cycles at `depth` grown from the initial cycle 0->1->2->0 (uint8 [k, 64])."""
    L = lib()
    L.orc_sym_prefixes.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_uint8), C.c_int64]
    L.orc_sym_prefixes.restype = C.c_int64
    k = L.orc_sym_prefixes(n, depth, None, 0)
    if k < 0:
        raise ValueError("bad request")
    out = np.zeros((k, 64), dtype=np.uint8)
    L.orc_sym_prefixes(n, depth, _ptr(out, C.c_uint8), k)
    return out


# ----------------------------------------------------------------------------- 9x9 (sym_fold9.cpp)


class Sym9Stats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("nodes", "page_walks", "walk_steps", "leaves")]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


CELLS9 = 128   # bytes per 9x9-capable cycle row


def sym9_default_center(n):
    return int(lib().orc9_default_center(n))


def sym9_seeds(n, c=None):
    """(rows uint8 [k, 128], multipliers, raw center count) -- sym_fold9.cpp center dedup."""
    L = lib()
    c = sym9_default_center(n) if c is None else c
    raw = C.c_int64()
    k = L.orc9_seeds(n, c, None, None, C.c_int64(0), C.byref(raw))
    if k < 0:
        raise ValueError("bad n / center")
    out = np.zeros((k, CELLS9), dtype=np.uint8)
    mul = np.zeros(k, dtype=np.uint64)
    L.orc9_seeds(n, c, out.ctypes.data_as(C.POINTER(C.c_uint8)), mul.ctypes.data_as(C.POINTER(C.c_uint64)),
                 C.c_int64(k), C.byref(raw))
    return out, mul, int(raw.value)


def sym9_cases(n, expanded, start, max_cases, c=None):
    L = lib()
    c = sym9_default_center(n) if c is None else c
    out = np.zeros((max_cases, CELLS9), dtype=np.uint8)
    mul = np.zeros(max_cases, dtype=np.uint64)
    k = L.orc9_cases(n, c, expanded, C.c_uint64(start), C.c_uint64(max_cases),
                     out.ctypes.data_as(C.POINTER(C.c_uint8)), mul.ctypes.data_as(C.POINTER(C.c_uint64)))
    if k < 0:
        raise ValueError("bad arguments")
    return out[:k], mul[:k]


def sym9_solve(n, level, cycles, mults=None, threads=8):
    """(per-case counts, per-case nodes, sum mult*count (int), stats)."""
    L = lib()
    cycles = np.ascontiguousarray(cycles, dtype=np.uint8)
    k = cycles.shape[0]
    pc = np.zeros(k, dtype=np.uint64)
    pn = np.zeros(k, dtype=np.uint64)
    tot = (C.c_uint64 * 2)()
    st = Sym9Stats()
    mp = None
    if mults is not None:
        mults = np.ascontiguousarray(mults, dtype=np.uint64)
        mp = mults.ctypes.data_as(C.POINTER(C.c_uint64))
    rc = L.orc9_solve(n, level, cycles.ctypes.data_as(C.POINTER(C.c_uint8)), mp, C.c_int64(k), threads,
                      pc.ctypes.data_as(C.POINTER(C.c_uint64)), pn.ctypes.data_as(C.POINTER(C.c_uint64)), tot,
                      C.byref(st))
    if rc != 0:
        raise ValueError(f"orc9_solve rc={rc}")
    return pc, pn, int(tot[0]) | (int(tot[1]) << 64), st.as_dict()


def sym9_answer(n, expanded, start, max_cases, c=None, threads=8):
    """sym_fold9 <n> threads max start with MF_STREAM_CASES=1 MF_EXPANDED_N=expanded."""
    cyc, mul = sym9_cases(n, expanded, start, max_cases, c)
    _, _, tot, st = sym9_solve(n, expanded, cyc, mul, threads)
    return tot * n * n, st
