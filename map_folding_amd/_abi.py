"""ctypes binding of libmapfold.so (include/mapfold.h).

The library is the product: nothing here computes a count on the CPU.  If the
shared library has not been built, loading raises immediately; there is no
fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("MAPFOLD_LIB") or os.path.join(_HERE, "libmapfold.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "mapfold.h")
HEADER_PATHS = [HEADER_PATH, os.path.join(os.path.dirname(_HERE), "include", "mapfold_cost.h")]

MF_OK = 0
MF_ERR_ARG = 1
MF_ERR_DEPTH = 2
MF_ERR_HIP = 3
MF_ERR_NO_DEVICE = 4
MF_ERR_CAPACITY = 5
MF_ERR_INTERNAL = 6

MF_SCHED_STATIC = 0
MF_SCHED_RESPLIT = 1


class MFResult(C.Structure):
    _fields_ = [
        ("count_lo", C.c_uint64), ("count_hi", C.c_uint64),
        ("nodes", C.c_uint64), ("walks", C.c_uint64), ("steps", C.c_uint64),
        ("leaves", C.c_uint64), ("cands", C.c_uint64), ("items", C.c_uint64),
        ("rounds", C.c_uint32), ("status", C.c_uint32),
        ("kernel_ms", C.c_double), ("span_ms", C.c_double), ("total_ms", C.c_double),
        ("inputs", C.c_uint64),
    ]

    @property
    def count(self) -> int:
        return int(self.count_lo) + (int(self.count_hi) << 64)

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["count"] = self.count
        return d


class MFSolveOptions(C.Structure):
    _fields_ = [
        ("schedule", C.c_int32), ("block", C.c_int32), ("waves_per_cu", C.c_int32),
        ("budget", C.c_int32), ("tail", C.c_int32), ("count_events", C.c_int32),
        ("core", C.c_int32), ("graph", C.c_int32),
        ("queue_capacity", C.c_uint64), ("order", C.POINTER(C.c_uint32)), ("order_len", C.c_uint64),
        ("deepen", C.c_int32), ("deepen_levels", C.POINTER(C.c_uint8)), ("deepen_levels_len", C.c_uint64),
    ]


class MFCostTrainOptions(C.Structure):
    _fields_ = [
        ("epochs", C.c_int32), ("batch", C.c_int32), ("lr", C.c_float), ("beta1", C.c_float),
        ("beta2", C.c_float), ("eps", C.c_float), ("seed", C.c_uint64),
    ]


MF_COST_FEATURES = 80
MF_COST_HIDDEN = 128
MF_COST_PARAMS = 43521


class MFError(RuntimeError):
    def __init__(self, status: int, where: str, message: str):
        super().__init__(f"{where}: status {status} ({message})")
        self.status = status


_lib = None


def lib() -> C.CDLL:
    """Load libmapfold.so; raise if it was not built (no CPU fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                          "or `make -C map_folding_amd/csrc`")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    i32, i64, u32, u64, u8, f32 = C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_uint8, C.c_float
    sig = {
        "mf_status_string": (C.c_char_p, [C.c_int]),
        "mf_last_error": (C.c_char_p, []),
        "mf_init": (C.c_int, [C.c_int]),
        "mf_device_cus": (C.c_int, []),
        "mf_shutdown": (None, []),
        "mf_lunnon_frontier": (C.c_int, [P(i32), i32, i32, P(u8), i64, P(i64)]),
        "mf_lunnon_solve": (C.c_int, [P(i32), i32, i32, P(u8), i64, P(MFSolveOptions), P(u64), P(u64),
                                      P(MFResult)]),
        "mf_gpu_folding": (C.c_int, [P(i32), i32, i32, P(MFSolveOptions), P(MFResult), P(i64)]),
        "mf_lunnon_batch": (C.c_int, [P(i32), P(i32), i32, P(MFSolveOptions), P(u64), P(MFResult)]),
        "mf_sym_seeds": (C.c_int, [i32, P(u8), P(u64), i64, P(i64), P(i64)]),
        "mf_sym_set_numbering": (C.c_int, [i32, i32, i32]),
        "mf_sym_get_numbering": (None, [P(i32), P(i32), P(i32)]),
        "mf_lunnon_check_states": (C.c_int, [P(i32), i32, i32, P(u8), i64, P(i64)]),
        "mf_sym_cases": (C.c_int, [i32, i32, u64, u64, P(u8), P(u64), P(i64)]),
        "mf_sym_case_total": (C.c_int, [i32, i32, P(u64)]),
        "mf_sym_cases_slice": (C.c_int, [i32, i32, u64, u64, u64, u64, P(u8), P(u64), P(i64)]),
        "mf_sym_seed_case_counts": (C.c_int, [i32, i32, u64, u64, P(u64), i64, P(i64)]),
        "mf_sym_fold_slice": (C.c_int, [i32, i32, u64, u64, u64, u64, P(MFSolveOptions), P(u64), P(MFResult)]),
        "mf_sym_solve": (C.c_int, [i32, i32, P(u8), P(u64), i64, P(MFSolveOptions), P(u64), P(u64),
                                   P(MFResult)]),
        "mf_sym_fold": (C.c_int, [i32, i32, u64, u64, P(MFSolveOptions), P(u64), P(MFResult)]),
        "mf_sym_solve_noalloc": (C.c_int, [i32, i32, P(u8), P(u64), i64, P(MFSolveOptions), P(u64), P(u64),
                                           P(MFResult)]),
        "mf_sym_frontier_create": (C.c_int, [i32, i32, P(u8), P(u64), i64, P(C.c_void_p)]),
        "mf_lunnon_frontier_create": (C.c_int, [P(i32), i32, i32, P(u8), i64, P(C.c_void_p)]),
        "mf_frontier_solve": (C.c_int, [C.c_void_p, P(MFSolveOptions), P(u64), P(u64), P(MFResult)]),
        "mf_lunnon_batch_frontier_create": (C.c_int, [P(i32), P(i32), i32, P(C.c_void_p)]),
        "mf_frontier_solve_maps": (C.c_int, [C.c_void_p, P(MFSolveOptions), P(u64), P(MFResult)]),
        "mf_wasm_input_ptr": (P(u8), []),
        "mf_wasm_output_ptr": (P(u32), []),
        "mf_wasm_case_stride": (u32, []),
        "mf_wasm_max_cases": (u32, []),
        "mf_wasm_solve_cases": (u32, [u32, u32, u32, u32]),
        "mf_frontier_destroy": (None, [C.c_void_p]),
        "mf_sym_expand": (C.c_int, [i32, i32, P(u8), i64, P(u8), P(u32), i64, P(i64)]),
        "mf_sym_pieces_create": (C.c_int, [i32, P(u8), P(u8), P(u32), P(u64), i64, i64, P(C.c_void_p)]),
        # learned cost model (include/mapfold_cost.h)
        "mf_cost_create": (C.c_int, [u64, P(C.c_void_p)]),
        "mf_cost_destroy": (None, [C.c_void_p]),
        "mf_cost_get_params": (C.c_int, [C.c_void_p, P(f32)]),
        "mf_cost_set_params": (C.c_int, [C.c_void_p, P(f32)]),
        "mf_cost_features_sym": (C.c_int, [i32, i32, P(u8), i64, P(f32)]),
        "mf_cost_features_lunnon": (C.c_int, [P(i32), i32, i32, P(u8), i64, P(f32)]),
        "mf_cost_gradient": (C.c_int, [C.c_void_p, P(f32), P(f32), i64, P(f32), P(f32)]),
        "mf_cost_train": (C.c_int, [C.c_void_p, P(f32), P(f32), i64, P(MFCostTrainOptions), P(f32)]),
        "mf_sym_estimate_work": (C.c_int, [i32, i32, P(u8), i64, i32, P(u64)]),
        "mf_cost_predict": (C.c_int, [C.c_void_p, P(f32), i64, P(f32)]),
        "mf_cost_predict_fp32": (C.c_int, [C.c_void_p, P(f32), i64, P(f32)]),
        "mf_cost_predict_frontier": (C.c_int, [C.c_void_p, C.c_void_p, P(f32), i64, P(i64)]),
        "mf_cost_features_frontier": (C.c_int, [C.c_void_p, P(f32), i64, P(i64)]),
        "mf_frontier_items": (u64, [C.c_void_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(status: int, where: str) -> None:
    if status != MF_OK:
        raise MFError(status, where, lib().mf_last_error().decode(errors="replace"))


def declared_symbols() -> list[str]:
    """Function names declared in include/mapfold.h and include/mapfold_cost.h."""
    import re
    text = "\n".join(open(p).read() for p in HEADER_PATHS)
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mf_[a-z0-9_]+)\s*\(", text)))
