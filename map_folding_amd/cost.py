"""Learned subtree-cost model and the schedules it drives (include/mapfold_cost.h).

The model predicts log2(1 + nodes) of a work item's DFS subtree.  It is used
only to order and split work: every count comes from the DFS kernels and is
the same under any order.  Training and inference run on the GPU in the
library's own kernels (FP32 forward/backward + Adam; BF16 MFMA inference).

Reference counterparts: the lookahead cost proxy `estimate_work` and the case
sort `score_and_sort_cases` (sym_fold.cpp:235-250, 267-314), and the per-case
node profiles of `MF_PROFILE_WORK` (experiments/9x9/sym_fold9.cpp:785-848)
that supply the training pairs.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import MF_COST_FEATURES, MF_COST_PARAMS, MFCostTrainOptions, check, lib

__all__ = [
    "CostModel", "features_sym", "features_lunnon", "node_target", "learned_order", "lookahead_order",
    "estimate_work", "sort_by_score", "wave_waste", "list_makespan", "spearman",
]


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and a.shape[1:] != shape:
        raise ValueError(f"expected rows of shape {shape}, got {a.shape}")
    return a


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def features_sym(n: int, level: int, cycles: np.ndarray) -> np.ndarray:
    """Feature rows of sym cases at `level` (64 cycle bytes per case)."""
    cycles = np.ascontiguousarray(cycles, dtype=np.uint8)
    out = np.zeros((cycles.shape[0], MF_COST_FEATURES), dtype=np.float32)
    check(lib().mf_cost_features_sym(n, level, cycles.ctypes.data_as(C.POINTER(C.c_uint8)), cycles.shape[0],
                                     _fp(out)), "mf_cost_features_sym")
    return out


def features_lunnon(dims, depth: int, states: np.ndarray) -> np.ndarray:
    """Feature rows of Lunnon frontier states b[0..depth]."""
    d = np.ascontiguousarray(np.asarray(dims, dtype=np.int32))
    states = np.ascontiguousarray(states, dtype=np.uint8)
    out = np.zeros((states.shape[0], MF_COST_FEATURES), dtype=np.float32)
    check(lib().mf_cost_features_lunnon(d.ctypes.data_as(C.POINTER(C.c_int32)), len(d), depth,
                                        states.ctypes.data_as(C.POINTER(C.c_uint8)), states.shape[0], _fp(out)),
          "mf_cost_features_lunnon")
    return out


def node_target(nodes) -> np.ndarray:
    """Regression target: log2(1 + subtree nodes)."""
    return np.log2(1.0 + np.asarray(nodes, dtype=np.float64)).astype(np.float32)


class CostModel:
    """Handle on a model held in device memory."""

    def __init__(self, seed: int = 0):
        h = C.c_void_p()
        check(lib().mf_cost_create(C.c_uint64(seed), C.byref(h)), "mf_cost_create")
        self._h = h

    # -- parameters
    @property
    def params(self) -> np.ndarray:
        p = np.zeros(MF_COST_PARAMS, dtype=np.float32)
        check(lib().mf_cost_get_params(self._h, _fp(p)), "mf_cost_get_params")
        return p

    @params.setter
    def params(self, p) -> None:
        p = _f32(p)
        if p.shape != (MF_COST_PARAMS,):
            raise ValueError(f"expected {MF_COST_PARAMS} parameters")
        check(lib().mf_cost_set_params(self._h, _fp(p)), "mf_cost_set_params")

    def save(self, path: str) -> None:
        np.save(path, self.params)

    def load(self, path: str) -> None:
        self.params = np.load(path, allow_pickle=False)

    # -- training
    def gradient(self, x, y):
        """FP32 loss and gradient over all rows (one forward + backward on the GPU)."""
        x = _f32(x, (MF_COST_FEATURES,))
        y = _f32(y)
        g = np.zeros(MF_COST_PARAMS, dtype=np.float32)
        loss = C.c_float()
        check(lib().mf_cost_gradient(self._h, _fp(x), _fp(y), x.shape[0], _fp(g), C.byref(loss)),
              "mf_cost_gradient")
        return float(loss.value), g

    def train(self, x, y, epochs: int = 20, batch: int = 512, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        x = _f32(x, (MF_COST_FEATURES,))
        y = _f32(y)
        o = MFCostTrainOptions()
        o.epochs, o.batch, o.lr, o.seed = epochs, batch, lr, seed
        losses = np.zeros(epochs, dtype=np.float32)
        check(lib().mf_cost_train(self._h, _fp(x), _fp(y), x.shape[0], C.byref(o), _fp(losses)), "mf_cost_train")
        return losses

    # -- inference
    def predict(self, x) -> np.ndarray:
        """BF16 MFMA inference."""
        x = _f32(x, (MF_COST_FEATURES,))
        out = np.zeros(x.shape[0], dtype=np.float32)
        check(lib().mf_cost_predict(self._h, _fp(x), x.shape[0], _fp(out)), "mf_cost_predict")
        return out

    def predict_fp32(self, x) -> np.ndarray:
        x = _f32(x, (MF_COST_FEATURES,))
        out = np.zeros(x.shape[0], dtype=np.float32)
        check(lib().mf_cost_predict_fp32(self._h, _fp(x), x.shape[0], _fp(out)), "mf_cost_predict_fp32")
        return out

    def close(self) -> None:
        if self._h:
            lib().mf_cost_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------------------------------------- schedules


def learned_order(pred) -> np.ndarray:
    """Largest predicted subtree first (stable)."""
    return np.argsort(-np.asarray(pred, dtype=np.float64), kind="stable").astype(np.uint32)


def estimate_work(n: int, level: int, cycles: np.ndarray, depth: int) -> np.ndarray:
    """The reference's lookahead proxy: descendants at level min(n^2, level+depth)
    of each case (estimate_work, sym_fold.cpp:235-250)."""
    cycles = np.ascontiguousarray(cycles, dtype=np.uint8)
    out = np.zeros(cycles.shape[0], dtype=np.uint64)
    check(lib().mf_sym_estimate_work(n, level, cycles.ctypes.data_as(C.POINTER(C.c_uint8)), cycles.shape[0], depth,
                                     out.ctypes.data_as(C.POINTER(C.c_uint64))), "mf_sym_estimate_work")
    return out


def sort_by_score(scores, mults, ascending: bool = False) -> np.ndarray:
    """score_and_sort_cases order (sym_fold.cpp:295-305): by score (descending
    unless `ascending`), ties by multiplier ascending, then by case index."""
    s = np.asarray(scores, dtype=np.float64)
    m = np.asarray(mults, dtype=np.float64)
    key = s if ascending else -s
    return np.lexsort((np.arange(len(s)), m, key)).astype(np.uint32)


def lookahead_order(scores, mults=None) -> np.ndarray:
    """The no-model baseline order: the lookahead proxy, descending."""
    if mults is None:
        mults = np.zeros(len(scores))
    return sort_by_score(scores, mults)


def wave_waste(nodes_in_order, lanes: int = 64) -> float:
    """Lane-slots per useful node when consecutive items share a wave and a wave
    lasts as long as its largest item (the static one-lane-per-item kernel):
    sum over waves of lanes * max / sum of nodes.  1.0 is perfect."""
    a = np.asarray(nodes_in_order, dtype=np.float64)
    pad = (-len(a)) % lanes
    if pad:
        a = np.concatenate([a, np.zeros(pad)])
    w = a.reshape(-1, lanes)
    return float((w.max(axis=1) * lanes).sum() / max(a.sum(), 1.0))


def list_makespan(nodes_in_order, workers: int) -> float:
    """Greedy list schedule of items, in the given order, on `workers` equal
    workers (each item to the earliest-free worker); makespan over the ideal
    sum/workers.  1.0 is perfect."""
    import heapq
    a = np.asarray(nodes_in_order, dtype=np.float64)
    heap = [0.0] * workers
    for v in a:
        t = heapq.heappop(heap)
        heapq.heappush(heap, t + v)
    return float(max(heap) / max(a.sum() / workers, 1e-300))


def spearman(a, b) -> float:
    ra = np.argsort(np.argsort(np.asarray(a, dtype=np.float64), kind="stable"), kind="stable").astype(np.float64)
    rb = np.argsort(np.argsort(np.asarray(b, dtype=np.float64), kind="stable"), kind="stable").astype(np.float64)
    ra -= ra.mean()
    rb -= rb.mean()
    return float((ra * rb).sum() / np.sqrt((ra * ra).sum() * (rb * rb).sum()))
