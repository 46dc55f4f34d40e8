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
    "CostModel", "features_sym", "features_frontier", "features_lunnon", "node_target", "learned_order", "class_order", "lookahead_order",
    "learned_levels",
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


def features_frontier(frontier) -> np.ndarray:
    """Feature rows of every item of a resident sym frontier, computed on the
    device (the values of features_sym)."""
    n = int(lib().mf_frontier_items(frontier._h))
    out = np.zeros((max(n, 1), MF_COST_FEATURES), dtype=np.float32)
    rows = C.c_int64()
    check(lib().mf_cost_features_frontier(frontier._h, _fp(out), out.shape[0], C.byref(rows)),
          "mf_cost_features_frontier")
    return out[:rows.value]


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

    def predict_frontier(self, frontier) -> np.ndarray:
        """Predictions for every item of a resident sym frontier, features and
        BF16 inference both on the device."""
        n = int(lib().mf_frontier_items(frontier._h))
        out = np.zeros(max(n, 1), dtype=np.float32)
        rows = C.c_int64()
        check(lib().mf_cost_predict_frontier(self._h, frontier._h, _fp(out), out.shape[0], C.byref(rows)),
              "mf_cost_predict_frontier")
        return out[:rows.value]

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


def class_order(pred, levels, width: float = 2.0) -> np.ndarray:
    """Largest size class first, where a class spans `width` units of predicted
    log2 nodes; inside a class, by level, then in the given order.

    The given order of split pieces is the splitter's: level by level, siblings
    next to each other. Keeping siblings together inside a class puts subtrees
    of the same depth and a similar shape in one wave, so its lanes' walks and
    backtracking depths differ less than under a strict largest-first sort
    (8x8 depth-36 chunk of 50,000 cases, 3.77M pieces: 962-964 ms at width 2
    or 3, 968 ms at width 8, 987 ms strictly largest first, 982 ms in the
    splitter's order alone)."""
    p = np.asarray(pred, dtype=np.float64)
    lv = np.asarray(levels, dtype=np.int64)
    return np.lexsort((np.arange(len(p)), lv, -np.floor(p / width))).astype(np.uint32)


def learned_levels(pred, base: int, theta: float, cap: int) -> np.ndarray:
    """Per-case deepening targets from predicted log2(1 + nodes): every case
    goes to `base`; a case predicted above `theta` goes one level deeper per
    unit of log2 nodes above it, at most `cap` levels more.  The large
    subtrees then enter the rounds as more, smaller pieces (8x8 depth-36 chunk
    of 50,000 cases, base 46: 902.5 ms at theta 21 / cap 4, 907.5 ms with the
    uniform target)."""
    p = np.asarray(pred, dtype=np.float64)
    return (base + np.clip(np.floor(p - theta), 0, cap)).astype(np.uint8)


def estimate_work(n: int, level: int, cycles: np.ndarray, depth: int) -> np.ndarray:
    """The reference's lookahead proxy: descendants at level min(n^2, level+depth)
    of each case (estimate_work, sym_fold.cpp:235-250); rows of case_bytes(n)."""
    from . import case_bytes
    cycles = np.ascontiguousarray(cycles, dtype=np.uint8)
    if cycles.ndim != 2 or cycles.shape[1] != case_bytes(n):
        raise ValueError(f"cycles must be [num, {case_bytes(n)}] for n={n}")
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


# ----------------------------------------------------------------------------- split schedules


class Pieces:
    """Subtrees at mixed levels that together cover a list of cases.

    cycles [P, 64], levels [P], origins [P] (case index), mults [P],
    score [P] (predicted log2 size, or the proxy used to split), host_nodes:
    nodes expanded on the host while splitting (the parents of split pieces)."""

    def __init__(self, cycles, levels, origins, mults, score, host_nodes, num_origins):
        self.cycles, self.levels, self.origins, self.mults = cycles, levels, origins, mults
        self.score, self.host_nodes, self.num_origins = score, host_nodes, num_origins

    def __len__(self):
        return len(self.levels)

    def frontier(self, n: int, ordered: bool = True):
        """Resident frontier of the pieces; `ordered` lays them out in
        `order()` once, so solves need no per-call order permutation."""
        from . import SymPieces
        if not ordered:
            return SymPieces(n, self.cycles, self.levels, self.origins, self.mults, self.num_origins)
        o = self.order()
        return SymPieces(n, self.cycles[o], self.levels[o], self.origins[o], self.mults[o], self.num_origins)

    def order(self) -> np.ndarray:
        """Largest size class first, same-level siblings together (class_order)."""
        return class_order(self.score, self.levels)


def split_cases(n: int, level: int, cycles, mults, score_fn, threshold: float, max_extra: int = 8,
                max_pieces: int = 4_000_000) -> Pieces:
    """Replace every case whose score exceeds `threshold` by its children,
    level by level (at most `max_extra` levels, never past level n^2 - 2),
    until no piece scores above it or the piece budget is reached.

    score_fn(level, cycles) -> scores (e.g. a model's predicted log2 nodes).
    Counts are unaffected: a node's subtree is the union of its children's."""
    from . import sym_expand
    cycles = np.ascontiguousarray(cycles, dtype=np.uint8)
    K = cycles.shape[0]
    cur = (cycles, np.arange(K, dtype=np.uint32), np.asarray(mults, dtype=np.uint64))
    done = []
    host_nodes = 0
    total = K
    lvl = level
    for extra in range(max_extra + 1):
        c, o, m = cur
        if c.shape[0] == 0:
            break
        s = np.asarray(score_fn(lvl, c), dtype=np.float32)
        split = s > threshold
        if extra == max_extra or lvl + 1 > n * n - 2:
            split[:] = False
        keep = ~split
        done.append((c[keep], np.full(int(keep.sum()), lvl, np.uint8), o[keep], m[keep], s[keep]))
        if not split.any():
            break
        ch, par = sym_expand(n, lvl, c[split])
        if total - int(split.sum()) + len(ch) > max_pieces:
            done.append((c[split], np.full(int(split.sum()), lvl, np.uint8), o[split], m[split], s[split]))
            break
        host_nodes += int(split.sum())
        total += len(ch) - int(split.sum())
        cur = (ch, o[split][par], m[split][par])
        lvl += 1
    cat = lambda i: np.concatenate([d[i] for d in done]) if done else np.zeros(0)
    return Pieces(cat(0).reshape(-1, 64).astype(np.uint8), cat(1).astype(np.uint8), cat(2).astype(np.uint32),
                  cat(3).astype(np.uint64), cat(4).astype(np.float32), host_nodes, K)


def learned_split(model: "CostModel", n: int, level: int, cycles, mults, threshold: float, **kw) -> Pieces:
    """Split by the network's predicted log2(1 + nodes)."""
    return split_cases(n, level, cycles, mults, lambda l, c: model.predict(features_sym(n, l, c)), threshold, **kw)


def lookahead_split(n: int, level: int, cycles, mults, threshold: float, depth: int = 8, **kw) -> Pieces:
    """The same splitter driven by the reference's lookahead proxy instead of
    the network: log2(1 + descendants at level + depth)."""
    return split_cases(n, level, cycles, mults,
                       lambda l, c: np.log2(1.0 + estimate_work(n, l, c, depth).astype(np.float64)),
                       threshold, **kw)


# ----------------------------------------------------------------------------- training set


def sym_pairs(n: int, level: int, start: int, count: int, stride: int = 1):
    """(features, log2(1 + nodes), cycles, mults, nodes) for the cases
    [start, start + count*stride) of the stream, every `stride`-th one, with
    node counts measured by the GPU DFS (per-item profile of one solve)."""
    from . import SolveOptions, sym_cases, sym_solve
    cyc, mul = sym_cases(n, level, start, count * stride)
    cyc, mul = cyc[::stride], mul[::stride]
    _, _, nodes = sym_solve(n, level, cyc, mul, options=SolveOptions(count_events=False), per_item=True)
    return features_sym(n, level, cyc), node_target(nodes), cyc, mul, nodes


def lunnon_pairs(dims, depth: int, stride: int = 1):
    """(features, log2(1 + inserts), states, inserts) of a Lunnon frontier."""
    from . import SolveOptions, lunnon_frontier, lunnon_solve
    st = lunnon_frontier(dims, depth)[::stride]
    _, _, nodes = lunnon_solve(dims, depth, st, options=SolveOptions(count_events=False), per_item=True)
    return features_lunnon(dims, depth, st), node_target(nodes), st, nodes


# Default training windows: the 6x6 depth-18 Lunnon frontier (112,024 states),
# every 10th 7x7 depth-24 case (43,314), and 3 x 15,000 8x8 depth-36 cases
# from windows away from [120,000,000, +50,000) -- about 200k pairs.
DEFAULT_SYM_WINDOWS = ((7, 24, 0, 43314, 10), (8, 36, 0, 15000, 1), (8, 36, 60_000_000, 15000, 1),
                       (8, 36, 200_000_000, 15000, 1))
DEFAULT_LUNNON = (([6, 6], 18),)


def default_training_set(sym_windows=DEFAULT_SYM_WINDOWS, lunnon=DEFAULT_LUNNON):
    xs, ys = [], []
    for n, level, start, count, stride in sym_windows:
        x, y = sym_pairs(n, level, start, count, stride)[:2]
        xs.append(x)
        ys.append(y)
    for dims, depth in lunnon:
        x, y = lunnon_pairs(dims, depth)[:2]
        xs.append(x)
        ys.append(y)
    return np.concatenate(xs), np.concatenate(ys)


def train_default(seed: int = 0, epochs: int = 12, batch: int = 512, lr: float = 1e-3, data=None):
    """Collect the default pairs with the GPU DFS and train a model on the GPU.
    Returns (model, per-epoch losses, number of pairs)."""
    x, y = default_training_set() if data is None else data
    m = CostModel(seed=seed)
    losses = m.train(x, y, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return m, losses, len(y)


def main(argv=None) -> int:
    """python -m map_folding_amd.cost train --out PARAMS.npy [--seed S] [--epochs E]

    Collects the default training pairs with the GPU DFS, trains the model on
    the GPU and saves its FP32 parameters (np.save; load with CostModel.load
    or bench.py --cost-params)."""
    import argparse
    import json
    import time
    from . import init
    ap = argparse.ArgumentParser(prog="python -m map_folding_amd.cost")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--out", required=True)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--epochs", type=int, default=12)
    a = ap.parse_args(argv)
    init(0)
    t0 = time.time()
    data = default_training_set()
    t1 = time.time()
    m, losses, pairs = train_default(seed=a.seed, epochs=a.epochs, data=data)
    m.save(a.out)
    print(json.dumps({"out": a.out, "pairs": pairs, "collect_s": t1 - t0, "train_s": time.time() - t1,
                      "epoch_loss": [float(x) for x in losses]}), flush=True)
    return 0


if __name__ == "__main__":
    import sys
    sys.exit(main())
