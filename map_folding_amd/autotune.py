"""Autotuner for the scheduling knobs of the GPU DFS.

Every knob here only changes how work is cut and scheduled -- the Lunnon
partition depth per map, the re-split budget and tail, the resident waves per
CU, the sym deepening target, the learned split threshold -- never a count, so the tuner picks the
fastest configuration by measuring real solves on the GPU (best of `repeats`
kernel times) and checks every trial's result against the first one.

The reference fixes these by hand (`partition_depth` on the command line of
`gpu_folding.cu:209-225`, `MF_EXPANDED_N` of `sym_fold.cpp:655-660`); its
lanes-per-subtree is one, as here (one lane owns one subtree: the walks of a
node are serial pointer chases that a second lane could not shorten).

Results are plain dicts, so a tuned configuration can be saved as JSON and
replayed without searching (`save` / `load`).
"""
from __future__ import annotations

import json
import time

from . import LunnonBatchFrontier, SolveOptions, lunnon_frontier_size

__all__ = ["tune_lunnon_batch", "tune_resplit", "tune_sym", "depths_for_target", "save", "load"]


def _best(fn, repeats):
    best = None
    for _ in range(repeats):
        r = fn()
        best = r if best is None or r[0] < best[0] else best
    return best


def depths_for_target(shapes, target: int, max_states: int = 2_000_000):
    """Per map: the shallowest depth whose frontier has at least `target`
    states (at most `max_states`, and at most p*q - 2)."""
    out = []
    for p, q in shapes:
        n = p * q
        best = 2
        for d in range(2, n - 1):
            k = lunnon_frontier_size([p, q], d)
            if k > max_states:
                break
            best = d
            if k >= target:
                break
        out.append(best)
    return out


def tune_lunnon_batch(shapes, targets=(64, 256, 1024, 4096), budgets=(16, 64, 256), waves=(16, 32),
                      repeats=2, log=None, extra_depths=()):
    """Pick per-map frontier depths (by a target states-per-map, or one of the
    explicit per-map depth lists in `extra_depths`), then the re-split budget
    and resident waves, then whether the two dimensions of the (rectangular)
    maps walk in one loop (the default) or one after the other (core 4), for
    one batched launch of `shapes`.

    Returns {"depths", "budget", "waves_per_cu", "core", "kernel_ms", "trials"}."""
    say = log or (lambda m: None)
    trials = []
    ref_totals = None

    def solve_with(depths, budget, wv, core=0):
        nonlocal ref_totals
        t0 = time.time()
        b = LunnonBatchFrontier(shapes, depths)
        gen_s = time.time() - t0
        opt = SolveOptions(count_events=False, budget=budget, waves_per_cu=wv, core=core)

        def once():
            totals, r = b.solve_maps(opt)
            return r.kernel_ms, totals
        ms, totals = _best(once, repeats)
        b.close()
        if ref_totals is None:
            ref_totals = totals
        elif totals != ref_totals:
            raise RuntimeError("a tuning trial changed a map total")
        trials.append({"depths": list(depths), "budget": budget, "waves_per_cu": wv, "core": core, "kernel_ms": ms,
                       "gen_s": gen_s})
        say(f"  depths {min(depths)}..{max(depths)} budget {budget} waves {wv} core {core}: {ms:.1f} ms "
            f"(gen {gen_s:.2f}s)")
        return ms

    best = None
    candidates = [depths_for_target(shapes, t) for t in targets] + [list(d) for d in extra_depths]
    for d in candidates:
        if len(d) != len(shapes):
            raise ValueError("a depth list needs one depth per map")
        ms = solve_with(d, budgets[len(budgets) // 2], waves[-1])
        if best is None or ms < best[0]:
            best = (ms, d)
    depths = best[1]
    best_cfg = (best[0], budgets[len(budgets) // 2], waves[-1])
    for bu in budgets:
        for wv in waves:
            if (bu, wv) == best_cfg[1:]:
                continue
            ms = solve_with(depths, bu, wv)
            if ms < best_cfg[0]:
                best_cfg = (ms, bu, wv)
    core = 0
    ms = solve_with(depths, best_cfg[1], best_cfg[2], core=4)
    if ms < best_cfg[0]:
        best_cfg, core = (ms, best_cfg[1], best_cfg[2]), 4
    return {"depths": depths, "budget": best_cfg[1], "waves_per_cu": best_cfg[2], "core": core,
            "kernel_ms": best_cfg[0], "trials": trials}


def tune_resplit(frontier, budgets=(32, 64, 128), tail_mult=(1, 4), waves=(20, 32), repeats=1, log=None):
    """Budget, tail (as a multiple of the budget) and resident waves of the
    persistent re-split schedule on a resident frontier, then paired walks
    (the default) against unpaired ones (core 4: one page or one dimension at
    a time).

    Returns {"budget", "tail", "waves_per_cu", "core", "kernel_ms", "trials"}."""
    say = log or (lambda m: None)
    trials = []
    ref = None
    best = None
    for bu in budgets:
        for tm in tail_mult:
            for wv in waves:
                opt = SolveOptions(count_events=False, budget=bu, tail=bu * tm, waves_per_cu=wv)

                def once():
                    r, _, _ = frontier.solve(opt)
                    return r.kernel_ms, r.count
                ms, count = _best(once, repeats)
                if ref is None:
                    ref = count
                elif count != ref:
                    raise RuntimeError("a tuning trial changed the count")
                trials.append({"budget": bu, "tail": bu * tm, "waves_per_cu": wv, "kernel_ms": ms})
                say(f"  budget {bu} tail {bu * tm} waves {wv}: {ms:.1f} ms")
                if best is None or ms < best["kernel_ms"]:
                    best = {"budget": bu, "tail": bu * tm, "waves_per_cu": wv, "core": 0, "kernel_ms": ms}
    opt = SolveOptions(count_events=False, budget=best["budget"], tail=best["tail"], waves_per_cu=best["waves_per_cu"],
                       core=4)

    def unpaired():
        r, _, _ = frontier.solve(opt)
        return r.kernel_ms, r.count
    ms, count = _best(unpaired, repeats)
    if count != ref:
        raise RuntimeError("a tuning trial changed the count")
    trials.append({"budget": best["budget"], "tail": best["tail"], "waves_per_cu": best["waves_per_cu"], "core": 4,
                   "kernel_ms": ms})
    say(f"  unpaired walks (core 4): {ms:.1f} ms")
    if ms < best["kernel_ms"]:
        best.update(core=4, kernel_ms=ms)
    best["trials"] = trials
    return best


def tune_sym(frontier, pred=None, level=None, deepens=(45, 46, 47, 48), thetas=(19, 21, 23), caps=(2, 4),
             budgets=(64, 128), lanes=(1, 2), waves=32, repeats=1, log=None):
    """The symmetry DFS's partition depth and split threshold on a resident
    case frontier: the device deepening target (every case expanded to that
    level before the rounds: fewer stack levels per lane, more resident
    waves), then -- given the cost model's predictions `pred` (log2 nodes per
    case) -- the learned threshold theta and the cap on extra levels for the
    cases predicted large (cost.learned_levels), then the re-split budget,
    then the lanes per subtree (1, or 2: core 3, a lane pair splitting each
    node's page walks).  Each stage keeps the best of the previous one.
    `level` is the cases' level (deepening targets at or below it are
    skipped).

    Returns {"deepen", "theta", "cap", "budget", "lanes", "kernel_ms",
    "trials"}; theta/cap are None when the learned levels did not beat the
    uniform target."""
    from . import cost
    say = log or (lambda m: None)
    trials = []
    ref = None

    def trial(name, deepen, levels, budget, core=0):
        nonlocal ref
        opt = SolveOptions(count_events=False, budget=budget, tail=budget, waves_per_cu=waves, deepen=deepen,
                           deepen_levels=levels, core=core)

        def once():
            r, _, _ = frontier.solve(opt)
            return r.kernel_ms, (r.count, r.nodes)
        ms, res = _best(once, repeats)
        if ref is None:
            ref = res
        elif res != ref:
            raise RuntimeError("a tuning trial changed the count")
        trials.append({"trial": name, "deepen": deepen, "budget": budget, "core": core, "kernel_ms": ms})
        say(f"  {name}: {ms:.1f} ms")
        return ms

    budget = budgets[0]
    best = (trial("no deepening", -1, None, budget), -1)
    for d in deepens:
        if level is not None and d <= level:
            continue
        ms = trial(f"deepen {d}", d, None, budget)
        if ms < best[0]:
            best = (ms, d)
    deepen = best[1]
    best_ms = best[0]
    theta = cap = None
    if pred is not None and deepen > 0:
        for th in thetas:
            for cp in caps:
                ms = trial(f"deepen {deepen} + learned theta {th} cap {cp}", deepen,
                           cost.learned_levels(pred, deepen, th, cp), budget)
                if ms < best_ms:
                    best_ms, theta, cap = ms, th, cp
    levels = None if theta is None else cost.learned_levels(pred, deepen, theta, cap)
    for bu in budgets[1:]:
        ms = trial(f"budget {bu}", deepen, levels, bu)
        if ms < best_ms:
            best_ms, budget = ms, bu
    nlanes = 1
    for ln in lanes:
        if ln == 2:
            ms = trial("two lanes per subtree", deepen, levels, budget, core=3)
            if ms < best_ms:
                best_ms, nlanes = ms, 2
    return {"deepen": deepen, "theta": theta, "cap": cap, "budget": budget, "lanes": nlanes, "kernel_ms": best_ms,
            "trials": trials}


def save(cfg: dict, path: str) -> None:
    with open(path, "w") as f:
        json.dump(cfg, f, indent=1, sort_keys=True)


def load(path: str) -> dict:
    with open(path) as f:
        return json.load(f)
