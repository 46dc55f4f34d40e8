"""Autotuner: trials never change a count, and the pick is the fastest trial."""
import pytest


def test_depths_for_target(mf):
    from map_folding_amd import autotune
    shapes = [(3, 4), (2, 6), (4, 4)]
    d = autotune.depths_for_target(shapes, 50)
    for (p, q), depth in zip(shapes, d):
        assert 2 <= depth <= p * q - 2
        k = mf.lunnon_frontier_size([p, q], depth)
        assert k >= 50 or depth == p * q - 2 or mf.lunnon_frontier_size([p, q], depth + 1) > 2_000_000
        if depth > 2:
            assert mf.lunnon_frontier_size([p, q], depth - 1) < 50


@pytest.mark.gpu
def test_tune_lunnon_batch_keeps_totals(gpu, golden):
    from map_folding_amd import autotune
    shapes = [(3, 4), (2, 6), (4, 4), (3, 5), (2, 8)]
    extra = [[6, 5, 8, 7, 9]]
    cfg = autotune.tune_lunnon_batch(shapes, targets=(16, 128), budgets=(16, 64), waves=(16, 32), repeats=1,
                                     extra_depths=extra)
    assert cfg["kernel_ms"] == min(t["kernel_ms"] for t in cfg["trials"])
    # Every depth candidate (two targets and the explicit list) is timed once
    # before the budget / waves sweep on the best of them.
    assert [t["depths"] for t in cfg["trials"][:3]][2] == extra[0]
    with pytest.raises(ValueError):
        autotune.tune_lunnon_batch(shapes, targets=(), budgets=(16,), waves=(32,), repeats=1, extra_depths=[[6]])
    # The last trial walks the dimensions one after the other (core 4).
    assert cfg["trials"][-1]["core"] == 4 and cfg["core"] in (0, 4)
    b = gpu.LunnonBatchFrontier(shapes, cfg["depths"])
    for core in (0, 4):
        totals, _ = b.solve_maps(gpu.SolveOptions(budget=cfg["budget"], waves_per_cu=cfg["waves_per_cu"], core=core))
        assert totals == [golden["rectangles"][f"{p}x{q}"] for p, q in shapes]
    b.close()


@pytest.mark.gpu
def test_tune_resplit_keeps_count(gpu, golden):
    from map_folding_amd import autotune
    cyc, mul = gpu.sym_cases(7, 24, 200000, 2000)
    f = gpu.SymFrontier(7, 24, cyc, mul)
    ref, _, _ = f.solve()
    cfg = autotune.tune_resplit(f, budgets=(16, 64), tail_mult=(1, 4), waves=(20, 32))
    assert len(cfg["trials"]) == 9 and cfg["kernel_ms"] == min(t["kernel_ms"] for t in cfg["trials"])
    assert cfg["trials"][-1]["core"] == 4
    r, _, _ = f.solve(gpu.SolveOptions(budget=cfg["budget"], tail=cfg["tail"], waves_per_cu=cfg["waves_per_cu"],
                                       core=cfg["core"]))
    assert r.count == ref.count
    f.close()


class _FakeSymFrontier:
    """Stands in for a resident frontier: a fixed count, and a kernel time
    that depends on the options as `ms_of` says."""

    def __init__(self, ms_of):
        self.ms_of = ms_of
        self.seen = []

    def solve(self, opt):
        import types
        self.seen.append(opt)
        return types.SimpleNamespace(kernel_ms=self.ms_of(opt), count=7, nodes=11), None, None


def test_tune_sym_stages_and_lanes():
    """Each stage keeps the best of the previous one; the lanes stage times
    core 3 at the chosen schedule and keeps it only when it is faster."""
    import numpy as np
    from map_folding_amd import autotune

    def ms(opt):
        t = 100.0 - (5.0 if opt.deepen == 46 else 0.0) - (2.0 if opt.deepen_levels is not None else 0.0)
        return t + (30.0 if opt.core == 3 else 0.0) + (1.0 if opt.budget == 128 else 0.0)
    f = _FakeSymFrontier(ms)
    pred = np.array([10.0, 25.0, 30.0])
    cfg = autotune.tune_sym(f, pred=pred, level=36, deepens=(45, 46), thetas=(21,), caps=(4,), budgets=(64, 128))
    assert (cfg["deepen"], cfg["theta"], cfg["cap"], cfg["budget"], cfg["lanes"]) == (46, 21, 4, 64, 1)
    assert cfg["trials"][-1]["trial"] == "two lanes per subtree" and cfg["trials"][-1]["core"] == 3
    assert f.seen[-1].core == 3 and f.seen[-1].deepen == 46 and f.seen[-1].deepen_levels is not None
    # A faster pair kernel is kept.
    f2 = _FakeSymFrontier(lambda o: ms(o) - (60.0 if o.core == 3 else 0.0))
    assert autotune.tune_sym(f2, pred=pred, level=36, deepens=(46,), thetas=(21,), caps=(4,))["lanes"] == 2
    # Without the pair stage, no core-3 trial runs.
    f3 = _FakeSymFrontier(ms)
    assert autotune.tune_sym(f3, level=36, deepens=(46,), lanes=(1,))["lanes"] == 1
    assert all(o.core == 0 for o in f3.seen)


def test_tune_resplit_unpaired_stage():
    """After the budget / tail / waves sweep, one trial runs unpaired walks
    (core 4) at the chosen schedule and is kept only when it is faster."""
    from map_folding_amd import autotune

    def ms(opt):
        return 50.0 + (3.0 if opt.budget == 16 else 0.0) + (4.0 if opt.core == 4 else 0.0)
    f = _FakeSymFrontier(ms)
    cfg = autotune.tune_resplit(f, budgets=(16, 64), tail_mult=(1,), waves=(32,))
    assert (cfg["budget"], cfg["core"], cfg["kernel_ms"]) == (64, 0, 50.0)
    assert f.seen[-1].core == 4 and f.seen[-1].budget == 64
    f2 = _FakeSymFrontier(lambda o: ms(o) - (8.0 if o.core == 4 else 0.0))
    cfg2 = autotune.tune_resplit(f2, budgets=(16, 64), tail_mult=(1,), waves=(32,))
    assert (cfg2["core"], cfg2["kernel_ms"]) == (4, 46.0)
