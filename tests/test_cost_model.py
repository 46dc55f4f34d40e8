"""Learned subtree-cost model: features, the reference's lookahead proxy, the
GPU training/inference kernels against the CPU FP32 reference, and what the
trained network gains over scheduling without it."""
import numpy as np
import pytest

from map_folding_amd import cost


# ----------------------------------------------------------------------------- CPU


@pytest.fixture(scope="module")
def cases7(oracle):
    return oracle.sym_cases(7, 24, 1000, 120)


def test_estimate_work_matches_oracle(oracle, cases7):
    """estimate_work (sym_fold.cpp:235-250) = descendants at level+depth."""
    cyc, _ = cases7
    for depth in (0, 1, 3, 6):
        got = cost.estimate_work(7, 24, cyc[:40], depth)
        want = [1 if depth == 0 else oracle.sym_count_to(7, c, 24, 24 + depth)[0] for c in cyc[:40]]
        assert got.tolist() == want
    # depth past the last level stops at n^2: the leaf count of the case.
    got = cost.estimate_work(5, 14, oracle.sym_prefixes(5, 14)[:30], 40)
    assert got.tolist() == [oracle.sym_count_to(5, c, 14, 25)[0] for c in oracle.sym_prefixes(5, 14)[:30]]


def test_sym_features_pinned_to_oracle(oracle, cases7):
    cyc, _ = cases7
    f = cost.features_sym(7, 24, cyc)
    assert f.shape == (len(cyc), 80) and np.isfinite(f).all()
    for k, c in enumerate(cyc[:30]):
        children = oracle.sym_count_to(7, c, 24, 25)[0]
        assert round(f[k, 66] * 16) == children
        # The next cell's option count is exactly get_good's popcount.
        assert round(f[k, 24] * 24) == children
        # Lookahead profile: log2(1 + descendants at level 24 + j).
        for j in range(1, 5):
            want = np.log2(1 + oracle.sym_count_to(7, c, 24, 24 + j)[0]) / 24
            assert abs(f[k, 71 + j] - want) < 1e-6
        # Placed cells carry their rank along the stacking cycle: a permutation.
        assert sorted(np.rint(f[k, :24] * 64).astype(int).tolist()) == list(range(1, 25))
        assert (f[k, 24:49] >= 0).all() and (f[k, 24:49] <= 1).all() and (f[k, 49:64] == 0).all()


def test_lunnon_feature_profile_consistent(oracle, mf):
    """Gaps of the filter-passing children of all depth-D states equal the gap
    counts of the depth-(D+1) frontier (two depths, one recursion)."""
    dims, D = [4, 4], 8
    s0, _ = oracle.lunnon_frontier(dims, D)
    s1, _ = oracle.lunnon_frontier(dims, D + 1)
    f0 = cost.features_lunnon(dims, D, s0)
    f1 = cost.features_lunnon(dims, D + 1, s1)
    lvl1 = np.rint(2 ** (f0[:, 73] * 24) - 1).sum()
    gaps_next = np.rint(f1[:, 66] * 16).sum()
    assert lvl1 == gaps_next
    assert np.rint(f0[:, 66] * 16).sum() >= len(s1)
    for k in range(20):
        assert sorted(np.rint(f0[k, :D + 1] * 64).astype(int).tolist()) == list(range(1, D + 2))


def test_sort_by_score_semantics():
    """score_and_sort_cases: score descending (or ascending), ties by
    multiplier ascending (sym_fold.cpp:295-305), then stream index."""
    scores = [5, 9, 5, 1, 9, 5]
    mults = [8, 16, 4, 8, 8, 4]
    assert cost.sort_by_score(scores, mults).tolist() == [4, 1, 2, 5, 0, 3]
    assert cost.sort_by_score(scores, mults, ascending=True).tolist() == [3, 2, 5, 0, 4, 1]


def test_class_order_semantics():
    """Pieces: size classes of width 2 (log2 nodes) largest first, then level,
    then the splitter's order; a permutation whatever the scores."""
    pred = [3.9, 9.0, 2.0, 8.1, 9.5, 3.0, 8.9]
    levels = [37, 36, 36, 38, 37, 36, 36]
    # classes: floor(p/2) = 1, 4, 1, 4, 4, 1, 4
    assert cost.class_order(pred, levels).tolist() == [1, 6, 4, 3, 2, 5, 0]
    assert cost.class_order(pred, levels, width=100.0).tolist() == [1, 2, 5, 6, 0, 4, 3]
    rng = np.random.default_rng(3)
    p, lv = rng.normal(8, 4, 5000), rng.integers(36, 44, 5000)
    o = cost.class_order(p, lv)
    assert sorted(o.tolist()) == list(range(5000))
    k = np.floor(p[o] / 2)
    assert (np.diff(k) <= 0).all()
    same = np.diff(k) == 0
    assert (np.diff(lv[o])[same] >= 0).all()
    tie = same & (np.diff(lv[o]) == 0)
    assert (np.diff(o.astype(np.int64))[tie] > 0).all()


def test_schedule_metrics():
    assert cost.wave_waste([4] * 128) == pytest.approx(1.0)
    assert cost.wave_waste([8] + [0] * 63 + [8] + [0] * 63) == pytest.approx(64.0)
    assert cost.list_makespan([5, 3, 3, 1], 2) == pytest.approx(1.0)
    assert cost.list_makespan([1, 3, 3, 5], 2) == pytest.approx(8 / 6)
    assert cost.spearman([1, 2, 3], [10, 20, 30]) == pytest.approx(1.0)


# ----------------------------------------------------------------------------- GPU


@pytest.fixture(scope="module")
def ref():
    from oracle import cost_ref
    return cost_ref


def _data(rows, seed):
    rng = np.random.default_rng(seed)
    x = rng.random((rows, 80), dtype=np.float32)
    y = (rng.standard_normal(rows) * 3 + 10).astype(np.float32)
    return x, y


@pytest.mark.gpu
@pytest.mark.parametrize("rows", [1, 77, 300, 1024])
def test_gradient_matches_cpu_reference(gpu, ref, rows):
    m = cost.CostModel(seed=rows)
    x, y = _data(rows, rows)
    loss, g = m.gradient(x, y)
    want_loss, want = ref.loss_and_grad(m.params, x, y)
    assert loss == pytest.approx(want_loss, rel=1e-5)
    # FP32 sums in different orders: tolerance 1e-4 of the largest entry per block.
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in ref.SHAPES])
    for (name, _), lo, hi in zip(ref.SHAPES, offs[:-1], offs[1:]):
        scale = max(np.abs(want[lo:hi]).max(), 1e-6)
        np.testing.assert_allclose(g[lo:hi], want[lo:hi], rtol=0, atol=1e-4 * scale, err_msg=name)
    m.close()


@pytest.mark.gpu
def test_forward_fp32_and_bf16_inference(gpu, ref):
    m = cost.CostModel(seed=3)
    x, _ = _data(1000, 7)
    p = m.params
    f32 = m.predict_fp32(x)
    np.testing.assert_allclose(f32, ref.predict_fp32(p, x), rtol=1e-5, atol=1e-4)
    b16 = m.predict(x)
    want = ref.predict_bf16(p, x)
    # Same BF16 roundings; FP32 accumulation order differs, which can move an
    # activation across a BF16 rounding boundary: allow 1% of the output scale.
    scale = np.abs(want).max()
    np.testing.assert_allclose(b16, want, rtol=0, atol=1e-2 * scale)
    assert np.abs(b16 - f32).max() < 0.05 * scale
    m.close()


@pytest.mark.gpu
def test_training_reduces_loss_and_is_deterministic(gpu):
    rng = np.random.default_rng(1)
    x = rng.random((4096, 80), dtype=np.float32)
    v = rng.standard_normal(80).astype(np.float32)
    y = (x @ v + 4.0).astype(np.float32)
    losses = []
    params = []
    for _ in range(2):
        m = cost.CostModel(seed=11)
        losses.append(m.train(x, y, epochs=15, batch=256, lr=2e-3, seed=5))
        params.append(m.params)
        m.close()
    l = losses[0]
    assert l[-1] < 0.05 * l[0], l
    np.testing.assert_array_equal(losses[0], losses[1])
    np.testing.assert_array_equal(params[0], params[1])


@pytest.mark.gpu
def test_learned_model_gains_over_no_model(gpu):
    """Pairs (case, node count) logged by the GPU DFS on two 8x8 depth-36
    windows train the network on the GPU; on a third, disjoint window its
    predictions rank the measured subtree sizes far better than the
    reference's depth-8 lookahead proxy, and ordering by them cuts the
    one-lane-per-case kernel's idle lane-slots against stream order."""
    import map_folding_amd as mf
    def pairs(start, k):
        cyc, mul = mf.sym_cases(8, 36, start, k)
        _, _, nodes = mf.sym_solve(8, 36, cyc, mul, options=mf.SolveOptions(count_events=False), per_item=True)
        return cyc, mul, cost.features_sym(8, 36, cyc), nodes
    tr = [pairs(0, 5000), pairs(60_000_000, 5000)]
    x = np.concatenate([t[2] for t in tr])
    y = cost.node_target(np.concatenate([t[3] for t in tr]))
    m = cost.CostModel(seed=0)
    m.train(x, y, epochs=20, batch=256, lr=1e-3, seed=0)
    cyc, mul, xe, nodes = pairs(120_000_000, 8000)
    pred = m.predict(xe)
    rho_model = cost.spearman(pred, nodes)
    proxy = cost.estimate_work(8, 36, cyc, 8)
    rho_proxy = cost.spearman(proxy, nodes)
    assert rho_model > 0.7 and rho_model > rho_proxy + 0.15, (rho_model, rho_proxy)
    stream = cost.wave_waste(nodes)
    learned = cost.wave_waste(nodes[cost.learned_order(pred)])
    lookahead = cost.wave_waste(nodes[cost.lookahead_order(proxy, mul)])
    assert learned < 0.6 * stream and learned < lookahead, (stream, learned, lookahead)
    m.close()


@pytest.mark.gpu
@pytest.mark.parametrize("n,e,start,num", [(8, 36, 120000000, 3000), (7, 24, 200000, 2000), (6, 16, 0, 1225)])
def test_device_features_equal_host_features(gpu, n, e, start, num):
    """The feature rows computed on the device from a resident frontier equal
    the host's (cost.cpp sym_row) -- exactly for the ranks, options and
    scalars, to float rounding for the log2 lookahead profile -- and the
    on-device prediction equals the prediction of the host rows."""
    import os
    from map_folding_amd import cost
    if n == 6:
        cyc, mul, _ = gpu.sym_seeds(6)
    else:
        cyc, mul = gpu.sym_cases(n, e, start, num)
    f = gpu.SymFrontier(n, e, cyc, mul)
    xd = cost.features_frontier(f)
    xh = cost.features_sym(n, e, cyc)
    assert xd.shape == xh.shape
    assert np.array_equal(xd[:, :72], xh[:, :72])
    assert np.allclose(xd[:, 72:], xh[:, 72:], rtol=0, atol=1e-6)
    m = cost.CostModel(seed=0)
    m.load(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "map_folding_amd",
                        "cost_default.npy"))
    pd = m.predict_frontier(f)
    ph = m.predict(xh)
    assert np.allclose(pd, ph, rtol=0, atol=1e-3)
    f.close()
    m.close()
