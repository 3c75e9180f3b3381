"""Batched self-joins (one pipeline, one persistent screen per precision
family) against the CPU FP64 oracle, join by join, bit-exact; hipGraph
replays against the eager run."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def tj():
    import torch
    assert torch.cuda.is_available(), "no GPU visible"
    import tensorjoin_variants_amd as T
    T.load_library()
    return T


def mixture(rng, n, d, k=4):
    c = rng.standard_normal((k, d)) * 3.0
    return (c[rng.integers(0, k, n)] + rng.standard_normal((n, d))).astype(np.float32)


def eps_for(rng, X, q=0.01):
    from tensorjoin_variants_amd.thresholds import percentile_eps
    n = X.shape[0]
    if n < 2:
        return 1.0
    return percentile_eps(X, None, q, min(4000, n * (n - 1) // 2), rng)


def check_all(oracle, sets, Ts, got, name):
    assert len(got) == len(sets)
    for k, (X, T, g) in enumerate(zip(sets, Ts, got)):
        want = oracle.join(X, None, float(T))
        assert g.shape == want.shape, (name, k, X.shape, g.shape, want.shape)
        assert np.array_equal(g.astype(np.int64), want), (name, k, X.shape)


def mixed_sets(seed, sizes, dims):
    rng = np.random.default_rng(seed)
    sets, Ts = [], []
    for n in sizes:
        d = int(rng.choice(dims))
        X = mixture(rng, n, d)
        sets.append(X)
        e = eps_for(rng, X)
        Ts.append(e * e)
    return sets, np.asarray(Ts)


SIZES = [0, 1, 2, 3, 31, 64, 65, 127, 200, 257, 511, 700, 1500, 2100]


@pytest.mark.parametrize("prec,prune", [("fp16", False), ("fp8", False), ("fp16", True), ("fp8", True)])
def test_batch_mixed_widths(tj, oracle, prec, prune):
    sets, Ts = mixed_sets(11, SIZES * 2, [1, 3, 16, 17, 31, 64, 100, 128, 200, 256])
    say("batch mixed widths", prec, "prune" if prune else "", len(sets))
    got, st = tj.batched_join(sets, T=Ts, precision=prec, audit=True, return_stats=True, prune=prune)
    say("batch stats", {k: st[k] for k in ("out", "candidates", "ambiguous", "violations")})
    assert st["violations"] == 0
    check_all(oracle, sets, Ts, got, f"mixed {prec}")


def test_batch_mixed_precisions(tj, oracle):
    sets, Ts = mixed_sets(12, SIZES, [16, 64, 256])
    precs = ["fp16" if k % 3 else "fp8" for k in range(len(sets))]
    say("batch mixed precisions", precs)
    got, st = tj.batched_join(sets, T=Ts, precision=precs, audit=True, return_stats=True)
    assert st["violations"] == 0
    assert st["plan"]["items_fp8"] > 0 and st["plan"]["items_fp16"] > 0 and st["plan"]["launches"] >= 2
    check_all(oracle, sets, Ts, got, "mixed precisions")


def test_batch_matches_single_join(tj):
    sets, Ts = mixed_sets(13, [300, 1000, 64, 2000], [16, 64, 256])
    say("batch vs single join")
    got = tj.batched_join(sets, T=Ts, precision="fp16")
    for X, T, g in zip(sets, Ts, got):
        single = tj.join(X, None, T=float(T), precision="fp16", out_dtype=np.int64)
        assert np.array_equal(g.astype(np.int64), single)


@pytest.mark.parametrize("prune", [False, True])
def test_batch_edge_cases(tj, oracle, prune):
    rng = np.random.default_rng(14)
    # duplicates with T = 0, all-equal rows, a set of one row, empty sets
    A = mixture(rng, 300, 64)
    A[150:] = A[:150]
    B = np.ones((40, 16), np.float32)
    C = mixture(rng, 1, 256)
    D = np.zeros((0, 64), np.float32)
    E = mixture(rng, 900, 256) * np.float32(1e-3)  # small magnitudes (join-unit scaling)
    F = mixture(rng, 500, 3) * np.float32(1e4)     # large magnitudes, tiny width
    sets = [A, B, C, D, E, F]
    Ts = np.array([0.0, 0.0, 1.0, 1.0, float(eps_for(rng, E)) ** 2, float(eps_for(rng, F)) ** 2])
    say("batch edge cases", "prune" if prune else "")
    got, st = tj.batched_join(sets, T=Ts, precision="fp16", audit=True, return_stats=True, prune=prune)
    assert st["violations"] == 0
    check_all(oracle, sets, Ts, got, "edge")
    assert got[1].shape[0] == 40 * 39 // 2 and got[3].shape == (0, 2) and got[2].shape == (0, 2)


def test_batch_all_tiny(tj, oracle):
    rng = np.random.default_rng(15)
    sets = [mixture(rng, n, 16) for n in (0, 1, 0, 1)]
    say("batch of sets with < 2 rows")
    got = tj.batched_join(sets, T=[1.0] * 4, precision="fp16")
    assert all(g.shape == (0, 2) for g in got)


def test_batch_capacity_retry(tj, oracle):
    from tensorjoin_variants_amd import BatchPlan
    sets, Ts = mixed_sets(16, [1200, 900, 1500], [16, 64])
    say("batch capacity retry")
    plan = BatchPlan([x.shape[0] for x in sets], [x.shape[1] for x in sets], "fp16", 16, 16)
    got, st = plan.run(sets, Ts)
    assert plan.out_capacity >= st["out"] and plan.cand_capacity >= st["candidates"]
    check_all(oracle, sets, Ts, got, "retry")


def test_batch_uneven_workload_reduced(tj, oracle):
    """The batched workload's distribution (log-uniform sizes, d in {16, 64, 256},
    4-cluster mixtures, eps at the 0.5th percentile) at 192 joins of <= 2048 rows."""
    from tensorjoin_variants_amd import workloads
    ws = workloads.batched_uneven(joins=192, nmax=2048, samples=4000)
    sets = [w.R for w in ws]
    Ts = np.array([w.T for w in ws])
    say("batched_uneven reduced", sum(x.shape[0] for x in sets))
    got, st = tj.batched_join(sets, T=Ts, precision="fp16", audit=True, return_stats=True, prune=True)
    say("stats", {k: st[k] for k in ("out", "candidates", "ambiguous", "violations", "max_err")})
    assert st["violations"] == 0
    check_all(oracle, sets, Ts, got, "batched_uneven")


def test_batch_graph_replay(tj):
    import torch
    from tensorjoin_variants_amd import BatchPlan
    from tensorjoin_variants_amd.batch import initial_capacities
    sets, Ts = mixed_sets(17, [800, 1600, 300, 64, 2048], [16, 64, 256])
    ns, ds = [x.shape[0] for x in sets], [x.shape[1] for x in sets]
    oc, cc = initial_capacities(ns)
    plan = BatchPlan(ns, ds, "fp16", oc, cc)
    eager, est = plan.run(sets, Ts)
    xs, ptrs, Td = plan.inputs(sets, Ts)
    say("batch graph capture")
    plan.capture(ptrs, Td)
    for rep in range(3):
        plan.pairs.fill_(-1)
        plan._stats.fill_(0)
        junk = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        say("batch graph replay", rep)
        plan.replay()
        got, st = plan.results()
        assert st["out"] == est["out"] and st["candidates"] == est["candidates"]
        for a, b in zip(got, eager):
            assert np.array_equal(a, b)
        del junk


def test_plan_graph_replay(tj):
    import torch
    rng = np.random.default_rng(18)
    X = mixture(rng, 3000, 128)
    e = eps_for(rng, X)
    from tensorjoin_variants_amd.join import Plan, _as_device, _initial_capacities
    oc, ac = _initial_capacities(3000, None, True)
    plan = Plan(3000, None, 128, "certified", "fp16", oc, ac)
    Xd = _as_device(X)
    res = plan.run(Xd, None, e * e, np.int64)
    eager = res.numpy()
    pairs = torch.empty((plan.out_capacity, 2), dtype=torch.int64, device="cuda")
    from tensorjoin_variants_amd import _lib
    from tensorjoin_variants_amd.join import _ptr
    L = _lib.load()
    say("plan graph capture")
    _lib.check(L.tj_plan_capture(plan._plan, _ptr(Xd), None, e * e, _ptr(plan._ws), _ptr(pairs), 1,
                                 _ptr(plan._stats)), "tj_plan_capture")
    for rep in range(3):
        pairs.fill_(-1)
        torch.cuda.synchronize()
        say("plan graph replay", rep)
        _lib.check(L.tj_plan_replay(plan._plan, torch.cuda.current_stream().cuda_stream), "tj_plan_replay")
        torch.cuda.synchronize()
        n = int(plan._stats[0].item())
        assert n == eager.shape[0]
        assert np.array_equal(pairs[:n].cpu().numpy(), eager)


def test_batch_uneven_full_sizes_pruned(tj, oracle):
    """Production mode (fast accepts on) at the batched workload's full size
    range (n up to 16384), tile pruning on: items sweep up to four column
    ranges back to back, which share the workgroup's LDS ring (a refill
    before every wave has left the previous range lost pairs)."""
    from tensorjoin_variants_amd import workloads
    ws = workloads.batched_uneven(joins=96, samples=4000, seed_offset=11)
    sets = [w.R for w in ws]
    Ts = np.array([w.T for w in ws])
    say("batched_uneven full sizes", len(sets), sum(x.shape[0] for x in sets))
    for rep in range(2):
        got, st = tj.batched_join(sets, T=Ts, precision="fp16", return_stats=True, prune=True)
        say("stats", rep, {k: st[k] for k in ("out", "candidates", "pruned", "violations")})
        assert st["pruned"] > 0
        check_all(oracle, sets, Ts, got, "batched_uneven full sizes")


def separated(rng, n, d, k, gap):
    """k Gaussian clusters (unit spread) whose centres are ~gap apart."""
    c = rng.standard_normal((k, d)) * gap / np.sqrt(2.0 * d)
    return (c[rng.integers(0, k, n)] + rng.standard_normal((n, d))).astype(np.float32)


@pytest.mark.parametrize("prec", ["fp16", "fp8"])
def test_tile_pruning_matches_unpruned_and_oracle(tj, oracle, prec):
    """Far-apart clusters: most panel pairs are certified apart and skipped;
    results equal the unpruned batch and the oracle, join by join."""
    rng = np.random.default_rng(21)
    sets, Ts = [], []
    for n, d in [(3000, 16), (2600, 64), (2200, 256), (1800, 32), (700, 128), (300, 16)]:
        if prec == "fp8" and d < 64:
            d = 64
        X = separated(rng, n, d, 4, 12.0 * np.sqrt(d))
        sets.append(X)
        Ts.append(float(eps_for(rng, X, 0.02)) ** 2)
    Ts = np.asarray(Ts)
    say("tile pruning", prec, [x.shape for x in sets])
    got, st = tj.batched_join(sets, T=Ts, precision=prec, return_stats=True, prune=True)
    ref, st0 = tj.batched_join(sets, T=Ts, precision=prec, return_stats=True)
    say("pruned pairs", st["pruned"], "of", st["pairs"], "candidates", st["candidates"], st0["candidates"])
    assert st["pruned"] > st["pairs"] // 4 and st0["pruned"] == 0
    for g, r in zip(got, ref):
        assert np.array_equal(g, r)
    check_all(oracle, sets, Ts, got, f"pruned {prec}")


def test_tile_pruning_keeps_cross_cluster_pairs(tj, oracle):
    """T large enough that pairs across neighbouring clusters are inside:
    the ball bound must keep every such panel pair."""
    rng = np.random.default_rng(22)
    sets = [separated(rng, 2500, 16, 6, 12.0), separated(rng, 1500, 64, 3, 20.0)]
    Ts = []
    for X in sets:
        M = oracle.d2_matrix(X[:400], X[:400])
        Ts.append(float(np.quantile(M[M > 0], 0.4)))  # reaches into other clusters
    Ts = np.asarray(Ts)
    say("tile pruning, wide T")
    got, st = tj.batched_join(sets, T=Ts, precision="fp16", audit=True, return_stats=True, prune=True)
    assert st["violations"] == 0
    check_all(oracle, sets, Ts, got, "wide T")


def test_tile_pruning_degenerate_rows(tj, oracle):
    """All-equal rows, duplicated rows at T = 0, and huge / tiny magnitudes:
    the balls degenerate (radius 0) or overflow, pruning stays sound."""
    rng = np.random.default_rng(23)
    A = np.ones((1200, 16), np.float32)
    B = separated(rng, 1000, 64, 4, 96.0)
    B[500:] = B[:500]
    C = separated(rng, 900, 32, 4, 68.0) * np.float32(1e30)
    D = separated(rng, 900, 32, 4, 68.0) * np.float32(1e-30)
    # +-3e38 coordinates: panel centres overflow (inf or NaN); no pruning there
    E = (rng.choice([-3.0e38, 3.0e38], size=(700, 16)) * rng.choice([1.0, 0.5], size=(700, 16))).astype(np.float32)
    E[350:] = E[:350]
    sets = [A, B, C, D, E]
    Ts = np.array([0.0, 0.0, float(eps_for(rng, C, 0.02)) ** 2, float(eps_for(rng, D, 0.02)) ** 2,
                   float(np.float64(3.0e38) ** 2)])
    say("tile pruning, degenerate rows")
    got, st = tj.batched_join(sets, T=Ts, precision="fp16", audit=True, return_stats=True, prune=True)
    assert st["violations"] == 0
    check_all(oracle, sets, Ts, got, "degenerate")
    assert got[0].shape[0] == 1200 * 1199 // 2
