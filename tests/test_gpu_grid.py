"""Uniform-grid join (d <= 3) against the CPU FP64 oracle, bit-exact pair sets."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def tj():
    import torch
    assert torch.cuda.is_available()
    import tensorjoin_variants_amd as T
    T.load_library()
    return T


def check(tj, oracle, R, S, T, name):
    say(name, R.shape, None if S is None else S.shape, T)
    got, st = tj.join(R, S, T=T, method="grid", out_dtype=np.int64, return_stats=True)
    want = oracle.join(R, S, T) if R.shape[0] * (R.shape[0] if S is None else S.shape[0]) < 4e8 else oracle.grid_join(R, S, T)
    assert got.shape == want.shape, (name, got.shape, want.shape)
    assert np.array_equal(got, want), name
    return st


@pytest.mark.parametrize("d", [1, 2, 3])
def test_grid_random(tj, oracle, d):
    rng = np.random.default_rng(30 + d)
    R = rng.random((5000, d)).astype(np.float32)
    S = rng.random((4000, d)).astype(np.float32)
    eps = {1: 0.0005, 2: 0.01, 3: 0.03}[d]
    st = check(tj, oracle, R, S, eps * eps, f"rxs d{d}")
    assert st["grid_candidates"] < R.shape[0] * S.shape[0] / 10
    check(tj, oracle, R, None, eps * eps, f"self d{d}")


def test_grid_boundaries(tj, oracle):
    # points on an exact lattice of step eps: many pairs at exactly d2 = T
    g = np.arange(30, dtype=np.float32) * np.float32(0.125)
    X = np.stack(np.meshgrid(g, g, g[:5], indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    check(tj, oracle, X, None, 0.125 ** 2, "lattice self T=eps^2")
    check(tj, oracle, X, X[::3].copy(), 0.125 ** 2, "lattice rxs")
    # duplicates and T = 0
    rng = np.random.default_rng(8)
    Y = rng.random((500, 3)).astype(np.float32)
    Y[250:] = Y[:250]
    check(tj, oracle, Y, None, 0.0, "duplicates T=0")
    # negative / mixed-sign coordinates, large offsets
    Z = (rng.standard_normal((3000, 2)) * 100 + 1e4).astype(np.float32)
    check(tj, oracle, Z, None, 4.0, "offset 1e4")
    # one dimension degenerate (all equal)
    W = rng.random((2000, 3)).astype(np.float32)
    W[:, 1] = 0.5
    check(tj, oracle, W, None, 0.02 ** 2, "flat dim")
    # a grid far sparser than 64 cells per point (no occupancy bitmap):
    # 8e9 cells for 3,000 points, pairs only between near-copies
    V = (rng.random((1500, 3)) * 1000).astype(np.float32)
    V = np.concatenate([V, V + rng.uniform(-0.2, 0.2, V.shape).astype(np.float32)]).astype(np.float32)
    check(tj, oracle, V, None, 0.25, "sparse grid self")
    check(tj, oracle, V[:1500], V[1500:], 0.25, "sparse grid rxs")


def test_grid_xwindow_lattice(tj, oracle):
    """Dense lattice blocks (512 points per cell): every neighbour row is a
    long range, so each query trims it to its x window by binary search on
    the S rows' sub-cell x position and skips rows beyond its y/z reach.
    Lattice coordinates (multiples of 2^-9, centred on 0) make d2 exact, with
    many pairs at exactly d2 = T on the window edges and cell faces.  The
    second set adds two far points, which leave fewer key bits for the x
    position (a coarser window)."""
    g = (np.arange(24, dtype=np.float64) - 12.0) * 2.0 ** -9
    X = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    T = 64.0 * 2.0 ** -18  # eps = 8 lattice steps
    check(tj, oracle, X, None, T, "lattice blocks self")
    check(tj, oracle, X[::7].copy(), X, T, "lattice blocks rxs")
    far = np.array([[500.0, 500.0, 500.0], [-500.0, 499.0, -498.0]], dtype=np.float32)
    Xf = np.concatenate([X, far]).astype(np.float32)
    check(tj, oracle, Xf, None, T, "lattice blocks + far points self")
    check(tj, oracle, Xf[::5].copy(), Xf, T, "lattice blocks + far points rxs")


def test_grid_edge_sizes(tj, oracle):
    rng = np.random.default_rng(9)
    X = rng.random((70, 3)).astype(np.float32)
    for nR, nS in [(1, 1), (1, 70), (70, 1)]:
        check(tj, oracle, X[:nR], X[:nS], 0.3 ** 2, f"{nR}x{nS}")
    check(tj, oracle, X[:1], None, 0.3 ** 2, "self 1")
    say("grid empty")
    assert tj.join(X[:0], X, T=1.0, method="grid").shape == (0, 2)


@pytest.mark.parametrize("d", [2, 3])
def test_grid_dense_blobs_shared_sweep(tj, oracle, d):
    """Tight blobs in a sparse field: waves whose lanes hold hundreds of
    candidate rows next to lanes with none (the wave-shared sweep), for R x S
    and the self-join's j > i filter."""
    rng = np.random.default_rng(40 + d)
    n = 20000
    X = rng.random((n, d)).astype(np.float32)
    centres = rng.random((6, d))
    k = n // 3
    X[:k] = (centres[rng.integers(0, 6, k)] + rng.normal(0, 0.002, (k, d))).astype(np.float32)
    Y = rng.random((n, d)).astype(np.float32)
    Y[: n // 4] = (centres[rng.integers(0, 6, n // 4)] + rng.normal(0, 0.002, (n // 4, d))).astype(np.float32)
    eps = 0.004
    st = check(tj, oracle, X, None, eps * eps, f"dense blobs self d{d}")
    say("dense blobs", d, st["out"], st["grid_candidates"])
    check(tj, oracle, Y, X, eps * eps, f"dense blobs rxs d{d}")


def test_grid_non_finite_rejected(tj):
    """The grid's bounds pass counts non-finite coordinates (in its 4-point
    vector body and in the tail) and the join is rejected."""
    rng = np.random.default_rng(50)
    X = rng.random((1001, 3)).astype(np.float32)
    Y = rng.random((700, 3)).astype(np.float32)
    for name, A, B, r, c, v in [("self nan body", X, None, 17, 1, np.nan),
                                ("self inf tail", X, None, 1000, 2, np.inf),
                                ("rxs -inf in S", X, Y, 5, 0, -np.inf)]:
        A2, B2 = A.copy(), None if B is None else B.copy()
        (A2 if B is None else B2)[r, c] = v
        say("grid non-finite", name)
        with pytest.raises(ValueError):
            tj.join(A2, B2, T=0.01, method="grid")


def test_grid_blobs_medium(tj, oracle):
    """The grid workload's distribution (90% uniform, 10% in tight blobs) at 1/16 size."""
    from tensorjoin_variants_amd import workloads
    w = workloads.grid_rxs(n=262_144)
    st = check(tj, oracle, w.R, w.S, w.T, "grid_rxs 262144")
    say("grid stats", st["out"], st["grid_candidates"])


def test_grid_full_size(tj, oracle):
    """The full grid workload (4,194,304 x 4,194,304, d = 3) against the CPU grid oracle."""
    from tensorjoin_variants_amd import workloads
    w = workloads.grid_rxs()
    say("grid full: GPU")
    got, st = tj.join(w.R, w.S, T=w.T, method="grid", out_dtype=np.int64, return_stats=True)
    say("grid full: CPU oracle", st["out"], st["grid_candidates"])
    want = oracle.grid_join(w.R, w.S, w.T)
    assert got.shape == want.shape
    assert np.array_equal(got, want)


def test_grid_far_outlier(tj, oracle):
    """Outliers 1e10 away from points a few ulps apart: the cell side must
    widen so rounding of the cell coordinate cannot separate a pair.
    Checked against the brute-force oracle, not the grid oracle."""
    from test_oracle import far_outlier_sets
    for name, X, T in far_outlier_sets():
        say("grid far outlier", name)
        want = oracle.join(X, None, T)
        got = tj.join(X, None, T=T, method="grid", out_dtype=np.int64)
        assert np.array_equal(got, want), name
        Y = X[::-1].copy()
        got = tj.join(X, Y, T=T, method="grid", out_dtype=np.int64)
        assert np.array_equal(got, oracle.join(X, Y, T)), name


@pytest.mark.parametrize("d", [2, 3, 8])
def test_gdsjoin_matches_oracle(tj, oracle, d):
    """GDS-Join's Python call (GDS/gdsjoingpu.py:146-232): counts per point
    including itself, and every point's neighbours (ascending here)."""
    rng = np.random.default_rng(60 + d)
    n = 4000
    X = rng.random((n, d)).astype(np.float32)
    X[n // 2:] = (X[: n // 2] + rng.normal(0, 0.003, (n // 2, d))).astype(np.float32)
    eps = {2: 0.01, 3: 0.03, 8: 0.3}[d]
    say("gdsjoin", d)
    counts, table = tj.gdsjoin(X.reshape(-1), eps, d, "float")
    T = float(np.float32(eps)) ** 2
    want = oracle.join(X, X, T)  # every directed pair, self pairs included
    assert counts.dtype == np.uint32 and table.dtype == np.uint32
    assert np.array_equal(counts, np.bincount(want[:, 0], minlength=n))
    assert np.array_equal(table.astype(np.int64), want[:, 1])
    assert int(counts.min()) >= 1


def test_gdsjoin_double_input(tj, oracle):
    rng = np.random.default_rng(66)
    X = rng.random((500, 3)).astype(np.float32)
    say("gdsjoin double")
    c1, t1 = tj.gdsjoin(X.astype(np.float64).ravel(), 0.05, 3, "double")
    c2, t2 = tj.gdsjoin(X.ravel(), 0.05, 3, "float")
    want = oracle.join(X, X, 0.05 ** 2)
    assert np.array_equal(t1.astype(np.int64), want[:, 1])
    assert np.array_equal(c1, np.bincount(want[:, 0], minlength=500))
    # float32 EPSILON rounding differs from the double one only at ties
    assert abs(len(t1) - len(t2)) <= 4
    with pytest.raises(ValueError):
        tj.gdsjoin(np.full(6, 0.1), 0.05, 3, "double")  # 0.1 is not a float32


def test_grid_long_rows(tj, oracle):
    """Rows beyond the wave sort (2,048 ids): one workgroup per row sorts
    chunks of 4,096 ids in LDS and merges them (rows of up to 9,000 ids, so
    two merge passes)."""
    rng = np.random.default_rng(70)
    blob = lambda n: (0.5 + rng.normal(0, 0.001, (n, 3))).astype(np.float32)
    X = blob(5000)
    st = check(tj, oracle, X, None, 0.05 ** 2, "one blob self-join (rows of 0..4999 ids)")
    assert st["out"] == 5000 * 4999 // 2
    R, S = blob(1500), blob(9000)
    S[::7] += np.float32(0.2)  # a few S points far away
    check(tj, oracle, R, S, 0.05 ** 2, "one blob R x S (rows of ~7,700 ids)")


def test_grid_capacity_retry(tj, oracle):
    from tensorjoin_variants_amd.join import Plan, _as_device
    rng = np.random.default_rng(71)
    X = rng.random((3000, 2)).astype(np.float32)
    T = 0.02 ** 2
    say("grid capacity retry")
    plan = Plan(3000, None, 2, "grid", "none", out_capacity=16, amb_capacity=0)
    res = plan.run(_as_device(X), None, T)
    assert plan.out_capacity >= res.stats["out"] > 16
    assert np.array_equal(res.numpy().astype(np.int64), oracle.join(X, None, T))


def test_grid_stage_abi(tj, oracle):
    """tj_grid_join through ctypes, its own workspace query, R x S and self."""
    import ctypes as C
    import torch
    from tensorjoin_variants_amd import _lib
    L = _lib.load()
    rng = np.random.default_rng(72)
    R = rng.random((6000, 3)).astype(np.float32)
    S = rng.random((5000, 3)).astype(np.float32)
    T = 0.03 ** 2
    for name, A, B in [("rxs", R, S), ("self", R, None)]:
        say("grid stage", name)
        want = oracle.join(A, B, T)
        cap = len(want) + 10
        nS = 0 if B is None else B.shape[0]
        need = C.c_size_t()
        assert L.tj_grid_workspace_bytes(A.shape[0], nS, cap, C.byref(need)) == 0
        ws = torch.empty(need.value, dtype=torch.uint8, device="cuda")
        Ad = torch.from_numpy(A).cuda()
        Bd = None if B is None else torch.from_numpy(B).cuda()
        pairs = torch.full((cap, 2), -1, dtype=torch.int64, device="cuda")
        stats = torch.zeros(_lib.STATS_SLOTS, dtype=torch.int64, device="cuda")
        _lib.check(L.tj_grid_join(Ad.data_ptr(), A.shape[0], None if Bd is None else Bd.data_ptr(), nS, 3, T,
                                  ws.data_ptr(), need.value, cap, pairs.data_ptr(), 1, stats.data_ptr(),
                                  torch.cuda.current_stream().cuda_stream), "tj_grid_join")
        torch.cuda.synchronize()
        st = _lib.decode_stats(stats.cpu().numpy().view(np.uint64))
        assert st["out"] == len(want)
        assert np.array_equal(pairs[: len(want)].cpu().numpy(), want)
        assert bool((pairs[len(want):] == -1).all())  # nothing past the output
        small = C.c_size_t(need.value - 256)
        assert L.tj_grid_join(Ad.data_ptr(), A.shape[0], None if Bd is None else Bd.data_ptr(), nS, 3, T,
                              ws.data_ptr(), small.value, cap, pairs.data_ptr(), 1, stats.data_ptr(), None) == -4


def test_grid_plan_graph_replay(tj):
    """A grid plan captured into a hipGraph (tj_plan_capture) and replayed
    three times, eager work in between: every replay gives the eager pairs
    (the grid's clears are fill kernels, its counts device-side)."""
    import torch
    from tensorjoin_variants_amd import _lib
    from tensorjoin_variants_amd import workloads
    from tensorjoin_variants_amd.join import Plan, _as_device, _ptr
    w = workloads.grid_rxs(n=65_536)
    Rd, Sd = _as_device(w.R), _as_device(w.S)
    plan = Plan(w.R.shape[0], w.S.shape[0], 3, "grid", "none", out_capacity=1 << 20, amb_capacity=0)
    eager = plan.run(Rd, Sd, w.T, np.int64).numpy()
    pairs = torch.empty((plan.out_capacity, 2), dtype=torch.int64, device="cuda")
    L = _lib.load()
    say("grid plan graph capture")
    _lib.check(L.tj_plan_capture(plan._plan, _ptr(Rd), _ptr(Sd), w.T, _ptr(plan._ws), _ptr(pairs), 1,
                                 _ptr(plan._stats)), "tj_plan_capture")
    for rep in range(3):
        pairs.fill_(-1)
        junk = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        say("grid plan graph replay", rep)
        _lib.check(L.tj_plan_replay(plan._plan, torch.cuda.current_stream().cuda_stream), "tj_plan_replay")
        torch.cuda.synchronize()
        n = int(plan._stats[0].item())
        assert n == eager.shape[0]
        assert np.array_equal(pairs[:n].cpu().numpy(), eager)
        del junk
