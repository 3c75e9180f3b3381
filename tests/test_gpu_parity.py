"""Parity of the HIP join against the CPU FP64 oracle (bit-exact pair sets).

Every case prints its name before touching the GPU, so a failing run names
its case.  Sizes keep the oracle to seconds.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PRECS = ["fp16", "bf16", "fp8", "fp16x3", "bf16x3", "fp8x3"]


def say(*a):
    print("[case]", *a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def tj():
    import torch
    assert torch.cuda.is_available(), "no GPU visible"
    import tensorjoin_variants_amd as T
    T.load_library()
    return T


def oracle_pairs(oracle, R, S, T):
    return oracle.join(R, S, T)


def run_check(tj, oracle, R, S, T, prec, method="certified", audit=True, name=""):
    """The production path (screen fast accept on) and, for certified joins,
    the audit path (every certified or fast-acceptable pair re-checked in
    FP64), both bit-exact against the oracle."""
    want = oracle_pairs(oracle, R, S, T)
    modes = [False, True] if (method == "certified" and audit) else [audit]
    for mode in modes:
        say(name, prec, method, "audit" if mode else "fast", R.shape, None if S is None else S.shape, T)
        got, st = tj.join(R, S, T=T, method=method, precision=prec, out_dtype=np.int64, audit=mode,
                          return_stats=True)
        assert got.shape == want.shape, (name, prec, mode, got.shape, want.shape, st)
        assert np.array_equal(got, want), (name, prec, mode)
        if method == "certified":
            assert st["violations"] == 0, st
            assert st["cert_accept"] + st["fp32_accept"] + st["fp64_accept"] == st["out"], st
            assert st["fp32_accept"] + st["fp32_reject"] <= st["ambiguous"], st
            if not mode:
                # outside audit mode only what the FP32 filter leaves reaches FP64
                assert st["fp64_refined"] == st["ambiguous"] - st["fp32_accept"] - st["fp32_reject"], st
            else:
                assert st["fp64_refined"] >= st["ambiguous"], st
    return st


def sample_T(oracle, R, S, q):
    M = oracle.d2_matrix(R[:200], (R if S is None else S)[:200])
    v = M[M > 0] if S is None else M.ravel()
    return float(np.quantile(v, q)) if v.size else 1.0


def test_device(tj):
    import torch
    name = torch.cuda.get_device_properties(0).gcnArchName
    say("device", name)
    assert "gfx950" in name


@pytest.mark.parametrize("d", [1, 3, 32, 100])
def test_plain_fp64(tj, oracle, d):
    rng = np.random.default_rng(100 + d)
    R = rng.standard_normal((1100, d)).astype(np.float32)
    S = rng.standard_normal((900, d)).astype(np.float32)
    T = sample_T(oracle, R, S, 0.01)
    run_check(tj, oracle, R, S, T, "fp16", method="plain_fp64", name=f"plain rxs d{d}")
    run_check(tj, oracle, R, None, T, "fp16", method="plain_fp64", name=f"plain self d{d}")


@pytest.mark.parametrize("d", [1, 3, 32, 100, 128, 512])
def test_plain_fp32(tj, oracle, d):
    """The FP32-first control: FP32 GEMM-form tile, rigorous interval, FP64
    for the undecided pairs -- bit-exact like every method."""
    rng = np.random.default_rng(150 + d)
    R = rng.standard_normal((1100, d)).astype(np.float32)
    S = rng.standard_normal((900, d)).astype(np.float32)
    T = sample_T(oracle, R, S, 0.01)
    st = run_check(tj, oracle, R, S, T, "none", method="plain_fp32", name=f"plain32 rxs d{d}")
    assert st["cert_accept"] + st["fp64_accept"] == st["out"], st
    assert st["ambiguous"] == st["fp64_refined"], st
    run_check(tj, oracle, R, None, T, "none", method="plain_fp32", name=f"plain32 self d{d}")


def test_plain_fp32_hard_cases(tj, oracle):
    """Near ties (partners at exactly T), magnitudes 1e-30..1e30 (the
    interval widens to FP64 where binary32 products underflow or overflow),
    exact zeros and duplicates."""
    rng = np.random.default_rng(160)
    n, d = 1500, 64
    R = rng.standard_normal((n, d)).astype(np.float32)
    step = rng.standard_normal((n, d)).astype(np.float32)
    step /= np.linalg.norm(step, axis=1, keepdims=True)
    S = (R + (0.5 * (1.0 + 1e-3 * (rng.random(n) - 0.5)))[:, None].astype(np.float32) * step).astype(np.float32)
    dd = oracle.d2_pairs(R, S, np.stack([np.arange(n), np.arange(n)], 1))
    T = float(np.sort(dd)[n // 2])
    st = run_check(tj, oracle, R, S, T, "none", method="plain_fp32", name="plain32 near ties")
    assert st["fp64_refined"] > 0, st
    base = rng.standard_normal((400, 20)).astype(np.float32)
    for scale in (1e-30, 1e-8, 1e12, 1e30):
        X = (base * np.float32(scale)).astype(np.float32)
        run_check(tj, oracle, X, None, sample_T(oracle, X, None, 0.02), "none", method="plain_fp32",
                  name=f"plain32 scale {scale}")
    Z = np.zeros((300, 16), np.float32)
    Z[150:] = rng.standard_normal((150, 16)).astype(np.float32)
    Z[200:250] = Z[150:200]
    run_check(tj, oracle, Z, None, 0.0, "none", method="plain_fp32", name="plain32 zeros and duplicates T=0")


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("d", [3, 16, 32, 100, 128, 256])
def test_certified_random(tj, oracle, prec, d):
    rng = np.random.default_rng(1000 + d)
    R = rng.standard_normal((1000, d)).astype(np.float32)
    S = rng.standard_normal((777, d)).astype(np.float32)
    T = sample_T(oracle, R, S, 0.01)
    run_check(tj, oracle, R, S, T, prec, name=f"rxs d{d}")
    run_check(tj, oracle, R, None, T, prec, name=f"self d{d}")


@pytest.mark.parametrize("prec", ["fp16", "bf16", "fp8", "fp8x3"])
def test_certified_d512(tj, oracle, prec):
    rng = np.random.default_rng(512)
    R = rng.standard_t(3, (700, 512)).astype(np.float32)
    S = rng.standard_t(3, (600, 512)).astype(np.float32)
    T = sample_T(oracle, R, S, 0.002)
    run_check(tj, oracle, R, S, T, prec, name="rxs d512 student-t")


@pytest.mark.parametrize("prec", PRECS)
def test_clusters_near_threshold(tj, oracle, prec):
    """Clustered data with T at a sampled distance: many pairs sit near the boundary."""
    rng = np.random.default_rng(7)
    c = rng.standard_normal((8, 24)) * 4
    X = (c[rng.integers(0, 8, 2000)] + rng.standard_normal((2000, 24))).astype(np.float32)
    M = oracle.d2_matrix(X[:300], X[300:600])
    T = float(np.sort(M.ravel())[2000])  # exactly a realised distance: an equality pair
    run_check(tj, oracle, X, None, T, prec, name="clusters T=realised d2")


@pytest.mark.parametrize("prec", PRECS)
def test_edge_shapes(tj, oracle, prec):
    rng = np.random.default_rng(5)
    X = rng.standard_normal((70, 8)).astype(np.float32)
    for nR, nS in [(1, 1), (1, 70), (70, 1), (33, 65), (257, 31)]:
        run_check(tj, oracle, X[:nR] if nR <= 70 else rng.standard_normal((nR, 8)).astype(np.float32),
                  X[:nS] if nS <= 70 else None, 4.0, prec, name=f"shape {nR}x{nS}")
    Y = rng.standard_normal((257, 8)).astype(np.float32)
    run_check(tj, oracle, Y, None, 6.0, prec, name="self 257")
    run_check(tj, oracle, Y[:1], None, 6.0, prec, name="self 1")


@pytest.mark.parametrize("prec", PRECS)
def test_empty(tj, prec):
    say("empty", prec)
    R = np.zeros((0, 16), np.float32)
    S = np.ones((10, 16), np.float32)
    assert tj.join(R, S, T=1.0, precision=prec).shape == (0, 2)
    assert tj.join(S, R, T=1.0, precision=prec).shape == (0, 2)


@pytest.mark.parametrize("prec", PRECS)
def test_adversarial_fixtures(tj, oracle, prec):
    """Closed-form fixtures of CEJ/src/run_g3c_b_r1_adversarial.py:365-446."""
    rng = np.random.default_rng(365)
    base = rng.standard_normal((300, 32)).astype(np.float32)
    T = sample_T(oracle, base, None, 0.02)
    want = oracle.join(base, None, T)
    say("negate", prec)
    assert np.array_equal(tj.join(-base, None, T=T, precision=prec), want)
    for e in (-8, 8):
        say("scale", e, prec)
        assert np.array_equal(tj.join(base * np.float32(2.0 ** e), None, T=T * 4.0 ** e, precision=prec), want)
    perm = rng.permutation(32)
    signs = rng.choice(np.array([-1, 1], np.float32), 32)
    run_check(tj, oracle, np.ascontiguousarray(base[:, perm] * signs), None, T, prec, name="dim perm sign flip")
    # all zero: every pair inside at T = 2^-40
    Z = np.zeros((200, 16), np.float32)
    say("all zero", prec)
    got = tj.join(Z, None, T=2.0 ** -40, precision=prec)
    assert len(got) == 200 * 199 // 2
    # signed max accumulator: +-1 rows, block-diagonal answer at T = 0
    X = np.ones((128, 64), np.float32)
    X[64:] = -1
    run_check(tj, oracle, X, None, 0.0, prec, name="signed max")


@pytest.mark.parametrize("prec", PRECS)
def test_magnitudes(tj, oracle, prec):
    rng = np.random.default_rng(77)
    base = rng.standard_normal((400, 20)).astype(np.float32)
    for scale in (1e-30, 1e-8, 1e12, 1e30):
        X = (base * np.float32(scale)).astype(np.float32)
        T = sample_T(oracle, X, None, 0.02)
        run_check(tj, oracle, X, None, T, prec, name=f"scale {scale}")
    # rows of wildly different magnitude, sparse rows, exact zeros
    X = base * rng.choice(np.array([1e-6, 1.0, 1e6], np.float32), (400, 1))
    X[::7] = 0
    X[1::5, 3:] = 0
    X = X.astype(np.float32)
    T = sample_T(oracle, X, None, 0.05)
    run_check(tj, oracle, X, None, T, prec, name="mixed magnitudes")


@pytest.mark.parametrize("prec", PRECS)
def test_known_answers_on_gpu(tj, oracle, prec):
    import json
    from pathlib import Path
    gold = json.loads((Path(__file__).parent / "golden" / "kat.json").read_text())["cases"]
    for c in gold:
        x = np.asarray(c["x"], np.float32)[None]
        y = np.asarray(c["y"], np.float32)[None]
        say("kat", c["name"], prec)
        if prec in ("fp16x3", "bf16x3") and x.shape[1] > 256:
            with pytest.raises(tj.TJError):  # the 16-bit hi/lo split serves d <= 256
                tj.join(x, y, T=c["T"], precision=prec)
            continue
        got = tj.join(x, y, T=c["T"], precision=prec, out_dtype=np.int64)
        assert len(got) == int(c["accept"]), c["name"]


def test_non_finite_rejected(tj):
    X = np.ones((10, 4), np.float32)
    X[3, 2] = np.nan
    say("nan input")
    with pytest.raises(ValueError):
        tj.join(X, None, T=1.0)


def test_capacity_retry(tj, oracle):
    from tensorjoin_variants_amd.join import Plan, _as_device
    rng = np.random.default_rng(9)
    X = rng.standard_normal((500, 8)).astype(np.float32)
    T = sample_T(oracle, X, None, 0.2)
    say("capacity retry")
    plan = Plan(500, None, 8, "certified", "fp16", out_capacity=16, amb_capacity=16)
    res = plan.run(_as_device(X), None, T)
    assert plan.out_capacity >= res.stats["out"] > 16
    assert np.array_equal(res.numpy().astype(np.int64), oracle.join(X, None, T))


def test_sort_keys(tj):
    import ctypes as C
    import torch
    from tensorjoin_variants_amd import _lib
    L = _lib.load()
    rng = np.random.default_rng(3)
    cases = [(0, 10, "uniform"), (1, 5, "uniform"), (1000, 12, "uniform"), (300_001, 40, "uniform"),
             (70_000, 63, "uniform"), (2_000_003, 38, "uniform"), (100_000, 20, "three values"),
             (500_000, 30, "skewed")]
    for n, bits, kind in cases:
        say("sort", n, bits, kind)
        if kind == "three values":
            keys = rng.choice(np.array([5, 2 ** 19 + 3, 77777], dtype=np.uint64), n)
        elif kind == "skewed":  # most keys share their low digits
            keys = (rng.integers(0, 2 ** (bits - 16), n, dtype=np.uint64) << np.uint64(16)) | np.uint64(0x1234)
            keys[::7] = rng.integers(0, 2 ** bits, keys[::7].size, dtype=np.uint64)
        else:
            keys = rng.integers(0, 2 ** bits, n, dtype=np.uint64)
        cap = n + 100
        dev = torch.zeros(cap, dtype=torch.int64, device="cuda")
        dev[:n] = torch.from_numpy(keys.view(np.int64))
        cnt = torch.tensor([n], dtype=torch.int64, device="cuda")
        need = C.c_size_t()
        assert L.tj_sort_keys(None, C.byref(need), None, None, cap, bits, None) == 0
        temp = torch.empty(need.value, dtype=torch.uint8, device="cuda")
        s = torch.cuda.current_stream().cuda_stream
        assert L.tj_sort_keys(temp.data_ptr(), C.byref(need), dev.data_ptr(), cnt.data_ptr(), cap, bits, s) == 0
        torch.cuda.synchronize()
        assert np.array_equal(dev[:n].cpu().numpy().view(np.uint64), np.sort(keys))


def test_canonical_and_matrix(tj, oracle):
    rng = np.random.default_rng(21)
    X = rng.standard_normal((600, 16)).astype(np.float32)
    T = sample_T(oracle, X, None, 0.01)
    want = oracle.canonical_u64(oracle.join(X, None, T), 600)
    for m in ("F16", "S16", "F8", "S8", "F16X3", "F32", "P32"):
        say("Matrix.run", m)
        ids, stats = tj.Matrix(T).run(m, X)
        assert np.array_equal(ids, want), m
        assert int(np.count_nonzero(ids // 600 == ids % 600)) == 600


def test_clusters_selfjoin_full(tj, oracle):
    """Workload 1 at full size: 8192 points, d = 32, int32 i<j output."""
    from tensorjoin_variants_amd import workloads
    w = workloads.clusters_selfjoin()
    want = oracle.join(w.R, None, w.T)
    for prec in ("fp16", "fp8", "fp16x3", "bf16"):
        say("clusters_selfjoin full", prec)
        got, st = tj.join(w.R, None, T=w.T, precision=prec, out_dtype=np.int32, audit=True, return_stats=True)
        assert got.dtype == np.int32
        assert np.array_equal(got.astype(np.int64), want), prec
        assert st["violations"] == 0
    say("clusters_selfjoin plain")
    assert np.array_equal(tj.join(w.R, None, T=w.T, method="plain_fp64").astype(np.int64), want)


@pytest.mark.parametrize("prec", ["fp16", "fp8"])
def test_certified_column_splits(tj, oracle, prec):
    """R x S wide enough that the screen splits S into L2-sized column ranges."""
    rng = np.random.default_rng(77)
    R = rng.standard_normal((3000, 256)).astype(np.float32)
    S = rng.standard_normal((20000, 256)).astype(np.float32)
    T = sample_T(oracle, R, S, 0.002)
    from tensorjoin_variants_amd.join import Plan
    desc = Plan(3000, 20000, 256, "certified", prec, 1024, 1024).describe()
    say("column splits", prec, desc["col_splits"], desc["cols_per_split"])
    assert desc["col_splits"] >= 2
    run_check(tj, oracle, R, S, T, prec, name="column splits")


@pytest.mark.parametrize("d", [16, 100, 256])
def test_fp32_filter_near_ties(tj, oracle, d):
    """Pairs whose squared distance sits within a few binary32 ulps of T:
    the certified FP32 filter must hand the undecidable ones to the
    sequential FP64 form (0 misses either way), and decide the rest."""
    rng = np.random.default_rng(300 + d)
    n = 3000
    R = rng.standard_normal((n, d)).astype(np.float32)
    step = rng.standard_normal((n, d)).astype(np.float32)
    step /= np.linalg.norm(step, axis=1, keepdims=True)
    # partner j = i at distance 0.5 (1 +- 1e-3): most partners are inside
    # the FP16 band but far enough from T for binary32, a few are not
    scale = 0.5 * (1.0 + 2e-3 * (rng.random(n) - 0.5))
    S = (R + scale[:, None].astype(np.float32) * step).astype(np.float32)
    dd = oracle.d2_pairs(R, S, np.stack([np.arange(n), np.arange(n)], 1))
    T = float(np.sort(dd)[n // 2])
    # T is one partner's exact sequential d2: a tie (kept, d2 <= T), and
    # near-ties around it
    say("fp32 filter near ties", d, T)
    for mode in (False, True):
        got, st = tj.join(R, S, T=T, method="certified", precision="fp16", out_dtype=np.int64, audit=mode,
                          return_stats=True)
        want = oracle.join(R, S, T)
        assert np.array_equal(got, want), (d, mode, got.shape, want.shape)
        assert st["violations"] == 0, st
        assert st["fp32_accept"] + st["fp32_reject"] > 0 and st["fp64_refined"] > 0, st


@pytest.mark.parametrize("prec,d", [("fp16", 128), ("bf16", 96), ("fp16", 64)])
def test_four_block_screen_uneven_norms(tj, oracle, prec, d):
    """R x S at the widths of the four-block (unseeded) screen bodies with
    heavy-tailed rows (Student-t, nu = 3) scaled by factors spanning 2^-6 to
    2^6: the row levels t of a lane's 16 rows differ widely, so the per-lane
    level lets many tiles through that only the rows' own levels reject."""
    rng = np.random.default_rng(400 + d)

    def rows(n):
        z = rng.standard_normal((n, d)) / np.sqrt(rng.chisquare(3.0, (n, 1)) / 3.0)
        return (z * np.exp2(rng.uniform(-6, 6, (n, 1)))).astype(np.float32)

    R, S = rows(2600), rows(2100)
    T = sample_T(oracle, R, S, 0.002)
    run_check(tj, oracle, R, S, T, prec, name="four-block uneven norms")
