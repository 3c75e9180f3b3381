"""Inputs whose rows do not start on 16-byte boundaries.

The per-row passes (scale scan, row metadata, spatial order, panel balls,
FP32 filter) read rows with 16-byte loads when a row's start is 16-byte
aligned and fall back to 4-byte loads otherwise.  Every other test passes
freshly allocated device arrays (aligned); here each set is a view one
float into its buffer, so the rows start at addresses = 4 (mod 16) and the
fallback paths run.  Results are checked against the CPU oracle, bit-exact.
"""
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


def offset_view(X):
    """A device copy of X that starts one float into its allocation."""
    import torch
    buf = torch.zeros(X.size + 1, dtype=torch.float32, device="cuda")
    buf[1:] = torch.from_numpy(np.ascontiguousarray(X).ravel()).cuda()
    v = buf[1:].view(X.shape)
    assert v.data_ptr() % 16 == 4
    return buf, v


@pytest.mark.parametrize("prec,prune", [("fp16", True), ("fp16", False), ("fp8", True)])
def test_batch_unaligned_rows(tj, oracle, prec, prune):
    import torch
    from tensorjoin_variants_amd import BatchPlan
    from tensorjoin_variants_amd.batch import initial_capacities
    from tensorjoin_variants_amd.thresholds import percentile_eps
    rng = np.random.default_rng(29)
    sets, Ts = [], []
    for n, d in [(300, 16), (1000, 64), (2049, 256), (64, 16), (700, 64), (513, 256), (90, 32)]:
        X = mixture(rng, n, d)
        e = percentile_eps(X, None, 0.01, min(4000, n * (n - 1) // 2), rng)
        sets.append(X)
        Ts.append(e * e)
    ns, ds = [x.shape[0] for x in sets], [x.shape[1] for x in sets]
    oc, cc = initial_capacities(ns)
    plan = BatchPlan(ns, ds, prec, oc, cc, prune=prune)
    views = [offset_view(X) for X in sets]
    ptrs = torch.tensor([v.data_ptr() for _, v in views], dtype=torch.int64, device="cuda")
    Td = torch.from_numpy(np.asarray(Ts, dtype=np.float64)).cuda()
    say("batch unaligned", prec, prune)
    plan.enqueue(ptrs, Td)
    got, st = plan.results()
    assert st["invalid"] == 0
    assert st["out"] <= plan.out_capacity and st["candidates"] <= plan.cand_capacity
    for k, (X, T, g) in enumerate(zip(sets, Ts, got)):
        want = oracle.join(X, None, float(T))
        assert np.array_equal(g.astype(np.int64), want), (prec, prune, k, X.shape)


@pytest.mark.parametrize("prec,d", [("fp16", 128), ("fp8", 64), ("bf16x3", 32)])
def test_join_unaligned_rows(tj, oracle, prec, d):
    from tensorjoin_variants_amd.thresholds import percentile_eps
    rng = np.random.default_rng(31 + d)
    R = rng.standard_normal((3000, d)).astype(np.float32)
    S = rng.standard_normal((2000, d)).astype(np.float32)
    e = percentile_eps(R, S, 0.002, 20000, rng)
    T = e * e
    (_, Rv), (_, Sv) = offset_view(R), offset_view(S)
    say("join unaligned", prec, d)
    got = tj.join(Rv, Sv, T=T, method="certified", precision=prec, out_dtype=np.int64)
    assert np.array_equal(got, oracle.join(R, S, T)), (prec, d)
