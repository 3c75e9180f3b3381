"""Seeded random sweep against the oracle: batches of uneven self-joins
(mixed widths, duplicates, FP16/FP8, tile pruning on and off), grid joins
(d = 1..3, scales 1e-3..1e3, dense blobs, self and R x S) and single
certified joins (every precision, d up to 512, audit and production
paths), with eps at random percentiles of the sampled distances (0 included:
the closest pairs and exact duplicates).  Bit-exact pair sets."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tj():
    import torch
    assert torch.cuda.is_available()
    import tensorjoin_variants_amd as T
    T.load_library()
    return T


def pick_T(X, Y, rng, q):
    n = X.shape[0]
    Z = X if Y is None else Y
    a = rng.integers(0, n, 2000)
    b = rng.integers(0, Z.shape[0], 2000)
    return float(np.quantile(((X[a].astype(np.float64) - Z[b]) ** 2).sum(1), q))


@pytest.mark.parametrize("r", range(45))
def test_random_case(tj, oracle, r):
    rng = np.random.default_rng(7000 + r)
    kind = r % 3
    if kind == 0:
        sets, Ts = [], []
        for _ in range(int(rng.integers(1, 24))):
            n = int(rng.choice([1, 2, 3, 7, 31, 64, 65, 127, 200, 513, 1000, 2049]))
            d = int(rng.choice([1, 3, 8, 16, 17, 32, 63, 64, 100, 128, 200, 256]))
            c = rng.normal(0, 3, (4, d))
            X = (c[rng.integers(0, 4, n)] + rng.normal(0, 1, (n, d))).astype(np.float32)
            if rng.random() < 0.2:
                X[: n // 2] = X[n // 2: 2 * (n // 2)]
            sets.append(X)
            Ts.append(pick_T(X, None, rng, rng.choice([0.0, 0.01, 0.2])) if n > 1 else 1.0)
        prec, prune = str(rng.choice(["fp16", "fp8"])), bool(rng.random() < 0.5)
        print("[case] batch", r, len(sets), prec, prune, file=sys.stderr, flush=True)
        got = tj.batched_join(sets, T=np.array(Ts), precision=prec, prune=prune)
        for k, (X, T) in enumerate(zip(sets, Ts)):
            assert np.array_equal(got[k].astype(np.int64), oracle.join(X, None, T)), (r, k, X.shape)
    elif kind == 1:
        d = int(rng.integers(1, 4))
        nR, nS = int(rng.choice([1, 50, 999, 5000, 20000])), int(rng.choice([1, 77, 3000, 20000]))
        scale = float(rng.choice([1e-3, 1.0, 1e3]))
        R = (rng.random((nR, d)) * scale).astype(np.float32)
        S = (rng.random((nS, d)) * scale).astype(np.float32)
        if rng.random() < 0.5:
            cb = rng.random(d) * scale
            S[: nS // 3] = (cb + rng.normal(0, 0.003 * scale, (nS // 3, d))).astype(np.float32)
            R[: nR // 3] = (cb + rng.normal(0, 0.003 * scale, (nR // 3, d))).astype(np.float32)
        Sx = None if rng.random() < 0.4 else S
        T = pick_T(R, Sx, rng, rng.choice([0.0, 0.001, 0.01]))
        print("[case] grid", r, d, nR, None if Sx is None else nS, file=sys.stderr, flush=True)
        got = tj.join(R, Sx, T=T, method="grid", out_dtype=np.int64)
        assert np.array_equal(got, oracle.join(R, Sx, T)), r
    else:
        d = int(rng.choice([2, 16, 33, 64, 128, 200, 256, 512]))
        nR, nS = int(rng.choice([1, 100, 1000, 4000])), int(rng.choice([1, 100, 2000]))
        R = rng.standard_normal((nR, d)).astype(np.float32) * np.float32(rng.choice([1e-3, 1.0, 1e3]))
        S = rng.standard_normal((nS, d)).astype(np.float32) * np.float32(R.std() if nR > 1 else 1.0)
        Sx = None if rng.random() < 0.4 else S
        prec = str(rng.choice(["fp16", "bf16", "fp16x3", "bf16x3", "fp8", "fp8x3"]))
        if prec in ("fp16x3", "bf16x3") and d > 256:
            prec = prec[:-2]  # the 16-bit hi/lo split serves d <= 256
        T = pick_T(R, Sx, rng, rng.choice([0.0, 0.001, 0.05]))
        audit = bool(rng.random() < 0.3)
        print("[case] certified", r, d, nR, None if Sx is None else nS, prec, audit, file=sys.stderr, flush=True)
        got = tj.join(R, Sx, T=T, method="certified", precision=prec, out_dtype=np.int64, audit=audit)
        assert np.array_equal(got, oracle.join(R, Sx, T)), r
