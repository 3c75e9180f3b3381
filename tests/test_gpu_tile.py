"""Screen tiles (row blocks per wave x column split width) are output-neutral.

Each tile of a certified plan -- the default, two row blocks for the
four-block bodies, column splits from one 32-column block to all of S --
returns the oracle's pair set bit for bit, and the tile tuner leaves its
winner on the plan with the same result.  The reference fixes its tile
(CEJ/src/run_g2b_tensorjoin_public.py:41-42); here the tile is a plan knob.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available(), "no GPU visible"
    from tensorjoin_variants_amd import _lib
    _lib.load()
    return torch


def rows(seed, n, d):
    rng = np.random.default_rng(seed)
    X = np.abs(rng.standard_normal((n, d))).astype(np.float32)
    X[rng.random((n, d)) < 0.4] = 0.0
    X /= np.maximum(np.linalg.norm(X, axis=1, keepdims=True), 1e-30).astype(np.float32)
    return X


def threshold(oracle, R, S, q=2e-3):
    M = oracle.d2_matrix(R[:300], (R if S is None else S)[:300])
    return float(np.quantile(M[M > 0], q))


@pytest.mark.parametrize("prec,d,self_join", [("fp16", 128, False), ("bf16", 64, False), ("fp16", 96, False),
                                              ("fp16", 64, True), ("fp8", 128, False), ("fp16x3", 128, False)])
def test_every_tile_gives_the_oracle_pairs(env, oracle, prec, d, self_join):
    from tensorjoin_variants_amd.join import Plan, _as_device
    R = rows(70 + d, 3000, d)
    S = None if self_join else rows(71 + d, 2500, d)
    T = threshold(oracle, R, S)
    want = oracle.join(R, S, T)
    assert len(want) > 0
    Rd, Sd = _as_device(R), None if S is None else _as_device(S)
    plan = Plan(R.shape[0], None if S is None else S.shape[0], d, "certified", prec, 1 << 18, 1 << 20)
    rb0 = plan.describe()["row_blocks"]
    nS = R.shape[0] if S is None else S.shape[0]
    for rb in [0, rb0] + ([2] if rb0 == 4 else []):
        for cols in (0, 32, 1000, 2048, nS, 1 << 20):
            say("tile", prec, d, "self" if self_join else "rxs", rb, cols)
            plan.set_tile(rb, cols)
            desc = plan.describe()
            assert desc["row_blocks"] == (rb or rb0)
            assert desc["col_splits"] * desc["cols_per_split"] >= nS
            got = plan.run(Rd, Sd, T).numpy()
            assert np.array_equal(got, want), (prec, d, rb, cols)
    plan.close()


def test_tile_tuner(env, oracle):
    from tensorjoin_variants_amd.autotune import Autotuner
    from tensorjoin_variants_amd.join import Plan, _as_device
    R, S = rows(80, 4096, 128), rows(81, 4096, 128)
    T = threshold(oracle, R, S)
    want = oracle.join(R, S, T)
    Rd, Sd = _as_device(R), _as_device(S)
    plan = Plan(4096, 4096, 128, "certified", "fp16", 1 << 18, 1 << 20)
    plan.run(Rd, Sd, T)  # size the capacities
    tuner = Autotuner(cache=None)
    say("tile tuner")
    rep = tuner.tune_tile(plan, Rd, Sd, T, reps=1)
    tried = {(m["row_blocks"], m["cols_per_split"]) for m in rep["ms"]}
    assert {4, 2} <= {rb for rb, _ in tried} and len(tried) >= 4
    win = (rep["winner"]["row_blocks"], rep["winner"]["cols_per_split"])
    assert win in tried and plan.tile == win
    assert tuner.lookup_tile(128, 4096, 4096, "fp16") == win
    assert np.array_equal(plan.run(Rd, Sd, T).numpy(), want)
    plan.close()
