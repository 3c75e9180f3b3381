"""Host-to-host pipeline (tj_host_*) against the CPU FP64 oracle, bit-exact:
R x S over several chunks (with a shorter last chunk), capacity growth that
re-runs a chunk, self-joins in pair and canonical-id form, the grid and
plain methods, audit mode, files read through vecio, and bad input."""
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


def quantile_T(oracle, R, S, q):
    M = oracle.d2_matrix(R[:300], (R if S is None else S)[:300])
    v = M[M > 0] if S is None else M.ravel()
    return float(np.quantile(v, q))


@pytest.mark.parametrize("prec,out", [("fp16", "int32"), ("fp8", "int64"), ("bf16", "int32")])
def test_rxs_chunks(tj, oracle, prec, out):
    rng = np.random.default_rng(11)
    R = rng.standard_normal((2300, 48)).astype(np.float32)
    S = rng.standard_normal((1700, 48)).astype(np.float32)
    T = quantile_T(oracle, R, S, 0.01)
    want = oracle.join(R, S, T)
    say("host rxs", prec, out, R.shape, S.shape, "chunk 600")
    got, st = tj.join_host(R, S, T=T, precision=prec, out=out, chunk_rows=600, return_stats=True)
    assert got.dtype == (np.int32 if out == "int32" else np.int64)
    assert got.shape == want.shape and np.array_equal(got, want), (got.shape, want.shape)
    assert st["violations"] == 0 and st["out"] == want.shape[0]
    assert st["cert_accept"] + st["fp32_accept"] + st["fp64_accept"] == st["out"]
    assert st["pairs"] == 2300 * 1700
    assert st["host_ms"]["total"] > 0


def test_capacity_growth_reruns_chunk(tj, oracle):
    # every pair is inside: 6.0M pairs per chunk overflow the first capacity
    rng = np.random.default_rng(12)
    R = rng.standard_normal((3000, 8)).astype(np.float32)
    S = rng.standard_normal((2400, 8)).astype(np.float32)
    T = 1e6
    say("host capacity growth", R.shape, S.shape)
    got = tj.join_host(R, S, T=T, precision="fp16", chunk_rows=2500)
    ii, jj = np.meshgrid(np.arange(3000), np.arange(2400), indexing="ij")
    want = np.stack([ii.ravel(), jj.ravel()], 1)
    assert got.shape == want.shape and np.array_equal(got, want)
    # and a selective join on the same plan shape afterwards (capacities kept)
    T2 = quantile_T(oracle, R, S, 0.001)
    plan = tj.HostPlan(3000, 2400, 8, "certified", "fp16", chunk_rows=2500)
    for t in (T, T2):
        g, _ = plan.run(R, S, t)
        w = oracle.join(R, S, t)
        assert np.array_equal(g, w)
    plan.close()


@pytest.mark.parametrize("d", [5, 64])
def test_self_join_pairs_and_canonical(tj, oracle, d):
    from tensorjoin_variants_amd import vecio
    rng = np.random.default_rng(13 + d)
    X = rng.standard_normal((1500, d)).astype(np.float32)
    T = quantile_T(oracle, X, None, 0.01)
    want = oracle.join(X, None, T)
    say("host self", d, X.shape)
    got = tj.join_host(X, T=T, precision="fp16")
    assert np.array_equal(got, want)
    ids, st = tj.join_host(X, T=T, precision="fp16", out="canonical", return_stats=True)
    ref = oracle.canonical_u64(want, 1500)
    assert ids.dtype == np.uint64 and np.array_equal(ids, ref)
    rep = vecio.validate_canonical(ids, 1500, expected_count=ref.size, expected_hash=vecio.sha256_u64(ref))
    assert rep["exact_contract_pass"], rep


def test_grid_and_plain_methods(tj, oracle):
    rng = np.random.default_rng(14)
    R = rng.uniform(0, 10, (4000, 3)).astype(np.float32)
    S = rng.uniform(0, 10, (3000, 3)).astype(np.float32)
    T = 0.05
    want = oracle.join(R, S, T)
    say("host grid", R.shape, S.shape)
    assert np.array_equal(tj.join_host(R, S, T=T, method="grid"), want)
    say("host plain", R.shape, S.shape)
    assert np.array_equal(tj.join_host(R, S, T=T, method="plain_fp64", chunk_rows=1000), want)
    want_self = oracle.join(R, None, T)
    assert np.array_equal(tj.join_host(R, T=T, method="grid"), want_self)


def test_audit_and_small_shapes(tj, oracle):
    rng = np.random.default_rng(15)
    for nR, nS in [(1, 1), (1, 700), (700, 1), (33, 65)]:
        R = rng.standard_normal((nR, 20)).astype(np.float32)
        S = rng.standard_normal((nS, 20)).astype(np.float32)
        T = 40.0
        say("host small", nR, nS)
        got, st = tj.join_host(R, S, T=T, precision="fp16", audit=True, return_stats=True)
        assert np.array_equal(got, oracle.join(R, S, T)) and st["violations"] == 0
    empty = tj.join_host(np.zeros((0, 20), np.float32), rng.standard_normal((5, 20)).astype(np.float32), T=1.0)
    assert empty.shape == (0, 2)
    empty = tj.join_host(rng.standard_normal((5, 20)).astype(np.float32), np.zeros((0, 20), np.float32), T=1.0)
    assert empty.shape == (0, 2)
    one = tj.join_host(rng.standard_normal((1, 20)).astype(np.float32), T=1.0, out="canonical")
    assert np.array_equal(one, np.array([0], np.uint64))


def test_fvecs_file_to_canonical(tj, oracle, tmp_path):
    from tensorjoin_variants_amd import vecio
    rng = np.random.default_rng(16)
    X = (rng.standard_normal((900, 32)) * 0.3).astype(np.float32)
    p = tmp_path / "x.fvecs"
    vecio.write_fvecs(p, X)
    Y = vecio.read_fvecs(p)
    T = 0.25 ** 2 * 32
    ref = oracle.canonical_u64(oracle.join(X, None, T), 900)
    say("host fvecs canonical", Y.shape)
    ids = tj.join_host(Y, T=T, out="canonical")
    rep = vecio.validate_canonical(ids, 900, expected_count=ref.size, expected_hash=vecio.sha256_u64(ref))
    assert rep["exact_contract_pass"], rep


def test_rejects_non_finite(tj):
    R = np.ones((50, 16), np.float32)
    R[7, 3] = np.inf
    say("host non-finite")
    with pytest.raises(tj.TJError, match="non-finite"):
        tj.join_host(R, np.ones((40, 16), np.float32), T=1.0, precision="fp16")
