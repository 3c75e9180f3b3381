"""The d = 128 certified-screen workload (the metric's configuration) on the GPU.

Rows: half-normal |N(0,1)| coordinates, 40% exact zeros, L2-normalised; eps
at selectivity 1e-4 (workloads.certified_d128).  The screen's behaviour
depends on this distribution (all coordinates >= 0, norms ~1, many exact
zeros that the hi/lo split and the FP8 codes represent exactly), so it is
checked here on its own rows:

* at reduced size (32,768 x 8,192, same distribution and selectivity)
  against the CPU oracle, bit-exact, in the production and the audit path,
  for every screen precision;
* at full size (1,048,576 x 262,144) against the oracle on the first 2,048
  rows of R (every panel shape and column split of the full plan runs), and
  against the plain FP64 GPU method over all 2^38 pairs, with 0 certificate
  violations.

The exact-output gate follows CEJ/src/run_g3b_r1_gpu_analytic_certificate.py:378-432.
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


@pytest.fixture(scope="module")
def small(oracle):
    from tensorjoin_variants_amd import workloads
    w = workloads.certified_d128(nR=32_768, nS=8_192, samples=200_000)
    return w, oracle.join(w.R, w.S, w.T)


@pytest.mark.parametrize("prec", ["fp16", "bf16", "fp16x3", "bf16x3", "fp8", "fp8x3"])
def test_d128_workload_reduced(tj, small, prec):
    w, want = small
    # the selectivity gives ~27 K pairs here; a near-empty set would test nothing
    assert 5_000 < len(want) < 200_000, len(want)
    for audit in (False, True):
        say("d128 workload 32768x8192", prec, "audit" if audit else "fast")
        got, st = tj.join(w.R, w.S, T=w.T, method="certified", precision=prec, out_dtype=np.int32, audit=audit,
                          return_stats=True)
        assert got.dtype == np.int32
        assert got.shape == want.shape, (prec, audit, got.shape, want.shape, st)
        assert np.array_equal(got.astype(np.int64), want), (prec, audit)
        assert st["violations"] == 0, st
        assert st["cert_accept"] + st["fp32_accept"] + st["fp64_accept"] == st["out"], st
        say("  ambiguous", st["ambiguous"], "fp64", st["fp64_refined"], "max err", st["max_err"])


def test_d128_workload_full_size(tj, oracle):
    import torch
    from tensorjoin_variants_amd import workloads
    from tensorjoin_variants_amd.join import Plan, _as_device
    say("d128 workload full: generating")
    w = workloads.certified_d128()
    R, S = _as_device(w.R), _as_device(w.S)
    nR, nS = w.R.shape[0], w.S.shape[0]
    say("d128 workload full: certified fp16")
    plan = Plan(nR, nS, 128, "certified", "fp16", out_capacity=40_000_000, amb_capacity=8_000_000)
    res = plan.run(R, S, w.T, np.int32)
    st = res.stats
    assert st["violations"] == 0, st
    assert st["pairs"] == nR * nS
    cert = res.pairs
    say("d128 workload full: plain fp64", st["out"], st["ambiguous"])
    pplan = Plan(nR, nS, 128, "plain_fp64", "none", out_capacity=plan.out_capacity, amb_capacity=0)
    pres = pplan.run(R, S, w.T, np.int32)
    assert pres.stats["out"] == st["out"]
    assert bool(torch.equal(cert, pres.pairs))
    say("d128 workload full: oracle on the first 2048 rows")
    k = 2048
    want = oracle.join(w.R[:k], w.S, w.T)
    n_k = int(torch.searchsorted(cert[:, 0].contiguous(), torch.tensor([k], dtype=cert.dtype, device=cert.device)))
    got = cert[:n_k].cpu().numpy().astype(np.int64)
    assert np.array_equal(got, want)
    assert len(want) > 0


def test_d128_workload_graph_replay(tj, small):
    """The timed form of the metric join: the whole join captured once into
    a hipGraph (Plan.capture) and replayed, eager work between replays, on
    the reduced workload; every replay gives the oracle's pairs."""
    import torch
    from tensorjoin_variants_amd.join import Plan, _as_device
    w, want = small
    R, S = _as_device(w.R), _as_device(w.S)
    plan = Plan(w.R.shape[0], w.S.shape[0], 128, "certified", "fp16", out_capacity=1 << 17, amb_capacity=1 << 17)
    res = plan.run(R, S, w.T, np.int32)
    assert np.array_equal(res.numpy().astype(np.int64), want)
    pairs = torch.empty((plan.out_capacity, 2), dtype=torch.int32, device="cuda")
    say("d128 workload graph capture")
    plan.capture(R, S, w.T, pairs, "int32")
    for rep in range(3):
        pairs.fill_(-1)
        junk = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        say("d128 workload graph replay", rep)
        plan.replay()
        torch.cuda.synchronize()
        n = int(plan._stats[0].item())
        assert n == len(want)
        assert np.array_equal(pairs[:n].cpu().numpy().astype(np.int64), want)
        del junk
