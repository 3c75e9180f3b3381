"""tj_join_noalloc (the join whose scratch lives in a library-held device
array) against tj_join (caller-owned workspace) through the C ABI: bit-equal
pairs and equal counters, on two calls in a row, the second smaller, so the
second call runs on the scratch the first one left behind."""
import ctypes as C
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
    return torch, _lib, _lib.load()


def via_plan(env, R, S, T, method, prec, out_cap, amb_cap, flags):
    torch, lib, L = env
    plan = C.c_void_p()
    nS = 0 if S is None else S.shape[0]
    lib.check(L.tj_plan_create(C.byref(plan), R.shape[0], nS, R.shape[1], method, prec, out_cap, amb_cap, flags),
              "tj_plan_create")
    n = C.c_size_t()
    lib.check(L.tj_plan_workspace_bytes(plan, C.byref(n)), "workspace")
    ws = torch.empty(n.value, dtype=torch.uint8, device="cuda")
    stats = torch.zeros(lib.STATS_SLOTS, dtype=torch.int64, device="cuda")
    pairs = torch.full((out_cap, 2), -1, dtype=torch.int64, device="cuda")
    lib.check(L.tj_join(plan, R.data_ptr(), None if S is None else S.data_ptr(), T, ws.data_ptr(), pairs.data_ptr(),
                        lib.OUT["int64"], stats.data_ptr(), torch.cuda.current_stream().cuda_stream), "tj_join")
    torch.cuda.synchronize()
    L.tj_plan_destroy(plan)
    return pairs, stats, n.value


def via_noalloc(env, R, S, T, method, prec, out_cap, amb_cap, flags):
    torch, lib, L = env
    nS = 0 if S is None else S.shape[0]
    stats = torch.zeros(lib.STATS_SLOTS, dtype=torch.int64, device="cuda")
    pairs = torch.full((out_cap, 2), -1, dtype=torch.int64, device="cuda")
    lib.check(L.tj_join_noalloc(R.data_ptr(), R.shape[0], None if S is None else S.data_ptr(), nS, R.shape[1], method,
                                prec, T, out_cap, amb_cap, flags, pairs.data_ptr(), lib.OUT["int64"], stats.data_ptr(),
                                torch.cuda.current_stream().cuda_stream), "tj_join_noalloc")
    torch.cuda.synchronize()
    return pairs, stats


# (name, nR, nS (0 = self), d, method, precision, flags) -- each case runs a
# larger then a smaller input back to back
CASES = [("certified fp16 rxs", 3000, 2000, 128, 0, 1, 0),
         ("certified fp8x3 self audit", 2500, 0, 64, 0, 6, 1),
         ("plain fp64 rxs", 1500, 1200, 24, 1, 0, 0),
         ("grid self d3", 20000, 0, 3, 2, 0, 0)]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_noalloc_matches_join(env, oracle, case):
    torch, lib, L = env
    name, nR, nS, d, method, prec, flags = case
    rng = np.random.default_rng(800 + d)
    out_cap, amb_cap = 1 << 17, 1 << 17
    for step, shrink in enumerate((1, 3)):
        n1, n2 = nR // shrink, nS // shrink
        Rh = rng.random((n1, d)).astype(np.float32) if method == 2 else rng.standard_normal((n1, d)).astype(np.float32)
        Sh = None if nS == 0 else rng.standard_normal((n2, d)).astype(np.float32)
        if method == 2:
            T = 0.01 ** 2  # ~0.08 neighbours per point in the unit cube
        else:
            M = oracle.d2_matrix(Rh[:200], (Rh if Sh is None else Sh)[:200])
            T = float(np.quantile(M[M > 0], 0.004))
        R = torch.from_numpy(Rh).cuda()
        S = None if Sh is None else torch.from_numpy(Sh).cuda()
        say("noalloc", name, "call", step + 1, Rh.shape, None if Sh is None else Sh.shape)
        want_p, want_s, ws_bytes = via_plan(env, R, S, T, method, prec, out_cap, amb_cap, flags)
        assert ws_bytes <= lib.NOALLOC_SCRATCH_BYTES
        got_p, got_s = via_noalloc(env, R, S, T, method, prec, out_cap, amb_cap, flags)
        ws, gs = want_s.cpu().numpy(), got_s.cpu().numpy()
        n = int(ws[0])
        assert 0 < n <= out_cap
        assert int(gs[0]) == n
        # counters that do not depend on the order atomics land in
        keep = [0, 1, 2, 3, 4, 7, 9, 10, 11, 12, 16, 17, 18]
        assert np.array_equal(ws[keep], gs[keep]), (name, step, ws, gs)
        assert bool(torch.equal(got_p[:n], want_p[:n])), (name, step)
        assert np.array_equal(got_p[:n].cpu().numpy(), oracle.join(Rh, Sh, T)), (name, step)


def test_noalloc_rejects_oversize_and_flags(env):
    """Beyond the scratch array: TJ_ERR_WORKSPACE before any device work."""
    torch, lib, L = env
    one = C.c_void_p(1)
    say("noalloc oversize")
    # 1M x 128 FP16 rows need ~300 MiB of codes and metadata
    assert L.tj_join_noalloc(one, 1 << 20, one, 1 << 18, 128, 0, 1, 1.0, 1024, 1024, 0, one, 1, one, None) == -4
    assert b"scratch" in L.tj_last_error_string()
    assert L.tj_join_noalloc(one, 100, one, 100, 8, 0, 1, 1.0, 1024, 1024, 2, one, 1, one, None) == -1
