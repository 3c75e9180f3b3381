"""The exported pipeline stages chained through the C ABI (ctypes), the way a
caller composing its own pipeline would drive them -- the stage-by-stage
launches of TG/src/frozen_driver.py:69-77 driven from
TG/src/g2_operator_r1.py:27-41 (stage 1, stage 2, terminal):

    tj_scale_scan -> tj_params -> tj_rowmeta -> tj_screen -> tj_certify
    -> tj_sort_keys -> tj_unpack_pairs

and tj_refine_fp64 consuming the screen's own candidate queue (count at
stats[TJ_ST_CANDIDATES]).  Every result is compared bit for bit with the CPU
oracle.
"""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ST_OUT, ST_CAND, ST_VIOL = 0, 4, 7


def say(*a):
    print("[case]", *a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available(), "no GPU visible"
    from tensorjoin_variants_amd import _lib
    return torch, _lib, _lib.load()


def bits_for(n):
    b = 0
    while (1 << b) < n:
        b += 1
    return b


class Stages:
    """Buffers and calls of one staged join."""

    def __init__(self, env, R, S, T, prec, out_cap=1 << 16, cand_cap=1 << 18):
        torch, lib, L = env
        self.torch, self.lib, self.L = torch, lib, L
        self.R = torch.from_numpy(R).cuda()
        self.S = None if S is None else torch.from_numpy(S).cuda()
        self.nR, self.d = R.shape
        self.nS = 0 if S is None else S.shape[0]
        self.ncols = self.nR if S is None else self.nS
        self.jbits = bits_for(max(self.ncols, 2))
        self.T = float(T)
        self.prec = lib.PREC[prec]
        self.x3 = prec.endswith("x3")
        self.stats = torch.zeros(lib.STATS_SLOTS, dtype=torch.int64, device="cuda")
        self.maxabs = torch.zeros(1, dtype=torch.int64, device="cuda")
        self.params = torch.zeros(64, dtype=torch.int64, device="cuda")  # 512 bytes >= JoinParams
        self.out_cap, self.cand_cap = out_cap, cand_cap
        self.out = torch.zeros(out_cap, dtype=torch.int64, device="cuda")
        self.cand = torch.zeros(cand_cap, dtype=torch.int64, device="cuda")
        self.cand_s = torch.zeros(cand_cap, dtype=torch.float32, device="cuda")
        self.stream = torch.cuda.current_stream().cuda_stream

    def ok(self, rc, what):
        self.lib.check(rc, what)

    def side(self, x, n):
        torch, L = self.torch, self.L
        qb, mb = C.c_size_t(), C.c_size_t()
        self.ok(L.tj_rowmeta_sizes(n, self.d, self.prec, C.byref(qb), C.byref(mb)), "tj_rowmeta_sizes")
        qh = torch.zeros(max(qb.value, 1), dtype=torch.uint8, device="cuda")
        ql = torch.zeros(max(qb.value, 1), dtype=torch.uint8, device="cuda") if self.x3 else None
        meta = torch.zeros(max(mb.value // 4, 1), dtype=torch.float32, device="cuda")
        self.ok(L.tj_rowmeta(x.data_ptr(), n, self.d, self.prec, self.params.data_ptr(), qh.data_ptr(),
                             None if ql is None else ql.data_ptr(), meta.data_ptr(), self.stats.data_ptr(),
                             self.stream), "tj_rowmeta")
        assert mb.value == 48 * n
        return qh, ql, meta

    def metadata(self):
        L, st = self.L, self.stats.data_ptr()
        self.ok(L.tj_scale_scan(self.R.data_ptr(), self.nR, self.d, self.maxabs.data_ptr(), st, self.stream),
                "tj_scale_scan")
        if self.S is not None:
            self.ok(L.tj_scale_scan(self.S.data_ptr(), self.nS, self.d, self.maxabs.data_ptr(), st, self.stream),
                    "tj_scale_scan")
        self.ok(L.tj_params(self.maxabs.data_ptr(), self.T, self.d, self.prec, self.params.data_ptr(), st,
                            self.stream), "tj_params")
        self.a = self.side(self.R, self.nR)
        self.b = self.side(self.S, self.nS) if self.S is not None else (None, None, None)

    def screen(self, audit):
        qa_h, qa_l, ma = self.a
        qb_h, qb_l, mb = self.b
        p = lambda t: None if t is None else t.data_ptr()
        self.ok(self.L.tj_screen(self.prec, self.d, p(qa_h), p(qa_l), p(ma), self.nR, p(qb_h), p(qb_l), p(mb),
                                 self.nS, self.params.data_ptr(), self.out.data_ptr(), self.out_cap,
                                 self.cand.data_ptr(), self.cand_s.data_ptr(), self.cand_cap,
                                 self.lib.FLAG_AUDIT if audit else 0, self.stats.data_ptr(), self.stream),
                "tj_screen")

    def certify(self, audit):
        self.ok(self.L.tj_certify(self.R.data_ptr(), None if self.S is None else self.S.data_ptr(), self.d,
                                  self.ncols, self.cand.data_ptr(), self.cand_s.data_ptr(), self.cand_cap,
                                  self.a[2].data_ptr(), None if self.S is None else self.b[2].data_ptr(),
                                  self.params.data_ptr(), self.out.data_ptr(), self.out_cap,
                                  self.lib.FLAG_AUDIT if audit else 0, self.stats.data_ptr(), self.stream),
                "tj_certify")

    def sorted_pairs(self, keys, count_ptr, cap):
        torch, L = self.torch, self.L
        need = C.c_size_t()
        bits = bits_for(max(self.nR, 2)) + self.jbits
        self.ok(L.tj_sort_keys(None, C.byref(need), None, None, cap, bits, None), "tj_sort_keys query")
        temp = torch.empty(need.value, dtype=torch.uint8, device="cuda")
        self.ok(L.tj_sort_keys(temp.data_ptr(), C.byref(need), keys.data_ptr(), count_ptr, cap, bits, self.stream),
                "tj_sort_keys")
        pairs = torch.zeros((cap, 2), dtype=torch.int64, device="cuda")
        self.ok(L.tj_unpack_pairs(keys.data_ptr(), count_ptr, cap, self.jbits, pairs.data_ptr(), self.lib.OUT["int64"],
                                  self.stream), "tj_unpack_pairs")
        torch.cuda.synchronize()
        return pairs

    def st(self):
        self.torch.cuda.synchronize()
        return self.lib.decode_stats(self.stats.cpu().numpy().view(np.uint64))


CASES = [("fp16", 128), ("bf16", 100), ("fp16x3", 64), ("fp8", 256), ("fp8x3", 32)]


def data(d, seed):
    rng = np.random.default_rng(seed)
    R = np.abs(rng.standard_normal((1500, d))).astype(np.float32)
    S = np.abs(rng.standard_normal((1100, d))).astype(np.float32)
    R[rng.random(R.shape) < 0.4] = 0
    S[rng.random(S.shape) < 0.4] = 0
    return R, S


@pytest.mark.parametrize("self_join", [False, True])
@pytest.mark.parametrize("prec,d", CASES)
def test_chained_stages(env, oracle, prec, d, self_join):
    R, S = data(d, 600 + d)
    if self_join:
        S = None
    M = oracle.d2_matrix(R[:300], (R if S is None else S)[:300])
    T = float(np.quantile(M[M > 0], 0.01))
    want = oracle.join(R, S, T)
    for audit in (False, True):
        say("stages", prec, d, "self" if self_join else "rxs", "audit" if audit else "fast")
        g = Stages(env, R, S, T, prec)
        g.metadata()
        g.screen(audit)
        g.certify(audit)
        st = g.st()
        assert st["out"] <= g.out_cap and st["candidates"] <= g.cand_cap, st
        assert st["violations"] == 0, st
        pairs = g.sorted_pairs(g.out, g.stats.data_ptr() + 8 * ST_OUT, g.out_cap)
        got = pairs[: st["out"]].cpu().numpy()
        assert np.array_equal(got, want), (prec, d, self_join, audit, got.shape, want.shape)


@pytest.mark.parametrize("self_join", [False, True])
@pytest.mark.parametrize("prec,d", [("fp16", 128), ("fp8", 64)])
def test_refine_consumes_screen_queue(env, oracle, prec, d, self_join):
    """Audit mode routes every survivor of the fast reject to the candidate
    queue, so tj_refine_fp64 over that queue (its count read from
    stats[TJ_ST_CANDIDATES]) must give exactly the join."""
    torch = env[0]
    R, S = data(d, 700 + d)
    if self_join:
        S = None
    M = oracle.d2_matrix(R[:300], (R if S is None else S)[:300])
    T = float(np.quantile(M[M > 0], 0.01))
    want = oracle.join(R, S, T)
    say("refine over the screen queue", prec, d, self_join)
    g = Stages(env, R, S, T, prec)
    g.metadata()
    g.screen(audit=True)
    st = g.st()
    assert st["out"] == 0 and st["candidates"] >= len(want), st
    out2 = torch.zeros(g.out_cap, dtype=torch.int64, device="cuda")
    st2 = torch.zeros(env[1].STATS_SLOTS, dtype=torch.int64, device="cuda")
    g.ok(g.L.tj_refine_fp64(g.R.data_ptr(), None if g.S is None else g.S.data_ptr(), g.d, g.jbits, T,
                            g.cand.data_ptr(), g.stats.data_ptr() + 8 * ST_CAND, g.cand_cap, out2.data_ptr(),
                            g.out_cap, st2.data_ptr(), g.stream), "tj_refine_fp64")
    torch.cuda.synchronize()
    s2 = env[1].decode_stats(st2.cpu().numpy().view(np.uint64))
    assert s2["fp64_refined"] == st["candidates"] and s2["out"] == s2["fp64_accept"] == len(want), s2
    pairs = g.sorted_pairs(out2, st2.data_ptr() + 8 * ST_OUT, g.out_cap)
    assert np.array_equal(pairs[: s2["out"]].cpu().numpy(), want)


def test_refine_count_clamped_to_capacity(env, oracle):
    """A count beyond the queue's capacity reads only the capacity's entries."""
    torch, lib, L = env
    rng = np.random.default_rng(5)
    X = rng.standard_normal((40, 8)).astype(np.float32)
    keys = np.array([(i << 6) | j for i in range(40) for j in range(40) if i < j], np.int64)
    Xd = torch.from_numpy(X).cuda()
    kd = torch.from_numpy(keys).cuda()
    cap = 100
    cnt = torch.tensor([len(keys) + 1000], dtype=torch.int64, device="cuda")
    out = torch.zeros(len(keys), dtype=torch.int64, device="cuda")
    stats = torch.zeros(lib.STATS_SLOTS, dtype=torch.int64, device="cuda")
    say("refine clamp")
    lib.check(L.tj_refine_fp64(Xd.data_ptr(), None, 8, 6, 1e30, kd.data_ptr(), cnt.data_ptr(), cap, out.data_ptr(),
                               len(keys), stats.data_ptr(), torch.cuda.current_stream().cuda_stream), "refine")
    torch.cuda.synchronize()
    s = lib.decode_stats(stats.cpu().numpy().view(np.uint64))
    assert s["fp64_refined"] == cap and s["out"] == cap
    assert sorted(out[:cap].cpu().numpy().tolist()) == sorted(keys[:cap].tolist())
