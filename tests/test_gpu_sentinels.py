"""Sentinel prefix probe: no stage writes outside [0, capacity) of its
output buffers.

Restates the reference's prefix probe (TG/src/retained_ops.py:107-127):
seed 927, rows 4..1023 uniform in (-0.08, 0.08), a pair at exactly eps
(row 128 = eps e0 against the zero rows), a near-duplicate of it (row 129,
offsets 2^-28), the first 1,024 rows copied into a second block with its
first row zeroed, and every output buffer padded with 32 entries of -17 on
each side.  Here the buffers are the screen's output keys and candidate
queue (keys and screen values), the certified output after tj_certify, the
sorted keys and the unpacked pairs, and tj_join's pairs; each chain runs
with capacities exactly at the need and at half the need (overflow: the
stats report the need, nothing is written past the capacity), and the
results at the exact capacity match the oracle.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from test_gpu_stages import Stages, bits_for

pytestmark = pytest.mark.gpu

PAD = 32
SENT = -17


def say(*a):
    print("[case]", *a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available(), "no GPU visible"
    from tensorjoin_variants_amd import _lib
    return torch, _lib, _lib.load()


def probe_rows(n=3072, d=64):
    rng = np.random.default_rng(927)
    x = np.zeros((n, d), np.float32)
    x[4:1024] = rng.uniform(-0.08, 0.08, (1020, d)).astype(np.float32)
    eps = 0.3173828125  # 1300 / 4096: T = eps^2 is exact
    x[128] = 0
    x[128, 0] = eps
    x[129] = x[128]
    x[129, 1:4] = 2.0 ** -28
    x[2048:3072] = x[:1024]
    x[2048] = 0
    return x, eps * eps


def guarded(torch, n, dtype, fill=SENT):
    buf = torch.full((n + 2 * PAD,), fill, dtype=dtype, device="cuda")
    return buf, buf[PAD:PAD + n]


def intact(buf):
    return bool((buf[:PAD] == SENT).all()) and bool((buf[-PAD:] == SENT).all())


@pytest.mark.parametrize("prec", ["fp16", "fp8", "bf16x3"])
@pytest.mark.parametrize("self_join", [True, False])
def test_stage_buffers_stay_in_bounds(env, oracle, prec, self_join):
    torch, lib, L = env
    x, T = probe_rows()
    R, S = (x, None) if self_join else (x[:2000], x[1000:])
    want = oracle.join(R, S, T)
    assert len(want) > 0
    # the pair at exactly eps is in the answer
    if self_join:
        assert ((want[:, 0] == 128) & (want[:, 1] == 2048)).any()
    # size the buffers from an audit-free sizing run
    g = Stages(env, R, S, T, prec, out_cap=1 << 22, cand_cap=1 << 22)
    g.metadata(); g.screen(False); g.certify(False)
    need = g.st()
    assert need["out"] == len(want)
    for frac in (1.0, 0.5):
        out_cap = max(1, int(need["out"] * frac))
        cand_cap = max(1, int(need["candidates"] * frac))
        say("sentinels", prec, "self" if self_join else "rxs", out_cap, cand_cap)
        g = Stages(env, R, S, T, prec, out_cap=out_cap, cand_cap=cand_cap)
        ob, g.out = guarded(torch, out_cap, torch.int64)
        cb, g.cand = guarded(torch, cand_cap, torch.int64)
        sb, g.cand_s = guarded(torch, cand_cap, torch.float32)
        g.metadata()
        g.screen(False)
        st = g.st()
        assert intact(ob) and intact(cb) and intact(sb), ("screen", frac)
        assert st["candidates"] == need["candidates"]
        g.certify(False)
        st = g.st()
        assert intact(ob) and intact(cb) and intact(sb), ("certify", frac)
        if frac == 1.0:
            assert st["out"] == need["out"]
        else:  # a truncated queue: the caller reruns larger (Plan.run does)
            assert st["out"] <= need["out"] and st["candidates"] > cand_cap
        # sort + unpack over the (possibly truncated) output
        torch.cuda.synchronize()
        need_t = C.c_size_t()
        bits = bits_for(max(R.shape[0], 2)) + g.jbits
        lib.check(L.tj_sort_keys(None, C.byref(need_t), None, None, out_cap, bits, None), "sort query")
        temp = torch.empty(need_t.value, dtype=torch.uint8, device="cuda")
        lib.check(L.tj_sort_keys(temp.data_ptr(), C.byref(need_t), g.out.data_ptr(), g.stats.data_ptr(), out_cap, bits,
                                 g.stream), "sort")
        pb, pairs = guarded(torch, 2 * out_cap, torch.int64)
        lib.check(L.tj_unpack_pairs(g.out.data_ptr(), g.stats.data_ptr(), out_cap, g.jbits, pairs.data_ptr(),
                                    lib.OUT["int64"], g.stream), "unpack")
        torch.cuda.synchronize()
        assert intact(ob) and intact(pb), ("sort/unpack", frac)
        if frac == 1.0:
            assert np.array_equal(pairs.view(-1, 2).cpu().numpy(), want)


@pytest.mark.parametrize("method,prec", [("certified", "fp16"), ("plain_fp64", "none"), ("plain_fp32", "none"),
                                         ("grid", "none")])
def test_join_pairs_stay_in_bounds(env, oracle, method, prec):
    """tj_join's pairs buffer, at the exact need and at half of it."""
    torch, lib, L = env
    from tensorjoin_variants_amd.join import Plan, _as_device
    x, T = probe_rows()
    if method == "grid":
        x = x[:, :3].copy()
        T = 0.02 ** 2
    want = oracle.join(x, None, T)
    for frac in (1.0, 0.5):
        cap = max(1, int(len(want) * frac))
        say("join sentinels", method, cap)
        plan = Plan(x.shape[0], None, x.shape[1], method, prec, out_capacity=cap, amb_capacity=1 << 22)
        pb, pairs = guarded(torch, 2 * cap, torch.int64)
        plan.enqueue(_as_device(x), None, T, pairs, "int64")
        torch.cuda.synchronize()
        st = lib.decode_stats(plan._stats.cpu().numpy().view(np.uint64))
        assert intact(pb), (method, frac)
        assert st["out"] == len(want)
        if frac == 1.0:
            assert np.array_equal(pairs.view(-1, 2).cpu().numpy(), want)
        plan.close()
