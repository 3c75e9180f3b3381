"""Row metadata (tj_scale_scan -> tj_params -> tj_rowmeta) pinned directly,
through the C ABI, against exact rational arithmetic on the CPU.

For every precision, on rows of mixed magnitudes, exact zeros, values on
rounding ties and values below the flush thresholds:

* the join-unit exponent G is ilogb(max |x|) minus the precision's top
  exponent (14 for FP16, 0 otherwise);
* every code equals an independent round-to-nearest-even encoder written
  here from the formats' definitions (significand bits, flush threshold,
  FP8's per-row power-of-two scale and its 448 limit) -- for the hi codes and,
  for the hi/lo splits, the lo codes of the residual x - hi;
* the certificate's row terms enclose the exact values (Fractions):
  nlo <= ||x||^2 <= nhi, H >= ||hi||, L >= ||lo||, E >= ||x - hi - lo||;
* the FP8 scale exponents packed in the metadata are the encoder's.

The enclosure method is that of CEJ/NUMERICS_G9_INT14.md:22-32 (stored
bounds checked against exact rationals); the row pass it pins replaces
half_metadata (TG/src/gpu_kernels.py:33-51) and gpu_quantize_metadata
(G16/src/gpu_preparation.py:12-49).
"""
import ctypes as C
import math
import sys
from fractions import Fraction

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PRECS = ["fp16", "fp16x3", "bf16", "bf16x3", "fp8", "fp8x3"]
TOP = {"fp16": 14, "fp16x3": 14}


def say(*a):
    print("[case]", *a, file=sys.stderr, flush=True)


def rne(v: float, bits: int) -> float:
    """v rounded to `bits` significant bits, ties to even (unbounded exponent)."""
    if v == 0.0:
        return 0.0
    m, e = math.frexp(v)
    return math.ldexp(round(m * 2.0 ** bits), e - bits)  # round(): ties to even


def enc16(v: float, prec: str) -> float:
    if prec.startswith("fp16"):
        return math.copysign(0.0, v) if abs(v) < 2.0 ** -14 else rne(v, 11)
    return 0.0 if abs(v) < 2.0 ** -60 else rne(v, 8)


def row_exp(m: float) -> int:
    """Per-row FP8 exponent: m * 2^-e in (112, 448]."""
    if not m > 0.0:
        return 0
    e = math.frexp(m)[1] - 1 - 8
    if m * 2.0 ** -e > 448.0:
        e += 1
    return max(e, -100)


def enc8(w: float) -> float:
    if abs(w) < 2.0 ** -6:
        return 0.0
    q = rne(w, 4)
    return math.copysign(min(abs(q), 448.0), q)


def dec8(codes: np.ndarray) -> np.ndarray:
    c = codes.astype(np.int64)
    E, M = (c >> 3) & 15, c & 7
    mag = np.where(E == 0, 0.0, (1.0 + M / 8.0) * np.exp2(E - 7.0))
    return np.where(c >> 7, -mag, mag)


def data(prec: str, d: int) -> np.ndarray:
    """Rows of mixed magnitudes plus rows of rounding ties (built in join
    units for the G this data set gets)."""
    rng = np.random.default_rng(900 + d)
    X = (rng.standard_normal((60, d)) * np.exp2(rng.integers(-20, 21, (60, 1)))).astype(np.float32)
    X[3] = 0.0
    X[5, ::3] = 0.0
    X[7] = (rng.standard_normal(d) * 2.0 ** -60).astype(np.float32)
    top = TOP.get(prec, 0)
    big = float(np.max(np.abs(X)))
    G = math.frexp(big)[1] - 1 - top
    if prec.startswith("fp8"):
        # a row with max 1.5 in join units (per-row scale exponent -8, so the
        # e4m3 operand is w = 256 v): ties of 4-bit significands (272, 304,
        # -336), values at and below the 2^-6 flush, an ordinary rounding
        t = np.array([384.0, 272.0, 304.0, -336.0, 2 ** -6 * (1 + 2 ** -4), 2 ** -7, 3.0, 100.5]) / 256.0
    elif prec.startswith("bf16"):
        t = np.array([1.5, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -(1 + 5 * 2 ** -8), 2 ** -60 * 1.5, 2 ** -61, 0.75])
    else:
        t = np.array([1.5, 1 + 2 ** -11, 1 + 3 * 2 ** -11, -(1 + 5 * 2 ** -11), 2 ** -14 * (1 + 2 ** -11),
                      2 ** -15, -(2 ** -16), 0.75])
    ties = np.zeros((2, d))
    ties[0, : min(d, t.size)] = t[:d]
    ties[1, : min(d, t.size)] = -t[:d] * 2.0 ** -3
    Xt = (ties * 2.0 ** G).astype(np.float32)
    assert np.array_equal(Xt.astype(np.float64), ties * 2.0 ** G)  # exact in binary32
    return np.concatenate([X, Xt]).astype(np.float32)


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available(), "no GPU visible"
    from tensorjoin_variants_amd import _lib
    return torch, _lib, _lib.load()


@pytest.mark.parametrize("d", [16, 37, 100])
@pytest.mark.parametrize("prec", PRECS)
def test_rowmeta_codes_and_enclosures(env, prec, d):
    torch, lib, L = env
    say("rowmeta", prec, d)
    X = data(prec, d)
    n = X.shape[0]
    Xd = torch.from_numpy(X).cuda()
    s = torch.cuda.current_stream().cuda_stream
    stats = torch.zeros(lib.STATS_SLOTS, dtype=torch.int64, device="cuda")
    maxabs = torch.zeros(1, dtype=torch.int64, device="cuda")
    params = torch.zeros(64, dtype=torch.int64, device="cuda")
    P = lib.PREC[prec]
    lib.check(L.tj_scale_scan(Xd.data_ptr(), n, d, maxabs.data_ptr(), stats.data_ptr(), s), "scale")
    lib.check(L.tj_params(maxabs.data_ptr(), 1.0, d, P, params.data_ptr(), stats.data_ptr(), s), "params")
    qb, mb = C.c_size_t(), C.c_size_t()
    lib.check(L.tj_rowmeta_sizes(n, d, P, C.byref(qb), C.byref(mb)), "sizes")
    qh = torch.zeros(qb.value, dtype=torch.uint8, device="cuda")
    ql = torch.zeros(qb.value, dtype=torch.uint8, device="cuda")
    meta = torch.zeros(mb.value // 4, dtype=torch.float32, device="cuda")
    lib.check(L.tj_rowmeta(Xd.data_ptr(), n, d, P, params.data_ptr(), qh.data_ptr(), ql.data_ptr(), meta.data_ptr(),
                           stats.data_ptr(), s), "rowmeta")
    torch.cuda.synchronize()
    G = int(params.cpu().numpy().view(np.int32)[18])  # after nine binary64 fields
    top = TOP.get(prec, 0)
    assert G == math.frexp(float(np.max(np.abs(X))))[1] - 1 - top
    fp8 = prec.startswith("fp8")
    x3 = prec.endswith("x3")
    dp = qb.value // (n * (1 if fp8 else 2))
    M = meta.cpu().numpy().reshape(n, 12)
    H = qh.cpu().numpy().reshape(n, -1)
    Lo = ql.cpu().numpy().reshape(n, -1)
    if fp8:
        hi_codes, lo_codes = H, Lo
    else:
        hi16, lo16 = H.view(np.uint16), Lo.view(np.uint16)
    packed = M[:, 7].view(np.uint32)
    for r in range(n):
        v = [float(X[r, k]) * 2.0 ** -G for k in range(d)] + [0.0] * (dp - d)
        if fp8:
            e_hi = row_exp(max(abs(a) for a in v))
            hi = [enc8(a * 2.0 ** -e_hi) * 2.0 ** e_hi for a in v]
            got_hi = dec8(hi_codes[r]) * 2.0 ** e_hi
            res = [a - b for a, b in zip(v, hi)]
            e_lo = row_exp(max(abs(a) for a in res)) if x3 else 0
            lo = [enc8(a * 2.0 ** -e_lo) * 2.0 ** e_lo for a in res] if x3 else [0.0] * dp
            got_lo = dec8(lo_codes[r]) * 2.0 ** e_lo if x3 else np.zeros(dp)
            assert packed[r] == ((e_hi + 32768) | ((e_lo + 32768) << 16)), (prec, r)
        else:
            hi = [enc16(a, prec) for a in v]
            if prec.startswith("fp16"):
                got_hi = hi16[r].view(np.float16).astype(np.float64)
                want_codes = np.array(hi, dtype=np.float16).view(np.uint16)
                assert np.array_equal(hi16[r], want_codes), (prec, r)  # signed zeros included
            else:
                got_hi = (hi16[r].astype(np.uint32) << 16).view(np.float32).astype(np.float64)
            res = [a - b for a, b in zip(v, hi)]
            lo = [enc16(a, prec) for a in res] if x3 else [0.0] * dp
            if x3:
                got_lo = (lo16[r].view(np.float16).astype(np.float64) if prec.startswith("fp16")
                          else (lo16[r].astype(np.uint32) << 16).view(np.float32).astype(np.float64))
            else:
                got_lo = np.zeros(dp)
        assert np.array_equal(got_hi, np.array(hi)), (prec, d, r)
        assert np.array_equal(got_lo, np.array(lo)), (prec, d, r)
        # enclosures, exactly
        nlo, nhi, Hb, Eb, Lb = (Fraction(float(M[r, c])) for c in (0, 1, 2, 3, 4))
        fv = [Fraction(a) for a in v]
        fh = [Fraction(a) for a in hi]
        fl = [Fraction(a) for a in lo]
        nx = sum(a * a for a in fv)
        assert nlo <= nx <= nhi, (prec, r)
        assert Hb >= 0 and Hb * Hb >= sum(a * a for a in fh), (prec, r)
        assert Eb >= 0 and Eb * Eb >= sum((a - b - c) ** 2 for a, b, c in zip(fv, fh, fl)), (prec, r)
        if x3:
            assert Lb * Lb >= sum(a * a for a in fl), (prec, r)
        # and they are tight (a loose bound would still be sound, but would
        # widen every pair's band): within 2^-20 relative, where binary32 can
        # hold the bracket that finely (far above its subnormal range)
        if nx > Fraction(2) ** -100:
            assert nhi - nlo <= nx * Fraction(1, 2 ** 20), (prec, r)
