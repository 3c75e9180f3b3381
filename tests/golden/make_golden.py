"""Writes tests/golden/kat.json: closed-form known-answer vectors for the
sequential binary64 difference-form predicate.

No reference code is imported or run; every expected value here is fixed by
arithmetic (exact rationals via fractions.Fraction, and Python floats, which
are IEEE binary64 with round-to-nearest-even).

* g10_*   -- the four cases of CEJ/src/check_g10_exactness_contract_cpu.py:56-63
             (D = 512, x = (a, b, 0...), y = 0, T = 1).  The reference asserts
             that binary64 disagrees with the exact real predicate only for
             'just_outside' (:59).
* order_* -- a vector whose sequential sum (k ascending) differs from a
             pairwise/tree sum and from the reversed order: pins the summation
             order of CEJ/src/audit_g2b_accepted_fp64.py:54-57.
* fma_*   -- vectors whose decision flips if fl(s + fl(d*d)) is contracted to
             fma(d, d, s): pins "no FMA".
"""
from __future__ import annotations

import json
import random
import struct
from fractions import Fraction
from pathlib import Path

HERE = Path(__file__).resolve().parent


def f32(v: float) -> float:
    return struct.unpack("<f", struct.pack("<f", v))[0]


def seq(x, y):
    s = 0.0
    for a, b in zip(x, y):
        d = float(a) - float(b)
        s = s + d * d
    return s


def seq_fma(x, y):
    s = 0.0
    for a, b in zip(x, y):
        d = float(a) - float(b)
        s = float(Fraction(d) * Fraction(d) + Fraction(s))
    return s


def pairwise(vals):
    vals = list(vals)
    while len(vals) > 1:
        vals = [vals[i] + vals[i + 1] if i + 1 < len(vals) else vals[i] for i in range(0, len(vals), 2)]
    return vals[0]


def exact(x, y):
    return sum((Fraction(a) - Fraction(b)) ** 2 for a, b in zip(x, y))


def case(name, x, y, T, note=""):
    x = [f32(v) for v in x]
    y = [f32(v) for v in y]
    d2 = seq(x, y)
    ex = exact(x, y)
    return {
        "name": name, "x": x, "y": y, "T": T, "d2_seq_hex": d2.hex(), "accept": d2 <= T,
        "exact_accept": ex <= Fraction(T), "d2_fma_hex": seq_fma(x, y).hex(),
        "d2_pairwise_hex": pairwise([(a - b) * (a - b) for a, b in zip(x, y)]).hex(),
        "d2_reversed_hex": seq(x[::-1], y[::-1]).hex(), "note": note,
    }


def main():
    cases = []
    D = 512
    for name, a, b in (("g10_inside", 0.5, 0.0), ("g10_equality", 1.0, 0.0),
                       ("g10_just_outside", 1.0, 2.0 ** -27), ("g10_clearly_outside", 1.0, 2.0 ** -13)):
        cases.append(case(name, [a, b] + [0.0] * (D - 2), [0.0] * D, 1.0, "CEJ/src/check_g10_exactness_contract_cpu.py:56-63"))
    cases.append(case("order_small_terms_absorbed", [1.0] + [2.0 ** -27] * 4, [0.0] * 5, 1.0,
                      "sequential 1.0 (each 2^-54 absorbed); pairwise and reversed give 1+2^-52"))
    # search fp32 vectors where contraction to fma flips the decision at T = d2_seq
    rng = random.Random(927)
    found = 0
    while found < 4:
        d = rng.choice([2, 3, 4, 8])
        x = [f32(rng.uniform(-2, 2)) for _ in range(d)]
        y = [f32(rng.uniform(-2, 2)) for _ in range(d)]
        s, sf = seq(x, y), seq_fma(x, y)
        if s != sf:
            T = min(s, sf)  # the larger of the two lies outside
            cases.append(case(f"fma_{found}", x, y, T, "decision differs between fl(s+fl(d*d)) and fma(d,d,s)"))
            found += 1
    (HERE / "kat.json").write_text(json.dumps({"cases": cases}, indent=1))
    print(f"wrote {len(cases)} cases")


if __name__ == "__main__":
    main()
