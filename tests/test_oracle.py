"""Pins the CPU oracle (the checker) before anything is compared against it.

Each check is chosen so a plausible mistake in the oracle fails it:
binary32 instead of binary64 arithmetic (G10 'clearly_outside'), a tree or
reversed summation order (order case), FMA contraction (fma cases), a wrong
comparison (<, not <=: G10 'equality'), a dropped or transposed operand
(restatements and brute force disagree), a broken grid neighbourhood (grid vs
brute force), a wrong output format (canonical ids vs the reference's
construction).
"""
import json
from pathlib import Path

import numpy as np
import pytest

GOLD = json.loads((Path(__file__).parent / "golden" / "kat.json").read_text())["cases"]


@pytest.mark.parametrize("c", GOLD, ids=[c["name"] for c in GOLD])
def test_known_answers(oracle, c):
    x = np.asarray(c["x"], dtype=np.float32)
    y = np.asarray(c["y"], dtype=np.float32)
    d2 = oracle.d2(x, y)
    assert d2.hex() == c["d2_seq_hex"]
    assert oracle.d2_seq_py(x, y).hex() == c["d2_seq_hex"]
    assert oracle.d2_seq_np(x[None], y[None])[0, 0].hex() == c["d2_seq_hex"]
    assert (d2 <= c["T"]) == c["accept"]
    got = oracle.join(x[None], y[None], c["T"])
    assert len(got) == int(c["accept"])


def test_g10_disagreement_set():
    # CEJ/src/check_g10_exactness_contract_cpu.py:59: binary64 and the exact real
    # predicate disagree only on 'just_outside'
    g10 = [c for c in GOLD if c["name"].startswith("g10_")]
    assert [c["name"] for c in g10 if c["accept"] != c["exact_accept"]] == ["g10_just_outside"]


def test_order_and_fma_cases_discriminate():
    order = [c for c in GOLD if c["name"].startswith("order")][0]
    assert order["d2_reversed_hex"] != order["d2_seq_hex"]
    for c in GOLD:
        if c["name"].startswith("fma_"):
            assert c["d2_fma_hex"] != c["d2_seq_hex"]
            assert float.fromhex(c["d2_fma_hex"]) > c["T"] or float.fromhex(c["d2_seq_hex"]) > c["T"]


def test_restatements_agree(oracle):
    rng = np.random.default_rng(11)
    for d in (1, 3, 7, 32, 100):
        R = rng.standard_normal((60, d)).astype(np.float32)
        S = (rng.standard_normal((45, d)) * rng.choice([1e-3, 1.0, 1e3], (45, 1))).astype(np.float32)
        M = oracle.d2_matrix(R, S)
        assert np.array_equal(M, oracle.d2_seq_np(R, S))
        for i, j in [(0, 0), (5, 44), (59, 7)]:
            assert M[i, j] == oracle.d2_seq_py(R[i], S[j])


def test_join_is_brute_force(oracle):
    rng = np.random.default_rng(12)
    R = rng.standard_normal((301, 9)).astype(np.float32)
    S = rng.standard_normal((157, 9)).astype(np.float32)
    M = oracle.d2_seq_np(R, S)
    T = float(np.quantile(M, 0.02))
    got = oracle.join(R, S, T)
    ii, jj = np.nonzero(M <= T)
    assert np.array_equal(got, np.stack([ii, jj], 1))
    got = oracle.join(R, None, T)
    Mr = oracle.d2_seq_np(R, R)
    ii, jj = np.nonzero(np.triu(Mr <= T, 1))
    assert np.array_equal(got, np.stack([ii, jj], 1))
    assert oracle.join(R, S, T, count_only=True) == len(oracle.join(R, S, T))


def test_exact_invariances(oracle):
    """Negation and power-of-two scaling are exact in binary64: same pairs."""
    rng = np.random.default_rng(13)
    X = rng.standard_normal((400, 16)).astype(np.float32)
    T = float(np.quantile(oracle.d2_matrix(X[:100], X[100:]), 0.01))
    base = oracle.join(X, None, T)
    assert np.array_equal(oracle.join(-X, None, T), base)
    for e in (-8, 8):
        assert np.array_equal(oracle.join(X * np.float32(2.0 ** e), None, T * 4.0 ** e), base)
    perm = rng.permutation(400)
    got = oracle.join(X[perm], None, T)
    mapped = np.sort(perm[got], axis=1)  # row r of X[perm] is row perm[r] of X
    mapped = mapped[np.lexsort((mapped[:, 1], mapped[:, 0]))]
    assert np.array_equal(mapped, base)


def test_closed_form_fixtures(oracle):
    # all-zero input: every pair has d2 = 0 (CEJ/src/run_g3c_b_r1_adversarial.py:410-423)
    Z = np.zeros((50, 8), np.float32)
    got = oracle.join(Z, None, 2.0 ** -40)
    assert len(got) == 50 * 49 // 2
    # signed +-1 rows: block-diagonal answer at T = 0 (:424-446)
    X = np.ones((40, 8), np.float32)
    X[20:] = -1.0
    got = oracle.join(X, None, 0.0)
    ii, jj = np.triu_indices(40, 1)
    keep = (ii < 20) == (jj < 20)
    assert np.array_equal(got, np.stack([ii[keep], jj[keep]], 1))


def test_grid_join_matches_brute_force(oracle):
    rng = np.random.default_rng(14)
    for d in (1, 2, 3):
        X = rng.random((3000, d)).astype(np.float32)
        Y = rng.random((2500, d)).astype(np.float32)
        X[:10] = X[10:20]  # duplicates
        T = (0.01 if d > 1 else 0.0005) ** 2
        assert np.array_equal(oracle.grid_join(X, None, T), oracle.join(X, None, T))
        assert np.array_equal(oracle.grid_join(X, Y, T), oracle.join(X, Y, T))
    # points exactly eps apart on the cell boundaries
    g = (np.arange(40, dtype=np.float32) * np.float32(0.25))[:, None]
    X = np.concatenate([g, np.zeros_like(g), np.zeros_like(g)], 1)
    assert np.array_equal(oracle.grid_join(X, None, 0.0625), oracle.join(X, None, 0.0625))


def test_canonical_format(oracle):
    pairs = np.array([[0, 2], [1, 3]], np.int64)
    ids = oracle.canonical_u64(pairs, 4)
    assert ids.tolist() == [0, 2, 5, 7, 8, 10, 13, 15]


def far_outlier_sets():
    """Points a few ulps apart near 1.0 plus outliers 1e10 away: one axis then
    spans ~2^55 cells of side sqrt(T), where binary64 rounding of the cell
    coordinate moves points by several cells unless the side is widened."""
    out = [("3-point d1", np.array([[-1e10], [1 + 7 * 2 ** -23], [1 + 9 * 2 ** -23]], np.float32), 3e-7 ** 2)]
    rng = np.random.default_rng(15)
    for d in (2, 3):
        X = (1.0 + rng.integers(0, 16, (600, d)) * 2.0 ** -23).astype(np.float32)
        X[0, 0] = -1e10
        X[1, d - 1] = 1e10
        out.append((f"ulp cloud d{d}", X, 3e-7 ** 2))
    return out


@pytest.mark.parametrize("case", far_outlier_sets(), ids=lambda c: c[0])
def test_grid_join_far_outlier(oracle, case):
    _, X, T = case
    want = oracle.join(X, None, T)
    assert len(want) > 0
    assert np.array_equal(oracle.grid_join(X, None, T), want)
    Y = X[::-1].copy()
    assert np.array_equal(oracle.grid_join(X, Y, T), oracle.join(X, Y, T))
