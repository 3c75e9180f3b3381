"""Threshold selection rules (CPU): the mid-gap rule of
CEJ/src/prepare_g2_cifar.py:93-127 (calibrate_radius), the dyadic grid of
TS/src/select_thresholds.py:21-29 and the outward FP32 bracket of
CEJ/src/run_g3b_r1_gpu_analytic_certificate.py:237-245."""
from fractions import Fraction

import numpy as np
import pytest

from tensorjoin_variants_amd.thresholds import dyadic_eps, midgap_eps, outward_float32, percentile_eps, sample_d2


def test_midgap_separates_ranks(oracle):
    rng = np.random.default_rng(41)
    X = rng.standard_normal((300, 12)).astype(np.float32)
    M = oracle.d2_matrix(X, None)
    upper = M[np.triu_indices(300, 1)]
    u = np.unique(upper)
    for rank in (0, 10, 1000, 20000, u.size - 2):
        eps = midgap_eps(u, rank)
        T = eps * eps
        assert u[rank] < T < u[rank + 1]
        # no pair sits on the boundary: the answer is the pairs up to rank,
        # and it does not move when T moves by an ulp either way
        want = int(np.count_nonzero(upper <= u[rank]))
        for t in (T, np.nextafter(T, 0.0), np.nextafter(T, np.inf)):
            assert oracle.join(X, None, float(t), count_only=True) == want


def test_midgap_rejects_a_gap_below_rounding():
    # sqrt(1 + 2^-52) rounds to 1: no binary64 eps lies strictly inside
    with pytest.raises(ValueError):
        midgap_eps(np.array([1.0, 1.0 + 2.0 ** -52]), 0)


def test_dyadic_grid_is_exact():
    for m in (1, 7, 255, 2576, 4095):
        eps = dyadic_eps(m / 4096 + 1e-9)
        assert eps == m / 4096
        T = eps * eps
        assert Fraction(T) == Fraction(m * m, 2 ** 24)        # no rounding in binary64
        assert float(np.float32(eps)) == eps and float(np.float32(T)) == T  # nor in binary32
    # the reference's 'original' cell (TS/src/select_thresholds.py:26) is the
    # operator's EPS = 161/256, T = 25921/65536 (TG/src/g2_operator_r1.py:9)
    eps = dyadic_eps(0.6289)
    assert eps == 161 / 256 and eps * eps == 25921 / 65536 and round(eps * 4096) == 2576
    assert dyadic_eps(1e-9) == 1 / 4096  # m >= 1


def test_outward_float32_brackets():
    rng = np.random.default_rng(42)
    for v in list(rng.standard_normal(200) * 10.0 ** rng.integers(-30, 30, 200)) + [0.1, 1.0, 2.0 ** -140]:
        lo, hi = outward_float32(float(v))
        assert Fraction(float(lo)) <= Fraction(float(v)) <= Fraction(float(hi))
        assert hi == lo or np.nextafter(lo, np.float32(np.inf)) == hi


def test_percentile_eps_is_a_sampled_distance():
    rng = np.random.default_rng(43)
    X = rng.random((500, 4)).astype(np.float32)
    eps = percentile_eps(X, None, 0.01, 5000, np.random.default_rng(1))
    d2 = np.sort(sample_d2(X, None, 5000, np.random.default_rng(1)))
    assert eps * eps == pytest.approx(d2[int(np.ceil(0.01 * 5000)) - 1], rel=1e-15)
