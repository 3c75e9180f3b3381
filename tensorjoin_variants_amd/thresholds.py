"""Threshold selection for epsilon-joins.

* ``percentile_eps`` -- eps = q-quantile of sampled pair distances (the rank
  rule of CEJ/src/prepare_g2_cifar.py:93-127 applied to a random sample).
* ``midgap_eps``     -- eps halfway between the k-th and (k+1)-th distinct
  squared distances, so no pair sits on the boundary
  (CEJ/src/prepare_g2_cifar.py:93-127, ``calibrate_radius``).
* ``dyadic_eps``     -- snap eps to m / 4096 so eps and T = m^2 / 2^24 are exact
  in binary32 and binary64 (TS/src/select_thresholds.py:21-29).
* ``outward_float32`` -- the FP32 bracket of a binary64 value
  (CEJ/src/run_g3b_r1_gpu_analytic_certificate.py:237-245).
"""
from __future__ import annotations

import numpy as np


def sample_d2(R: np.ndarray, S: np.ndarray | None, samples: int, rng: np.random.Generator,
              chunk: int = 1 << 16) -> np.ndarray:
    """Squared distances (binary64, difference form) of uniformly sampled pairs."""
    nR = R.shape[0]
    if S is None:
        if nR < 2:
            return np.zeros(0)
        i = rng.integers(0, nR, samples)
        j = rng.integers(0, nR - 1, samples)
        j = j + (j >= i)
        other = R
    else:
        i = rng.integers(0, nR, samples)
        j = rng.integers(0, S.shape[0], samples)
        other = S
    out = np.empty(samples, dtype=np.float64)
    for a in range(0, samples, chunk):
        b = min(samples, a + chunk)
        delta = R[i[a:b]].astype(np.float64) - other[j[a:b]].astype(np.float64)
        out[a:b] = np.einsum("ij,ij->i", delta, delta)
    return out


def percentile_eps(R, S, quantile: float, samples: int, rng) -> float:
    """eps such that a fraction ``quantile`` of sampled pairs lie within eps."""
    d2 = sample_d2(R, S, samples, rng)
    if d2.size == 0:
        return 0.0
    k = max(0, min(d2.size - 1, int(np.ceil(quantile * d2.size)) - 1))
    return float(np.sqrt(np.partition(d2, k)[k]))


def midgap_eps(d2_sorted_unique: np.ndarray, rank: int) -> float:
    """eps halfway between distinct squared distances number ``rank`` and ``rank+1``."""
    lower = float(np.sqrt(d2_sorted_unique[rank]))
    upper = float(np.sqrt(d2_sorted_unique[rank + 1]))
    eps = lower + (upper - lower) / 2.0
    if not d2_sorted_unique[rank] < eps * eps < d2_sorted_unique[rank + 1]:
        raise ValueError("rounded eps escaped the gap")
    return eps


def dyadic_eps(eps: float, denominator: int = 4096) -> float:
    """Nearest m / denominator (m >= 1); with denominator 4096, T = m^2 / 2^24 is exact."""
    m = max(1, int(round(eps * denominator)))
    return m / denominator


def outward_float32(value: float) -> tuple[np.float32, np.float32]:
    r = np.float32(value)
    lo = r if float(r) <= value else np.nextafter(r, np.float32(-np.inf))
    hi = r if float(r) >= value else np.nextafter(r, np.float32(np.inf))
    return np.float32(lo), np.float32(hi)
