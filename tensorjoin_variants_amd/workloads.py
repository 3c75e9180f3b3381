"""Synthetic inputs for the five epsilon-join workloads, with fixed seeds.

Every generator is deterministic (NumPy PCG64, seed recorded in the returned
``Workload.meta``).  Thresholds come from ``thresholds.percentile_eps`` on
sampled pair distances, the selection rule of CEJ/src/prepare_g2_cifar.py:93-127
(a rank in sampled squared distances) reduced to a plain percentile as the
workload definitions ask.

Sizes (``scale=1.0``):
  clusters_selfjoin  N=8192, d=32, 16 Gaussian clusters, eps = 0.25th pct     (int32 i<j)
  grid_rxs           |R|=|S|=4,194,304, d=3, 90% uniform + 10% in 64 shared blobs, eps=0.002
  certified_d128     |R|=1,048,576, |S|=262,144, d=128, half-normal 40% zeros, L2-normalised,
                     eps at selectivity 1e-4
  heavytail_fp8      |R|=|S|=262,144, d=512, Student-t nu=3, eps at selectivity 1e-5
  batched_uneven     4096 self-joins, n log-uniform [64,16384], d in {16,64,256}
  gist_shape_selfjoin  N=60,000, d=512 self-join (the reference's public run's shape), ~64 neighbours/pt
``scale`` shrinks row counts for tests (selectivities are kept).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .thresholds import percentile_eps

SEED_BASE = 20261203


@dataclass
class Workload:
    name: str
    R: np.ndarray
    S: np.ndarray | None  # None = self-join (i < j)
    eps: float
    T: float  # fl64(eps^2): the squared threshold the predicate uses
    out_dtype: type = np.int64
    meta: dict = field(default_factory=dict)

    @property
    def pairs_evaluated(self) -> int:
        n = self.R.shape[0]
        return n * (n - 1) // 2 if self.S is None else n * self.S.shape[0]


def _rng(k: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(SEED_BASE + k))


def clusters_selfjoin(n: int = 8192, d: int = 32, clusters: int = 16, pct: float = 0.25,
                      samples: int = 100_000, seed_offset: int = 1) -> Workload:
    """16 Gaussian clusters, centres ~ N(0, 5^2 I), points = centre + N(0, I)."""
    rng = _rng(seed_offset)
    centres = rng.standard_normal((clusters, d)) * 5.0
    assign = rng.integers(0, clusters, n)
    X = (centres[assign] + rng.standard_normal((n, d))).astype(np.float32)
    eps = percentile_eps(X, None, pct / 100.0, samples, rng)
    return Workload("clusters_selfjoin", X, None, eps, eps * eps, np.int32,
                    dict(n=n, d=d, clusters=clusters, percentile=pct, samples=samples,
                         seed=SEED_BASE + seed_offset, assignment="uniform random"))


def _points_with_blobs(rng, n, d, uniform_frac, centres, sigma):
    blobs = centres.shape[0]
    nu = int(round(n * uniform_frac))
    nb = n - nu
    X = np.empty((n, d), dtype=np.float32)
    X[:nu] = rng.random((nu, d), dtype=np.float32)
    per = np.full(blobs, nb // blobs)
    per[: nb % blobs] += 1
    idx = np.repeat(np.arange(blobs), per)
    X[nu:] = (centres[idx] + sigma * rng.standard_normal((nb, d))).astype(np.float32)
    perm = rng.permutation(n)
    return X[perm]


def grid_rxs(n: int = 4_194_304, d: int = 3, eps: float = 0.002, seed_offset: int = 2) -> Workload:
    """90% uniform in [0,1]^3, 10% in 64 Gaussian blobs (sigma 0.003).  The
    64 blob centres (uniform in [0,1]^3) are drawn once and shared by R and
    S, so R's blob points meet S's: about 170 S neighbours per R blob point,
    ~7e7 output pairs at full size (the output-bound half of the workload);
    the points themselves are drawn independently."""
    rng = _rng(seed_offset)
    centres = rng.random((64, d))
    R = _points_with_blobs(rng, n, d, 0.9, centres, 0.003)
    S = _points_with_blobs(rng, n, d, 0.9, centres, 0.003)
    return Workload("grid_rxs", R, S, eps, eps * eps, np.int64,
                    dict(n=n, d=d, uniform_frac=0.9, blobs=64, sigma=0.003, shared_centres=True,
                         seed=SEED_BASE + seed_offset))


def _half_normal_sparse(rng, n, d, zero_frac, chunk=1 << 16):
    X = np.empty((n, d), dtype=np.float32)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        v = np.abs(rng.standard_normal((b - a, d), dtype=np.float32))
        v[rng.random((b - a, d), dtype=np.float32) < zero_frac] = 0.0
        norm = np.sqrt(np.sum(v.astype(np.float64) ** 2, axis=1))
        norm[norm == 0] = 1.0
        X[a:b] = (v / norm[:, None]).astype(np.float32)
    return X


def certified_d128(nR: int = 1_048_576, nS: int = 262_144, d: int = 128, selectivity: float = 1e-4,
                   samples: int = 1_000_000, seed_offset: int = 3) -> Workload:
    """Half-normal |N(0,1)| coordinates, 40% exact zeros, rows L2-normalised."""
    rng = _rng(seed_offset)
    R = _half_normal_sparse(rng, nR, d, 0.4)
    S = _half_normal_sparse(rng, nS, d, 0.4)
    eps = percentile_eps(R, S, selectivity, samples, rng)
    return Workload("certified_d128", R, S, eps, eps * eps, np.int32,
                    dict(nR=nR, nS=nS, d=d, zero_frac=0.4, selectivity=selectivity,
                         samples=samples, seed=SEED_BASE + seed_offset))


def _student_t(rng, n, d, nu, chunk=1 << 15):
    X = np.empty((n, d), dtype=np.float32)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        z = rng.standard_normal((b - a, d), dtype=np.float32)
        g = rng.standard_gamma(nu / 2.0, size=(b - a, d), dtype=np.float32)
        X[a:b] = z * np.sqrt(np.float32(nu / 2.0) / np.maximum(g, np.float32(1e-30)))
    return X


def heavytail_fp8(nR: int = 262_144, nS: int = 262_144, d: int = 512, selectivity: float = 1e-5,
                  samples: int = 1_000_000, seed_offset: int = 4) -> Workload:
    """Student-t (nu = 3) per coordinate: Z * sqrt(nu / chi2_nu)."""
    rng = _rng(seed_offset)
    R = _student_t(rng, nR, d, 3.0)
    S = _student_t(rng, nS, d, 3.0)
    eps = percentile_eps(R, S, selectivity, samples, rng)
    return Workload("heavytail_fp8", R, S, eps, eps * eps, np.int32,
                    dict(nR=nR, nS=nS, d=d, nu=3, selectivity=selectivity, samples=samples,
                         seed=SEED_BASE + seed_offset))


def _mixture_join(seed, n: int, d: int, pct: float, samples: int):
    rng = np.random.default_rng(seed)
    centres = (rng.standard_normal((4, d)) * 3.0).astype(np.float32)
    X = rng.standard_normal((n, d), dtype=np.float32)
    X += centres[rng.integers(0, 4, n)]
    eps = percentile_eps(X, None, pct / 100.0, min(samples, n * (n - 1) // 2), rng)
    return X, eps


def batched_uneven(joins: int = 4096, nmin: int = 64, nmax: int = 16384, dims=(16, 64, 256),
                   pct: float = 0.5, samples: int = 20_000, seed_offset: int = 5, threads: int = 0):
    """Independent self-joins of log-uniform size, 4-cluster mixtures (centres
    ~ N(0, 3^2 I), points = centre + N(0, I)), eps per join at the ``pct``-th
    percentile of sampled pair distances.  Join t draws from its own child
    seed, so the sets do not depend on how many threads generate them."""
    from concurrent.futures import ThreadPoolExecutor
    import os
    rng = _rng(seed_offset)
    sizes = np.exp(rng.uniform(np.log(nmin), np.log(nmax + 1), joins)).astype(np.int64)
    sizes = np.clip(sizes, nmin, nmax)
    ds = rng.choice(np.asarray(dims), joins)
    seeds = np.random.SeedSequence(SEED_BASE + seed_offset).spawn(joins)
    threads = threads or min(16, len(os.sched_getaffinity(0)))
    with ThreadPoolExecutor(max_workers=threads) as ex:
        made = list(ex.map(lambda t: _mixture_join(seeds[t], int(sizes[t]), int(ds[t]), pct, samples),
                           range(joins)))
    return [Workload(f"batched_uneven[{t}]", X, None, eps, eps * eps, np.int32,
                     dict(n=int(sizes[t]), d=int(ds[t]), percentile=pct))
            for t, (X, eps) in enumerate(made)]


def gist_shape_selfjoin(n: int = 60_000, d: int = 512, selectivity: float = 1.074e-3,
                        samples: int = 1_000_000, seed_offset: int = 6) -> Workload:
    """The shape of the reference's public CIFAR-GIST run (60,000 x 512
    self-join, about 64 neighbours per point: 1,933,039 upper pairs of
    1.8e9 at its published eps, CEJ/DECISION_G2B_FORMAL.md:17-21) on
    synthetic rows: non-negative half-normal, 40% zeros, L2-normalised.
    The CIFAR-GIST file itself is not in the reference archive."""
    rng = _rng(seed_offset)
    X = _half_normal_sparse(rng, n, d, 0.4)
    eps = percentile_eps(X, None, selectivity, samples, rng)
    return Workload("gist_shape_selfjoin", X, None, eps, eps * eps, np.int32,
                    dict(n=n, d=d, zero_frac=0.4, selectivity=selectivity, samples=samples,
                         seed=SEED_BASE + seed_offset))


REGISTRY = {
    "clusters_selfjoin": clusters_selfjoin,
    "gist_shape_selfjoin": gist_shape_selfjoin,
    "grid_rxs": grid_rxs,
    "certified_d128": certified_d128,
    "heavytail_fp8": heavytail_fp8,
}
