"""GDS-Join's Python call as a drop-in: ``gdsjoin(DATASET, EPSILON, NDIM, DTYPE)``.

The reference binds GDS-Join through ctypes (GDS/gdsjoingpu.py:146-232):
``gdsjoin(DATASET, EPSILON, NDIM, DTYPE, verbose)`` takes the points as one
flat array of N * NDIM coordinates and returns ``(numNeighbors, neighborTable)``:
numNeighbors[i] (uint32) counts the points within EPSILON of point i, i
itself included, and neighborTable (uint32) lists them, point after point.
Here the same call runs the grid join for NDIM <= 3 (tj_grid_neighbors) and
the certified MFMA join above that (its i < j pairs mirrored, plus the
self pairs, on the device).

Differences, all toward a stricter contract: the predicate is this
library's d2_seq <= fl64(EPSILON^2) (GDS-Join tests sqrt(d2) <= EPSILON in
its own summation order, GDS/kernel.cu:368-371), and every point's
neighbours come in ascending order (GDS-Join leaves them in the order its
atomics wrote them).  Coordinates are FP32 on the device: DTYPE "float"
takes them as FP32, DTYPE "double" accepts only values that FP32 represents
exactly (anything else raises, rather than joining rounded data).  EPSILON
is rounded to FP32 for DTYPE "float", as the reference's c_float(EPSILON).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .join import _as_device, _ptr, _stream_ptr, canonical_u64, join, torch


def _points(DATASET, NDIM: int, DTYPE: str) -> np.ndarray:
    if DTYPE not in ("float", "double"):
        raise ValueError(f"DTYPE must be 'float' or 'double', not {DTYPE!r}")
    flat = np.asarray(DATASET, dtype=np.float64 if DTYPE == "double" else np.float32).reshape(-1)
    if NDIM < 1 or flat.size % NDIM:
        raise ValueError(f"{flat.size} coordinates are not a multiple of NDIM={NDIM}")
    x32 = flat.astype(np.float32)
    if DTYPE == "double" and not np.array_equal(x32.astype(np.float64), flat, equal_nan=True):
        raise ValueError("DTYPE 'double': coordinates must be exactly representable in float32")
    return x32.reshape(-1, NDIM)


def neighbor_table(X, T: float, capacity: int | None = None):
    """Device (counts uint32[n], table uint32[total]) of the grid join's
    neighbour-table output (d <= 3), with capacity retry."""
    t = torch()
    L = _lib.load()
    Xd = _as_device(X)
    n, d = Xd.shape
    cap = int(capacity) if capacity is not None else max(1 << 16, 8 * n)
    stats = t.zeros(_lib.STATS_SLOTS, dtype=t.int64, device="cuda")
    counts = t.empty(max(n, 1), dtype=t.int32, device="cuda")
    for _ in range(4):
        need = C.c_size_t()
        _lib.check(L.tj_grid_workspace_bytes(n, 0, cap, C.byref(need)), "tj_grid_workspace_bytes")
        ws = t.empty(int(need.value), dtype=t.uint8, device="cuda")
        table = t.empty(max(cap, 1), dtype=t.int32, device="cuda")
        _lib.check(L.tj_grid_neighbors(_ptr(Xd), n, d, float(T), _ptr(ws), need.value, cap, _ptr(counts),
                                       _ptr(table), _ptr(stats), _stream_ptr()), "tj_grid_neighbors")
        t.cuda.current_stream().synchronize()
        st = _lib.decode_stats(stats.cpu().numpy().view(np.uint64))
        if st["invalid"]:
            raise ValueError(f"{st['invalid']} non-finite input coordinates")
        if st["out"] <= cap:
            return counts[:n], table[: st["out"]], st
        cap = int(st["out"] * 1.1) + 1024
    raise _lib.TJError("capacity retry limit reached")


def gdsjoin(DATASET, EPSILON, NDIM, DTYPE, verbose: bool = False):
    """(numNeighbors uint32[N], neighborTable uint32[sum]) of the self-join
    within EPSILON, self pairs included, each point's neighbours ascending."""
    X = _points(DATASET, int(NDIM), DTYPE)
    eps = float(np.float32(EPSILON)) if DTYPE == "float" else float(EPSILON)
    T = eps * eps
    n = X.shape[0]
    if n == 0:
        return np.zeros(0, np.uint32), np.zeros(0, np.uint32)
    if X.shape[1] <= 3:
        counts, table, st = neighbor_table(X, T)
        counts = counts.cpu().numpy().view(np.uint32)
        table = table.cpu().numpy().view(np.uint32)
    else:
        t = torch()
        method = "certified" if X.shape[1] <= 512 else "plain_fp64"
        pairs = join(X, None, T=T, method=method, out_dtype=np.int64)
        ids = t.from_numpy(canonical_u64(pairs, n).view(np.int64)).cuda()
        counts = t.bincount(ids // n, minlength=n).to(t.int32).cpu().numpy().view(np.uint32)
        table = (ids % n).to(t.int32).cpu().numpy().view(np.uint32)
    if verbose:
        print(f"[Python] Number of rows in file: {n}")
        print(f"[Python] Total number of neighbors: {int(counts.sum(dtype=np.uint64))}")
    return counts, table
