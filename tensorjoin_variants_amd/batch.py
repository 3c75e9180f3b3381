"""Batches of independent epsilon self-joins in one device pipeline.

``batched_join(sets, eps)`` joins every point set with its own threshold and
returns one sorted (i, j), i < j, array per set -- the same result as calling
``join(X, None, eps=e)`` per set.  All joins share one row-metadata pass, one
persistent screen launch per precision family over the panels of every join
(largest first), one certificate pass and one sort, with no host round trip
between stages; ``BatchPlan.capture`` records the sequence into a hipGraph.

The reference has no batched join API: it runs one join at a time and slices
the triangular tile list of that join into host batches of 4,096 tiles with a
blocking count read after every stage (CEJ/src/run_g5_tensorjoin_public.py:
377-381,410-535; TG/src/g2_operator_r1.py:45-57).
"""
from __future__ import annotations

import ctypes as C
import json

import numpy as np

from . import _lib
from .join import _as_device, _ptr, _stream_ptr, torch

BATCH_PRECISIONS = ("fp16", "fp8")


class BatchPlan:
    """Sizes, widths and per-join precisions of a batch, and its workspace."""

    def __init__(self, ns, ds, precisions, out_capacity: int = 1 << 20, cand_capacity: int = 1 << 20,
                 audit: bool = False, timing: bool = False, device: str = "cuda", prune: bool = False):
        self.ns = [int(n) for n in ns]
        self.ds = [int(d) for d in ds]
        if isinstance(precisions, str):
            precisions = [precisions] * len(self.ns)
        self.precisions = list(precisions)
        if not (len(self.ns) == len(self.ds) == len(self.precisions)) or not self.ns:
            raise ValueError("ns, ds and precisions must be non-empty and of equal length")
        for p in self.precisions:
            if p not in BATCH_PRECISIONS:
                raise ValueError(f"batch precision {p!r} not in {BATCH_PRECISIONS}")
        self.audit, self.timing, self.device, self.prune = audit, timing, device, prune
        self.pairs_total = sum(n * (n - 1) // 2 for n in self.ns)
        self._b = None
        self._make(int(out_capacity), int(cand_capacity))

    def _make(self, out_cap: int, cand_cap: int) -> None:
        L = _lib.load()
        self.close()
        J = len(self.ns)
        n = (C.c_int64 * J)(*self.ns)
        d = (C.c_int * J)(*self.ds)
        pr = (C.c_int * J)(*[_lib.PREC[p] for p in self.precisions])
        b = C.c_void_p()
        flags = (_lib.FLAG_AUDIT if self.audit else 0) | (_lib.FLAG_TIMING if self.timing else 0) | \
            (_lib.FLAG_PRUNE if self.prune else 0)
        _lib.check(L.tj_batch_create(C.byref(b), J, n, d, pr, out_cap, cand_cap, flags), "tj_batch_create")
        self._b = b
        self.out_capacity, self.cand_capacity = out_cap, cand_cap
        nbytes = C.c_size_t()
        _lib.check(L.tj_batch_workspace_bytes(b, C.byref(nbytes)), "tj_batch_workspace_bytes")
        t = torch()
        self._ws = t.empty(int(nbytes.value), dtype=t.uint8, device=self.device)
        self._stats = t.zeros(_lib.STATS_SLOTS, dtype=t.int64, device=self.device)
        self._offsets = t.zeros(J + 1, dtype=t.int64, device=self.device)
        self.pairs = t.empty((max(out_cap, 1), 2), dtype=t.int32, device=self.device)
        self.workspace_bytes = int(nbytes.value)
        self._captured = None

    def close(self) -> None:
        if self._b is not None:
            _lib.load().tj_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def describe(self) -> dict:
        buf = C.create_string_buffer(1024)
        _lib.check(_lib.load().tj_batch_describe(self._b, buf, 1024), "tj_batch_describe")
        return json.loads(buf.value.decode())

    def stage_times(self) -> dict:
        ms = (C.c_float * 4)()
        _lib.check(_lib.load().tj_batch_stage_times(self._b, ms, 4), "tj_batch_stage_times")
        return dict(zip(("metadata", "screen", "refine", "sort"), [float(v) for v in ms]))

    def inputs(self, sets, Ts):
        """Device copies of the point sets, and the pointer / threshold arrays."""
        t = torch()
        xs = [_as_device(np.zeros((0, d), np.float32) if x is None else x) for x, d in zip(sets, self.ds)]
        for k, (x, n, d) in enumerate(zip(xs, self.ns, self.ds)):
            if tuple(x.shape) != (n, d):
                raise ValueError(f"set {k}: shape {tuple(x.shape)} != plan ({n}, {d})")
        Ts = np.asarray(Ts, dtype=np.float64)
        if Ts.shape != (len(self.ns),) or not np.all(np.isfinite(Ts)) or np.any(Ts < 0):
            raise ValueError("thresholds must be finite, >= 0, one per set")
        ptrs = t.tensor([int(x.data_ptr()) if x.numel() else 0 for x in xs], dtype=t.int64, device=self.device)
        Td = t.from_numpy(Ts.copy()).to(self.device)
        return xs, ptrs, Td

    def enqueue(self, ptrs, Td, stream=None) -> None:
        rc = _lib.load().tj_batch_join(self._b, _ptr(ptrs), _ptr(Td), _ptr(self._ws), _ptr(self.pairs),
                                       _ptr(self._offsets), _ptr(self._stats), _stream_ptr(stream))
        _lib.check(rc, "tj_batch_join")

    def capture(self, ptrs, Td) -> None:
        """Record the whole pipeline for these inputs into a graph (replay())."""
        rc = _lib.load().tj_batch_capture(self._b, _ptr(ptrs), _ptr(Td), _ptr(self._ws), _ptr(self.pairs),
                                          _ptr(self._offsets), _ptr(self._stats))
        _lib.check(rc, "tj_batch_capture")
        self._captured = (ptrs, Td)

    def replay(self, stream=None) -> None:
        _lib.check(_lib.load().tj_batch_replay(self._b, _stream_ptr(stream)), "tj_batch_replay")

    def read_stats(self) -> dict:
        return _lib.decode_stats(self._stats.cpu().numpy().view(np.uint64))

    def results(self) -> tuple[list[np.ndarray], dict]:
        """Per-join pair arrays of the last completed run, and its stats."""
        torch().cuda.current_stream().synchronize()
        st = self.read_stats()
        n = st["out"]
        off = self._offsets.cpu().numpy()
        allp = self.pairs[:n].cpu().numpy()
        return [allp[off[k]:off[k + 1]] for k in range(len(self.ns))], st

    def run(self, sets, Ts, max_retries: int = 3):
        """Join every set with capacity retry.  Returns (list of pair arrays, stats)."""
        t = torch()
        xs, ptrs, Td = self.inputs(sets, Ts)
        for _ in range(max_retries + 1):
            self.enqueue(ptrs, Td)
            t.cuda.current_stream().synchronize()
            st = self.read_stats()
            if st["invalid"]:
                raise ValueError(f"{st['invalid']} non-finite input coordinates or thresholds")
            need_out, need_c = st["out"], st["candidates"]
            if need_out <= self.out_capacity and need_c <= self.cand_capacity:
                res, st = self.results()
                st["plan"] = self.describe()
                return res, st
            grow = lambda need, cap: max(cap, int(need * 1.25) + 1024)
            self._make(grow(need_out, self.out_capacity), grow(need_c, self.cand_capacity))
        raise _lib.TJError("capacity retry limit reached")


def initial_capacities(ns) -> tuple[int, int]:
    pairs = sum(n * (n - 1) // 2 for n in ns)
    rows = sum(ns)
    cap = int(min(max(pairs, 1), max(1 << 20, 16 * rows, pairs // 256)))
    cap = min(cap, (1 << 32) - 1)
    return cap, cap


def batched_join(sets, eps=None, *, T=None, precision="auto", audit: bool = False, return_stats: bool = False,
                 tuner=None, prune=False):
    """Exact epsilon self-joins of many point sets at once.

    sets: list of [n_t, d_t] float32 arrays (d_t <= 256, sizes and widths may
    differ).  eps / T: one value for all sets or one per set (T = fl64(eps^2)
    unless T is given).  precision: 'auto' (per-width choice of the batch
    tuner), 'fp16', 'fp8', or a list with one entry per set.  prune: order
    each set's rows spatially and skip panel pairs whose bounding balls are
    certified apart -- True, False or 'auto' (the batch tuner's measured
    choice); the result never depends on it.  Returns a list of
    sorted int32 (i, j), i < j, arrays -- identical to join(X, None, eps=e)
    per set.
    """
    J = len(sets)
    if J == 0:
        return ([], {}) if return_stats else []
    if (eps is None) == (T is None):
        raise ValueError("give exactly one of eps and T")
    if T is None:
        e = np.broadcast_to(np.asarray(eps, dtype=np.float64), (J,))
        Ts = e * e
    else:
        Ts = np.broadcast_to(np.asarray(T, dtype=np.float64), (J,)).copy()
    ns = [int(x.shape[0]) for x in sets]
    ds = [int(x.shape[1]) for x in sets]
    if precision == "auto" or prune == "auto":
        from .autotune import BatchTuner
        tuner = tuner or BatchTuner()
        if precision == "auto":
            precision = tuner.choose(ds, ns)
        if prune == "auto":
            prune = tuner.choose_prune()
    out_cap, cand_cap = initial_capacities(ns)
    plan = BatchPlan(ns, ds, precision, out_cap, cand_cap, audit, prune=prune)
    res, st = plan.run(sets, Ts)
    return (res, st) if return_stats else res
