"""Host side of the epsilon-join: plans, the join call, the reference format.

The join call keeps the reference's contract -- (R, S, eps) -> the sorted
list of pairs whose sequential binary64 squared distance is <= eps^2 -- with
the certified MFMA path behind it.  PyTorch is used only for device buffers
and the current stream; all computation runs in libtjoin.so.

Reference interfaces mirrored here:
* ``Matrix.run(method, x)`` of TG/src/g2_operator_r1.py:42-61, returning the
  sorted directed uint64 ids i*N+j including the N self pairs and a stats
  dict (the canonical form of TG/src/retained_ops.py:17-19).
* ``run_operator(vectors, threshold, blas)`` of
  G16/src/fp32_gpu_norm_operator.py:107-161 -> (canonical, record).
* The capacity-retry protocol of CEJ/src/run_g2a_tensorjoin.py:329-378: a
  fixed-capacity pass reports the needed count; the call reruns with a
  larger buffer instead of silently dropping pairs.
"""
from __future__ import annotations

import ctypes as C
import json
import math
from dataclasses import dataclass

import numpy as np

from . import _lib

_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as t
        _torch = t
    return _torch


def _stream_ptr(stream=None) -> int:
    t = torch()
    s = stream if stream is not None else t.cuda.current_stream()
    return int(s.cuda_stream)


def _as_device(x, device="cuda"):
    t = torch()
    if isinstance(x, t.Tensor):
        x = x.to(device=device, dtype=t.float32)
    else:
        x = t.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(device)
    if x.dim() != 2:
        raise ValueError("point sets must be 2-D [n, d]")
    return x.contiguous()


def _ptr(t) -> int | None:
    return None if t is None else int(t.data_ptr())


@dataclass
class JoinResult:
    pairs: object  # torch tensor [P, 2] (int32 or int64) on the device
    stats: dict

    def numpy(self) -> np.ndarray:
        return self.pairs.cpu().numpy()


class Plan:
    """Shapes + method + precision + capacities, and the caller-owned workspace.

    One plan serves any number of joins of the same shapes; capacities grow
    automatically when a join reports more pairs than fit.
    """

    def __init__(self, nR: int, nS: int | None, d: int, method: str = "certified",
                 precision: str = "fp16", out_capacity: int = 1 << 20, amb_capacity: int = 1 << 20,
                 audit: bool = False, device: str = "cuda", timing: bool = False,
                 tile: tuple[int, int] | None = None):
        self.nR, self.nS, self.d = int(nR), (None if nS is None else int(nS)), int(d)
        self.method, self.precision = method, precision
        self.audit = audit
        self.timing = timing
        self.device = device
        self.tile = None
        self._plan = None
        self._ws = None
        self._make(int(out_capacity), int(amb_capacity))
        if tile is not None:
            self.set_tile(*tile)

    def set_tile(self, row_blocks: int, cols_per_split: int) -> None:
        """Screen tile of a certified plan: 32-row blocks per wave and columns
        of S per work item (0 = the default of each).  Output-neutral."""
        _lib.check(_lib.load().tj_plan_set_tile(self._plan, int(row_blocks), int(cols_per_split)),
                   "tj_plan_set_tile")
        self.tile = (int(row_blocks), int(cols_per_split))

    def _make(self, out_cap: int, amb_cap: int) -> None:
        L = _lib.load()
        self.close()
        p = C.c_void_p()
        rc = L.tj_plan_create(C.byref(p), self.nR, 0 if self.nS is None else self.nS, self.d,
                              _lib.METHOD[self.method], _lib.PREC[self.precision if self.method == "certified" else "none"],
                              out_cap, amb_cap,
                              (_lib.FLAG_AUDIT if self.audit else 0) | (_lib.FLAG_TIMING if self.timing else 0))
        _lib.check(rc, "tj_plan_create")
        self._plan = p
        if self.tile is not None:  # a regrown plan keeps its tile
            _lib.check(L.tj_plan_set_tile(p, *self.tile), "tj_plan_set_tile")
        self.out_capacity, self.amb_capacity = out_cap, amb_cap
        nbytes = C.c_size_t()
        _lib.check(L.tj_plan_workspace_bytes(p, C.byref(nbytes)), "tj_plan_workspace_bytes")
        t = torch()
        self._ws = t.empty(int(nbytes.value), dtype=t.uint8, device=self.device)
        self._stats = t.zeros(_lib.STATS_SLOTS, dtype=t.int64, device=self.device)
        self.workspace_bytes = int(nbytes.value)

    def describe(self) -> dict:
        buf = C.create_string_buffer(1024)
        _lib.check(_lib.load().tj_plan_describe(self._plan, buf, 1024), "tj_plan_describe")
        return json.loads(buf.value.decode())

    def stage_times(self) -> dict:
        """ms of metadata / screen / refine / sort of the last join (timing plans only)."""
        ms = (C.c_float * 4)()
        _lib.check(_lib.load().tj_plan_stage_times(self._plan, ms, 4), "tj_plan_stage_times")
        return dict(zip(("metadata", "screen", "refine", "sort"), [float(v) for v in ms]))

    def close(self) -> None:
        if self._plan is not None:
            _lib.load().tj_plan_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def enqueue(self, R, S, T: float, pairs, out_format: str, stream=None) -> None:
        """Enqueue one join on the stream (no synchronisation)."""
        L = _lib.load()
        rc = L.tj_join(self._plan, _ptr(R), _ptr(S), float(T), _ptr(self._ws), _ptr(pairs),
                       _lib.OUT[out_format], _ptr(self._stats), _stream_ptr(stream))
        _lib.check(rc, "tj_join")

    def capture(self, R, S, T: float, pairs, out_format: str) -> None:
        """Record one join with these buffers into the plan's hipGraph
        (tj_plan_capture, synchronous); replay() launches it.  A later
        capture replaces the graph; capacities must not change in between."""
        L = _lib.load()
        rc = L.tj_plan_capture(self._plan, _ptr(R), _ptr(S), float(T), _ptr(self._ws), _ptr(pairs),
                               _lib.OUT[out_format], _ptr(self._stats))
        _lib.check(rc, "tj_plan_capture")

    def replay(self, stream=None) -> None:
        """Launch the captured join on the stream (no synchronisation)."""
        _lib.check(_lib.load().tj_plan_replay(self._plan, _stream_ptr(stream)), "tj_plan_replay")

    def run(self, R, S, T: float, out_dtype=np.int64, max_retries: int = 3) -> JoinResult:
        """Join with capacity retry; returns device pairs and the decoded stats."""
        t = torch()
        fmt = "int32" if np.dtype(out_dtype) == np.int32 else "int64"
        tdt = t.int32 if fmt == "int32" else t.int64
        for _ in range(max_retries + 1):
            pairs = t.empty((max(self.out_capacity, 1), 2), dtype=tdt, device=self.device)
            self.enqueue(R, S, T, pairs, fmt)
            t.cuda.current_stream().synchronize()
            stats = _lib.decode_stats(self._stats.cpu().numpy().view(np.uint64))
            if stats["invalid"]:
                raise ValueError(f"{stats['invalid']} non-finite input coordinates")
            # the screen's queue holds every candidate (certified inside + ambiguous band)
            need_out, need_amb = stats["out"], stats["candidates"]
            if need_out <= self.out_capacity and need_amb <= self.amb_capacity:
                stats["plan"] = self.describe()
                return JoinResult(pairs[:need_out], stats)
            grow = lambda need, cap: max(cap, int(need * 1.25) + 1024)
            self._make(grow(need_out, self.out_capacity), grow(need_amb, self.amb_capacity))
        raise _lib.TJError("capacity retry limit reached")


def _initial_capacities(nR: int, nS: int | None, self_join: bool) -> tuple[int, int]:
    pairs = nR * (nR - 1) // 2 if self_join else nR * (nS or 0)
    out = int(min(max(pairs, 1), max(1 << 20, min(pairs // 512, 1 << 26), 16 * max(nR, nS or 0))))
    amb = int(min(max(pairs, 1), 1 << 22))
    return out, amb


def join(R, S=None, eps: float | None = None, *, T: float | None = None, method: str = "auto",
         precision: str = "auto", out_dtype=None, audit: bool = False, return_stats: bool = False):
    """Exact epsilon-join.  Returns sorted (i, j) pairs [P, 2] (numpy).

    A self-join (S is None) returns i < j.  The predicate is
    d2_seq(R_i, S_j) <= T with T = fl64(eps^2) unless T is given.
    method: 'auto' (grid for d <= 3, else certified), 'certified' (MFMA
    screen + FP64 refinement), 'grid' (uniform grid, d <= 3), 'plain_fp64'
    (FP64 brute force), 'plain_fp32' (the FP32-first control: FP32 GEMM-form
    tile over every pair, rigorous interval, FP64 for what it leaves).
    """
    if (eps is None) == (T is None):
        raise ValueError("give exactly one of eps and T")
    if T is None:
        T = float(eps) * float(eps)
    Rn = R.shape[0]
    if Rn == 0 or (S is not None and S.shape[0] == 0):
        empty = np.zeros((0, 2), dtype=out_dtype or np.int32)
        return (empty, {"out": 0, "pairs": 0}) if return_stats else empty
    Rd = _as_device(R)
    Sd = None if S is None else _as_device(S)
    nR, d = Rd.shape
    nS = None if Sd is None else Sd.shape[0]
    if Sd is not None and Sd.shape[1] != d:
        raise ValueError("R and S differ in dimension")
    if method == "auto":
        method = "grid" if d <= 3 else "certified"
    if precision == "auto":
        from .autotune import choose_precision
        precision = choose_precision(d) if method == "certified" else "none"
    if out_dtype is None:
        out_dtype = np.int32 if max(nR, nS or 0) < 2**31 else np.int64
    out_cap, amb_cap = _initial_capacities(nR, nS, S is None)
    plan = Plan(nR, nS, d, method, precision, out_cap, amb_cap, audit)
    res = plan.run(Rd, Sd, T, out_dtype)
    pairs = res.numpy()
    return (pairs, res.stats) if return_stats else pairs


def canonical_u64(pairs, n: int) -> np.ndarray:
    """Upper (i<j) self-join pairs -> reference directed ids, on the GPU.

    Mirror and sort as OA/src/output.cu:8-31 (and CEJ/src/prepare_g2_cifar.py:136-141):
    every i*n+j and j*n+i plus the n diagonal ids, ascending.
    """
    t = torch()
    L = _lib.load()
    p = pairs if isinstance(pairs, t.Tensor) else t.from_numpy(np.ascontiguousarray(pairs, dtype=np.int64))
    p = p.to(device="cuda", dtype=t.int64).contiguous().reshape(-1, 2)
    m = p.shape[0]
    count = t.tensor([m], dtype=t.int64, device="cuda")
    need = C.c_size_t()
    _lib.check(L.tj_canonical_u64(None, C.byref(need), None, None, m, n, None, None), "tj_canonical_u64")
    temp = t.empty(int(need.value), dtype=t.uint8, device="cuda")
    ids = t.empty(max(2 * m + n, 1), dtype=t.int64, device="cuda")
    _lib.check(L.tj_canonical_u64(_ptr(temp), C.byref(need), _ptr(p), _ptr(count), m, n, _ptr(ids),
                                  _stream_ptr()), "tj_canonical_u64")
    t.cuda.current_stream().synchronize()
    return ids[: 2 * m + n].cpu().numpy().view(np.uint64)


def join_reference_compat(x, eps: float | None = None, *, T: float | None = None, precision: str = "auto"):
    """Reference output format: sorted directed uint64 ids i*N+j incl. self pairs."""
    pairs, stats = join(x, None, eps, T=T, precision=precision, out_dtype=np.int64, return_stats=True)
    return canonical_u64(pairs, np.asarray(x).shape[0] if not hasattr(x, "shape") else x.shape[0]), stats


class Matrix:
    """Drop-in for the reference operator ``Matrix`` (TG/src/g2_operator_r1.py:13-61).

    ``Matrix(T).run(method, x) -> (sorted directed uint64 ids incl. self, stats)``.
    Methods: 'F16' / 'S16' -> FP16 MFMA screen; 'F8' / 'S8' -> FP8 (e4m3) MFMA
    screen (the reference's F8 is an INT8 screen; here the 8-bit screen is
    OCP e4m3 with per-row power-of-two scales).  The split (S*) and fused
    (F*) variants of the reference differ only in kernel structure; both map
    to the fused screen here and return the same pairs.  'F32' -> the plain
    FP64 brute force; 'P32' -> the pedantic FP32-first control of
    G16/src/fp32_gpu_norm_operator.py:107-161 (FP32 GEMM-form tile,
    rigorous interval, FP64 terminal).
    """

    METHODS = {"F16": ("certified", "fp16"), "S16": ("certified", "fp16"),
               "F8": ("certified", "fp8"), "S8": ("certified", "fp8"),
               "F16X3": ("certified", "fp16x3"), "F32": ("plain_fp64", "none"),
               "P32": ("plain_fp32", "none")}

    def __init__(self, T: float):
        self.T = float(T)

    def run(self, m: str, x):
        if m not in self.METHODS:
            raise ValueError(f"method {m!r} not in {sorted(self.METHODS)}")
        method, prec = self.METHODS[m]
        n = x.shape[0]
        xd = _as_device(x)
        out_cap, amb_cap = _initial_capacities(n, None, True)
        plan = Plan(n, None, xd.shape[1], method, prec if method == "certified" else "fp16", out_cap, amb_cap)
        t0 = torch().cuda.Event(enable_timing=True)
        t1 = torch().cuda.Event(enable_timing=True)
        t0.record()
        res = plan.run(xd, None, self.T, np.int64)
        t1.record()
        t1.synchronize()
        ids = canonical_u64(res.pairs, n)
        stats = dict(res.stats)
        stats["seconds"] = t0.elapsed_time(t1) / 1e3
        return ids, stats
