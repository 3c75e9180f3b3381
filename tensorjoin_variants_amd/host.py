"""Host-to-host joins: pageable NumPy inputs -> sorted NumPy result.

``join_host`` / ``HostPlan`` run the library's host pipeline (tj_host_*,
csrc/host.hip): staged uploads through pinned buffers, the device join chunk
by chunk over R, and the result copied back while the next chunk computes.
This is the path the reference's public runs time end to end
(CEJ/src/run_g5_tensorjoin_public.py:372-547: pageable source -> device ->
stages -> sorted canonical ids on the host).  The result array is the
library's own allocation, handed to NumPy without a copy and released when
the array is garbage-collected.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class _Owned:
    """Keeps a tj_host_free'd block alive for the NumPy arrays viewing it."""

    def __init__(self, ptr: int):
        self.ptr = ptr

    def __del__(self):
        if self.ptr:
            _lib.load().tj_host_free(C.c_void_p(self.ptr))
            self.ptr = 0


def _wrap(ptr: int, count: int, dtype, cols: int) -> np.ndarray:
    if count == 0 or not ptr:
        if ptr:
            _lib.load().tj_host_free(C.c_void_p(ptr))
        return np.zeros((0, cols) if cols > 1 else (0,), dtype=dtype)
    ctype = {np.dtype(np.int32): C.c_int32, np.dtype(np.int64): C.c_int64,
             np.dtype(np.uint64): C.c_uint64}[np.dtype(dtype)]
    buf = (ctype * (count * cols)).from_address(ptr)
    buf._owner = _Owned(ptr)
    a = np.ctypeslib.as_array(buf)
    return a.reshape(count, cols) if cols > 1 else a


def _host_f32(x, name: str) -> np.ndarray:
    a = np.asarray(x)
    if a.ndim != 2:
        raise ValueError(f"{name} must be a 2-D array")
    return np.ascontiguousarray(a, dtype=np.float32)


class HostPlan:
    """Reusable host-to-host join of fixed shapes (tj_host_create).

    ``run(R, S, T, out)`` -> (result, stats): out = 'int32' / 'int64' gives
    sorted (i, j) rows; 'canonical' (self-joins) gives the reference's
    directed uint64 ids i*n+j with both directions and the n self pairs."""

    def __init__(self, nR: int, nS: int | None, d: int, method: str = "certified", precision: str = "fp16",
                 audit: bool = False, chunk_rows: int = 0, threads: int = 0):
        L = _lib.load()
        self.nR, self.nS, self.d = int(nR), (None if nS is None else int(nS)), int(d)
        h = C.c_void_p()
        _lib.check(L.tj_host_create(C.byref(h), self.nR, 0 if nS is None else self.nS, self.d,
                                    _lib.METHOD[method], _lib.PREC[precision],
                                    _lib.FLAG_AUDIT if audit else 0, int(chunk_rows), int(threads)),
                   "tj_host_create")
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.load().tj_host_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def times(self) -> dict:
        ms = (C.c_double * 4)()
        _lib.check(_lib.load().tj_host_times(self._h, ms, 4), "tj_host_times")
        return {"total": ms[0], "stage": ms[1], "wait_device": ms[2], "copy_out": ms[3]}

    def run(self, R, S, T: float, out: str = "int32"):
        L = _lib.load()
        Rh = _host_f32(R, "R")
        Sh = None if S is None else _host_f32(S, "S")
        if Rh.shape != (self.nR, self.d) or (Sh is not None and Sh.shape != (self.nS, self.d)):
            raise ValueError("input shapes differ from the plan's")
        if (Sh is None) != (self.nS is None):
            raise ValueError("S must be given exactly for R x S plans")
        fmt = _lib.OUT_U64_CANONICAL if out == "canonical" else _lib.OUT[out]
        res = C.c_void_p()
        count = C.c_int64()
        raw = np.zeros(_lib.STATS_SLOTS, dtype=np.uint64)
        _lib.check(L.tj_host_join(self._h, Rh.ctypes.data, None if Sh is None else Sh.ctypes.data, float(T), fmt,
                                  C.byref(res), C.byref(count), raw.ctypes.data), "tj_host_join")
        n = int(count.value)
        if out == "canonical":
            result = _wrap(res.value or 0, n, np.uint64, 1)
        else:
            result = _wrap(res.value or 0, n, np.int32 if out == "int32" else np.int64, 2)
        stats = _lib.decode_stats(raw)
        stats["host_ms"] = self.times()
        return result, stats


def join_host(R, S=None, eps: float | None = None, *, T: float | None = None, method: str = "auto",
              precision: str = "auto", out: str | None = None, audit: bool = False, return_stats: bool = False,
              chunk_rows: int = 0):
    """Host-to-host join of NumPy arrays: sorted int32/int64 (i, j) rows
    (self-joins: i < j), or with out='canonical' (self-joins) the reference's
    sorted directed uint64 ids incl. self pairs.  eps/T, method and precision
    as in join()."""
    if (eps is None) == (T is None):
        raise ValueError("give exactly one of eps and T")
    if T is None:
        T = float(eps) * float(eps)
    Rh = _host_f32(R, "R")
    Sh = None if S is None else _host_f32(S, "S")
    if Sh is not None and Sh.shape[1] != Rh.shape[1]:
        raise ValueError("R and S widths differ")
    d = Rh.shape[1]
    if method == "auto":
        method = "grid" if d <= 3 else "certified"
    if precision == "auto":
        from .autotune import choose_precision
        precision = choose_precision(d) if method == "certified" else "none"
    if out is None:
        out = "int32" if max(Rh.shape[0], 0 if Sh is None else Sh.shape[0]) < 2 ** 31 else "int64"
    if Sh is not None and Sh.shape[0] == 0:  # nS = 0 would name a self-join in the C API
        empty = np.zeros((0, 2), dtype=np.int64 if out == "int64" else np.int32)
        return (empty, {"out": 0, "pairs": 0}) if return_stats else empty
    plan = HostPlan(Rh.shape[0], None if Sh is None else Sh.shape[0], Rh.shape[1], method, precision, audit,
                    chunk_rows)
    try:
        res, st = plan.run(Rh, Sh, T, out)
    finally:
        plan.close()
    return (res, st) if return_stats else res
