"""CPU FP64 oracle for the exact epsilon-join -- TEST INFRASTRUCTURE ONLY.

This module is the checker.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py import it; the product package never does.

Three restatements of one predicate, d2_seq(x, y) <= T, where d2_seq is the
sequential binary64 difference form of CEJ/src/audit_g2b_accepted_fp64.py:51-57
(delta = f64(a) - f64(b); s = fl(s + fl(delta*delta)); k ascending; no FMA):

* ``d2_seq_py``   -- pure Python floats (binary64), one pair; tiny inputs.
* ``d2_seq_np``   -- NumPy, vectorised across pairs, sequential over k.
* ``join`` / ``d2_matrix`` / ``grid_join`` -- the C oracle (tjoracle.c,
  OpenMP over rows, SIMD across pairs only), the one used at scale.

Output order of a join is (i, j) ascending; a self-join emits i < j only
(the upper triangle of CEJ/src/prepare_g2_cifar.py:130-153 without the
diagonal).  ``canonical_u64`` converts to the reference's directed format
(every i*N+j including self pairs, sorted; TG/src/retained_ops.py:17-19).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libtjoracle.so"

_lib = None


def build(force: bool = False) -> Path:
    """Compile libtjoracle.so with the oracle's Makefile (gcc, no FMA contraction)."""
    if force or not LIB_PATH.exists() or LIB_PATH.stat().st_mtime < (HERE / "tjoracle.c").stat().st_mtime:
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=HERE, check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        i64, f64, ptr, cint = C.c_int64, C.c_double, C.c_void_p, C.c_int
        L.tjo_d2.restype = f64
        L.tjo_d2.argtypes = [ptr, ptr, i64]
        L.tjo_join.restype = i64
        L.tjo_join.argtypes = [ptr, i64, ptr, i64, i64, f64, cint, ptr, i64, cint, cint]
        L.tjo_d2_matrix.restype = None
        L.tjo_d2_matrix.argtypes = [ptr, i64, ptr, i64, i64, ptr, cint]
        L.tjo_d2_pairs.restype = None
        L.tjo_d2_pairs.argtypes = [ptr, ptr, i64, ptr, i64, ptr, cint]
        L.tjo_grid_join.restype = i64
        L.tjo_grid_join.argtypes = [ptr, i64, ptr, i64, i64, f64, cint, ptr, i64, cint]
        L.tjo_num_threads.restype = cint
        _lib = L
    return _lib


def _f32(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim != 2:
        raise ValueError("expected a 2-D array")
    return a


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


# ----------------------------------------------------------------- restatements
def d2_seq_py(x, y) -> float:
    """Sequential binary64 difference form on Python floats (one pair)."""
    s = 0.0
    for a, b in zip(x, y):
        delta = float(np.float32(a)) - float(np.float32(b))
        s = s + delta * delta
    return s


def d2_seq_np(R, S) -> np.ndarray:
    """[nR, nS] matrix of d2_seq, vectorised across pairs, sequential over k."""
    R = _f32(R).astype(np.float64)
    S = _f32(S).astype(np.float64)
    out = np.zeros((R.shape[0], S.shape[0]), dtype=np.float64)
    for k in range(R.shape[1]):
        delta = R[:, k][:, None] - S[:, k][None, :]
        out = out + delta * delta
    return out


def d2(x, y) -> float:
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    return float(lib().tjo_d2(_p(x), _p(y), x.size))


def d2_matrix(R, S=None, threads: int = 0) -> np.ndarray:
    R = _f32(R)
    S = R if S is None else _f32(S)
    out = np.empty((R.shape[0], S.shape[0]), dtype=np.float64)
    lib().tjo_d2_matrix(_p(R), R.shape[0], _p(S), S.shape[0], R.shape[1], _p(out), threads)
    return out


def d2_pairs(R, S, pairs, threads: int = 0) -> np.ndarray:
    R = _f32(R)
    S = R if S is None else _f32(S)
    pairs = np.ascontiguousarray(pairs, dtype=np.int64).reshape(-1, 2)
    out = np.empty(pairs.shape[0], dtype=np.float64)
    lib().tjo_d2_pairs(_p(R), _p(S), R.shape[1], _p(pairs), pairs.shape[0], _p(out), threads)
    return out


def join(R, S=None, T: float = 0.0, threads: int = 0, count_only: bool = False):
    """Brute-force exact join.  Returns int64 [P, 2] sorted pairs (or the count)."""
    R = _f32(R)
    self_join = S is None
    S_ = R if self_join else _f32(S)
    if S_.shape[1] != R.shape[1]:
        raise ValueError("dimension mismatch")
    L = lib()
    if count_only:
        return int(L.tjo_join(_p(R), R.shape[0], _p(S_), S_.shape[0], R.shape[1], float(T),
                              int(self_join), None, 0, threads, 1))
    cap = 1 << 16
    while True:
        out = np.empty((cap, 2), dtype=np.int64)
        n = L.tjo_join(_p(R), R.shape[0], _p(S_), S_.shape[0], R.shape[1], float(T),
                       int(self_join), _p(out), cap, threads, 0)
        if n < 0:
            raise MemoryError("oracle join failed")
        if n <= cap:
            return out[:n].copy()
        cap = int(n)


def grid_join(R, S=None, T: float = 0.0, threads: int = 0):
    """Uniform-grid exact join (d <= 3).  Same pair set as ``join``."""
    R = _f32(R)
    self_join = S is None
    S_ = R if self_join else _f32(S)
    L = lib()
    cap = 1 << 20
    while True:
        out = np.empty((cap, 2), dtype=np.int64)
        n = L.tjo_grid_join(_p(R), R.shape[0], _p(S_), S_.shape[0], R.shape[1], float(T),
                            int(self_join), _p(out), cap, threads)
        if n < 0:
            raise RuntimeError(f"oracle grid join failed ({n})")
        if n <= cap:
            return out[:n].copy()
        cap = int(n)


def num_threads() -> int:
    return int(lib().tjo_num_threads())


# ------------------------------------------------------------------ formats
def canonical_u64(pairs: np.ndarray, n: int) -> np.ndarray:
    """Upper (i<j) self-join pairs -> the reference's directed u64 IDs incl. self.

    Follows CEJ/src/prepare_g2_cifar.py:136-141: diagonal i*N+i, i*N+j and
    j*N+i for every i<j pair, then one ascending sort.
    """
    pairs = np.asarray(pairs, dtype=np.uint64).reshape(-1, 2)
    i, j = pairs[:, 0], pairs[:, 1]
    diag = np.arange(n, dtype=np.uint64)
    ids = np.concatenate([diag * np.uint64(n) + diag, i * np.uint64(n) + j, j * np.uint64(n) + i])
    ids.sort()
    return ids
