"""Precision / tile choice for the certified screen.

Every choice is certified (the output never depends on it), so the tuner
only trades speed.  ``choose_precision`` is the static default; ``Autotuner``
times candidate precisions on a sample of the actual join with HIP events and
caches the winner per (d, size class) in a JSON file.
"""
from __future__ import annotations

import json
import math
import os
from pathlib import Path

CACHE = Path(os.environ.get("TJ_AUTOTUNE_CACHE", str(Path.home() / ".cache" / "tjoin_autotune.json")))

CANDIDATES = ("fp16", "fp8", "bf16", "fp16x3", "fp8x3")


def choose_precision(d: int) -> str:
    """Static default: FP16 one-term up to d=512 (the widest FP16 instantiation)."""
    if d <= 512:
        return "fp16"
    raise ValueError(f"d={d} exceeds the screen's widest instantiation (512); use method='plain_fp64'")


def size_class(n: int) -> int:
    return int(math.log2(max(n, 1)))


class Autotuner:
    def __init__(self, cache: Path = CACHE):
        self.cache = cache
        try:
            self.table = json.loads(cache.read_text())
        except Exception:
            self.table = {}

    def key(self, d: int, nR: int, nS: int | None) -> str:
        return f"d{d}_r{size_class(nR)}_s{size_class(nS or nR)}"

    def lookup(self, d: int, nR: int, nS: int | None) -> str | None:
        return self.table.get(self.key(d, nR, nS))

    def tune(self, R, S, T: float, candidates=CANDIDATES, sample_rows: int = 16384, reps: int = 3) -> dict:
        """Time each candidate on the first `sample_rows` rows of R; cache the fastest."""
        from .join import Plan, torch
        t = torch()
        Rs = R[:sample_rows]
        nS = None if S is None else S.shape[0]
        times = {}
        for prec in candidates:
            try:
                plan = Plan(Rs.shape[0], nS, R.shape[1], "certified", prec, 1 << 20, 1 << 20)
                plan.run(Rs, S, T)  # warm + size
                best = float("inf")
                for _ in range(reps):
                    a, b = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
                    a.record()
                    plan.run(Rs, S, T)
                    b.record()
                    b.synchronize()
                    best = min(best, a.elapsed_time(b))
                times[prec] = best
            except Exception as e:  # unsupported width for this precision
                times[prec] = float("inf")
        win = min(times, key=times.get)
        self.table[self.key(R.shape[1], R.shape[0], nS)] = win
        try:
            self.cache.parent.mkdir(parents=True, exist_ok=True)
            self.cache.write_text(json.dumps(self.table, indent=1, sort_keys=True))
        except OSError:
            pass
        return {"winner": win, "ms": times}
