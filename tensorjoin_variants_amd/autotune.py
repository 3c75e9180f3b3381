"""Precision / tile choice for the certified screen.

Every choice is certified (the output never depends on it), so the tuner
only trades speed.  ``choose_precision`` is the static default; ``Autotuner``
times candidate precisions on a sample of the actual join with HIP events and
keeps the winner per (d, size class) in an in-memory table, and times screen
tiles (row blocks per wave x column split width) on a plan's own join and
keeps the winner per (d, size class, precision).  The table is
written to a JSON file only when one is named (the ``cache`` argument or the
TJ_AUTOTUNE_CACHE environment variable); by default tuning leaves no file.
"""
from __future__ import annotations

import json
import math
import os
from pathlib import Path


def default_cache() -> Path | None:
    """The tuning table's file: TJ_AUTOTUNE_CACHE if set, else none (memory only)."""
    v = os.environ.get("TJ_AUTOTUNE_CACHE")
    return Path(v) if v else None


class _Table:
    """Winner table, loaded from and saved to ``cache`` when one is given."""

    def __init__(self, cache: Path | None):
        self.cache = default_cache() if cache is None else Path(cache)
        self.table = {}
        if self.cache is not None:
            try:
                self.table = json.loads(self.cache.read_text())
            except (OSError, ValueError):
                self.table = {}

    def save(self) -> None:
        if self.cache is None:
            return
        try:
            self.cache.parent.mkdir(parents=True, exist_ok=True)
            self.cache.write_text(json.dumps(self.table, indent=1, sort_keys=True))
        except OSError:
            pass


CANDIDATES = ("fp16", "fp8", "bf16", "fp16x3", "fp8x3")


def choose_precision(d: int) -> str:
    """Static default: FP16 one-term up to d=512 (the widest FP16 instantiation)."""
    if d <= 512:
        return "fp16"
    raise ValueError(f"d={d} exceeds the screen's widest instantiation (512); use method='plain_fp64'")


def size_class(n: int) -> int:
    return int(math.log2(max(n, 1)))


class Autotuner(_Table):
    def __init__(self, cache: Path | None = None):
        super().__init__(cache)

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
        self.save()
        return {"winner": win, "ms": times}

    def tile_key(self, d: int, nR: int, nS: int | None, precision: str) -> str:
        return self.key(d, nR, nS) + f"_{precision}_tile"

    def lookup_tile(self, d: int, nR: int, nS: int | None, precision: str) -> tuple[int, int] | None:
        v = self.table.get(self.tile_key(d, nR, nS, precision))
        return None if v is None else (int(v[0]), int(v[1]))

    @staticmethod
    def tile_candidates(plan) -> list[tuple[int, int]]:
        """Row blocks per wave (the body's default and, for the four-block
        bodies, two) x column split widths (the default rule's, halved,
        doubled, quartered, and all of S in one split)."""
        desc = plan.describe()
        rb0, cps = desc["row_blocks"], desc["cols_per_split"]
        nS = plan.nS or plan.nR
        rbs = [rb0] + ([2] if rb0 == 4 else [])
        widths = sorted({w for w in (cps // 4, cps // 2, cps, 2 * cps, 4 * cps, nS) if 1024 <= w <= nS} | {cps})
        return [(rb, w) for rb in rbs for w in widths]

    def tune_tile(self, plan, R, S, T: float, candidates=None, reps: int = 2, out=None, fmt: str = "int64") -> dict:
        """Time each screen tile on ``plan``'s own join (the full R and S it
        was made for) and keep the fastest per (d, size class, precision).
        Leaves the winner set on the plan."""
        from .join import torch
        t = torch()
        if plan.method != "certified":
            raise ValueError("only a certified plan has a screen tile")
        if candidates is None:
            plan.set_tile(0, 0)
            candidates = self.tile_candidates(plan)
        if out is None:
            out = t.empty((max(plan.out_capacity, 1), 2), dtype=t.int64 if fmt == "int64" else t.int32,
                          device=plan.device)
        times = {}
        for rb, w in candidates:
            plan.set_tile(rb, w)
            plan.enqueue(R, S, T, out, fmt)  # warm
            best = float("inf")
            for _ in range(reps):
                a, b = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
                a.record()
                plan.enqueue(R, S, T, out, fmt)
                b.record()
                b.synchronize()
                best = min(best, a.elapsed_time(b))
            times[(rb, w)] = best
        win = min(times, key=times.get)
        plan.set_tile(*win)
        self.table[self.tile_key(plan.d, plan.nR, plan.nS, plan.precision)] = list(win)
        self.save()
        return {"winner": {"row_blocks": win[0], "cols_per_split": win[1]},
                "ms": [{"row_blocks": rb, "cols_per_split": w, "ms": ms} for (rb, w), ms in times.items()]}


def width_class(d: int) -> int:
    """Padded width of the 16-bit batch screen (16, 32, 64, 128, 256)."""
    for w in (16, 32, 64, 128, 256):
        if d <= w:
            return w
    raise ValueError(f"d={d} exceeds the batch screen's widest instantiation (256)")


class BatchTuner(_Table):
    """Per-width screen precision for batches of joins.

    ``tune`` times each batch precision (FP16, FP8) end to end on a sample of
    the batch's own joins -- the largest few of every width class, where the
    pair work is -- and keeps the faster per width class (kept in the same
    kind of table as ``Autotuner``).  ``choose`` maps a batch to per-join
    precisions from that table (FP16 where a class was never tuned).  The
    choice never changes the output, only the time.
    """

    CANDIDATES = ("fp16", "fp8")

    def __init__(self, cache: Path | None = None):
        super().__init__(cache)

    @staticmethod
    def key(w: int) -> str:
        return f"batch_w{w}"

    def choose(self, ds, ns=None) -> list[str]:
        return [self.table.get(self.key(width_class(d)), "fp16") for d in ds]

    def choose_prune(self) -> bool:
        """The measured choice of tile pruning for batches (off if never tuned)."""
        return bool(self.table.get("batch_prune", False))

    def tune_prune(self, sets, Ts, precisions, joins: int = 24, reps: int = 2) -> dict:
        """Times the batch pipeline on its ``joins`` largest sets with and
        without tile pruning (spatial row order + panel bounding balls) and
        keeps the faster.  Pruning pays where the screen's column sweep
        dominates; where the survivors of the near pairs dominate, the order
        and bound passes cost more than the skipped columns save."""
        import numpy as np
        from .batch import BatchPlan, initial_capacities
        from .join import torch
        t = torch()
        Ts = np.broadcast_to(np.asarray(Ts, dtype=np.float64), (len(sets),))
        idx = sorted(range(len(sets)), key=lambda k: -sets[k].shape[0])[:joins]
        sub = [sets[k] for k in idx]
        ns = [x.shape[0] for x in sub]
        ds = [x.shape[1] for x in sub]
        precs = [precisions[k] for k in idx] if not isinstance(precisions, str) else precisions
        times = {}
        for prune in (False, True):
            oc, cc = initial_capacities(ns)
            plan = BatchPlan(ns, ds, precs, oc, cc, prune=prune)
            plan.run(sub, Ts[idx])
            xs, ptrs, Td = plan.inputs(sub, Ts[idx])
            best = float("inf")
            for _ in range(reps):
                a, b = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
                a.record()
                plan.enqueue(ptrs, Td)
                b.record()
                b.synchronize()
                best = min(best, a.elapsed_time(b))
            times["prune" if prune else "no_prune"] = best
            plan.close()
        win = times["prune"] < times["no_prune"]
        self.table["batch_prune"] = win
        self.save()
        return {"prune": win, "ms": times, "joins": len(idx)}

    def tune(self, sets, Ts, per_class: int = 8, reps: int = 2) -> dict:
        import numpy as np
        from .batch import BatchPlan, initial_capacities
        from .join import torch
        t = torch()
        Ts = np.broadcast_to(np.asarray(Ts, dtype=np.float64), (len(sets),))
        classes: dict[int, list[int]] = {}
        for k, x in enumerate(sets):
            classes.setdefault(width_class(x.shape[1]), []).append(k)
        report = {}
        for w, idx in sorted(classes.items()):
            idx = sorted(idx, key=lambda k: -sets[k].shape[0])[:per_class]
            sub = [sets[k] for k in idx]
            ns = [x.shape[0] for x in sub]
            ds = [x.shape[1] for x in sub]
            times = {}
            for prec in self.CANDIDATES:
                oc, cc = initial_capacities(ns)
                plan = BatchPlan(ns, ds, prec, oc, cc)
                plan.run(sub, Ts[idx])  # warm + size the capacities
                xs, ptrs, Td = plan.inputs(sub, Ts[idx])
                best = float("inf")
                for _ in range(reps):
                    a, b = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
                    a.record()
                    plan.enqueue(ptrs, Td)
                    b.record()
                    b.synchronize()
                    best = min(best, a.elapsed_time(b))
                times[prec] = best
                plan.close()
            win = min(times, key=times.get)
            self.table[self.key(w)] = win
            report[w] = {"winner": win, "ms": times, "joins": len(idx)}
        self.save()
        return report
