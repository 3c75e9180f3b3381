#!/usr/bin/env python3
"""Benchmark: exact epsilon-join pair evaluations per second on one MI355X.

Default workload (the metric's configuration): the d = 128 certified-screen
join -- |R| = 1,048,576, |S| = 262,144, d = 128, half-normal coordinates with
40% exact zeros, rows L2-normalised, eps at selectivity 1e-4 -- inputs
resident in HBM.  A step is one full join: row metadata, MFMA screen,
per-pair certificate + FP64 refinement of the ambiguous band, radix sort,
unpack to int32 (i, j) rows.

Measured in the same run:
* the plain GPU methods on the same workload -- the FP64 brute force (no
  certificate; the metric's denominator) and the FP32-first control (FP32
  GEMM-form tile, rigorous interval, FP64 for what it leaves) -- and the
  pair-set difference of each against the certified result (must be 0);
* the screen kernel against the MFMA peak of its precision (roofline);
* the CPU oracle (C, OpenMP, sequential FP64) on a bounded row sample;
* the host-to-host pipeline (pageable NumPy inputs -> sorted pageable
  result, PCIe and host copies included) on the same join -- reported
  beside the headline value, never as it.

Other workloads (--workload): clusters_selfjoin (8192 x 32 self-join),
grid_rxs (4,194,304^2, d = 3, uniform grid), heavytail_fp8 (262,144^2,
d = 512, Student-t, FP8 screen), batched_uneven (4,096 self-joins of
log-uniform size in [64, 16384], d in {16, 64, 256}, one pipeline replayed
as a hipGraph), gist_shape_selfjoin (60,000 x 512 self-join: the shape of
the reference's public host-to-host run).  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "exact ε-join pair evals/s: MFMA-certified vs plain GPU, 0 pair diffs vs FP64"
PEAK = {"fp16": 2.5e15, "bf16": 2.5e15, "fp16x3": 2.5e15, "bf16x3": 2.5e15, "fp8": 5.0e15, "fp8x3": 5.0e15}
HBM = 8.0e12


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def measured_traffic(workload: str):
    """HBM bytes per launch of the dominant kernel, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs (scaled by a copy of known size);
    None when this workload has no such measurement."""
    try:
        t = json.loads((ROOT / "profiles" / "traffic.json").read_text())[workload]
    except (OSError, KeyError, ValueError):
        return None, None
    return float(t["traffic_bytes"]), f"{t['profile']} (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE on {t['kernel']})"


def cpu_baseline(w, seconds_target: float = 12.0) -> dict:
    """Oracle (sequential FP64, OpenMP) on the first k rows of R against all of S."""
    from oracle import oracle
    oracle.build()
    threads = len(os.sched_getaffinity(0))
    S = w.S
    nS = w.R.shape[0] if S is None else S.shape[0]
    if w.name == "grid_rxs":
        t0 = time.perf_counter()
        got = oracle.grid_join(w.R, w.S, w.T, threads=threads)
        dt = time.perf_counter() - t0
        pairs = w.R.shape[0] * nS
        return {"value": pairs / dt, "unit": "pair-evals/s (logical)", "cores": threads, "kind": "port",
                "sample": f"full grid join on the CPU (uniform grid, sequential FP64 predicate), {len(got)} pairs, "
                          f"{dt:.2f} s"}
    k = 16
    t0 = time.perf_counter()
    oracle.join(w.R[:k], S, w.T, threads=threads, count_only=True)
    dt = time.perf_counter() - t0
    k2 = int(min(w.R.shape[0], max(k, k * seconds_target / max(dt, 1e-3))))
    # a workload smaller than the target is joined whole, repeatedly
    reps = 1 if k2 < w.R.shape[0] else max(1, int(min(50, 2.0 / max(dt * k2 / k, 1e-4))))
    t0 = time.perf_counter()
    for _ in range(reps):
        cnt = oracle.join(w.R[:k2], S if S is not None else w.R, w.T, threads=threads, count_only=True)
    dt = (time.perf_counter() - t0) / reps
    pairs = k2 * nS
    return {"value": pairs / dt, "unit": "pair-evals/s", "cores": threads, "kind": "port",
            "sample": f"first {k2} rows of R x all {nS} rows of S ({pairs:.3e} pairs, {cnt} inside), "
                      f"{dt:.3f} s per pass, {reps} pass(es)"}


def make_workload(name: str, scale: float):
    from tensorjoin_variants_amd import workloads
    if name == "certified_d128":
        return workloads.certified_d128(nR=int(1_048_576 * scale), nS=int(262_144 * scale))
    if name == "heavytail_fp8":
        return workloads.heavytail_fp8(nR=int(262_144 * scale), nS=int(262_144 * scale))
    if name == "clusters_selfjoin":
        return workloads.clusters_selfjoin(n=int(8192 * scale))
    if name == "grid_rxs":
        return workloads.grid_rxs(n=int(4_194_304 * scale))
    if name == "gist_shape_selfjoin":
        return workloads.gist_shape_selfjoin(n=int(60_000 * scale))
    raise SystemExit(f"unknown workload {name}")


def timed(plan, R, S, T, out, fmt, steps, warmup, graph=False):
    """ms per join over `steps` joins enqueued back to back (hipGraph replays
    of the whole join when `graph`, else eager launches)."""
    import torch
    if graph:
        plan.capture(R, S, T, out, fmt)
        step = plan.replay
    else:
        step = lambda: plan.enqueue(R, S, T, out, fmt)
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="certified_d128",
                    choices=["certified_d128", "heavytail_fp8", "clusters_selfjoin", "grid_rxs", "batched_uneven",
                             "gist_shape_selfjoin"])
    ap.add_argument("--precision", default=None)
    ap.add_argument("--scale", type=float, default=1.0, help="row-count scale (1.0 = full workload)")
    ap.add_argument("--variants", default=None,
                    help="other screen precisions timed on the same join (comma list; '' for none; "
                         "default: the 16-bit hi/lo splits for certified_d128)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-tune-tile", action="store_true", help="run the screen's default tile (no tile tuning)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-host", action="store_true", help="skip the host-to-host pipeline leg")
    ap.add_argument("--prune", dest="prune", action="store_true", default=None,
                    help="batched workload: force tile pruning on (default: measured choice)")
    ap.add_argument("--no-prune", dest="prune", action="store_false", help="batched workload: force tile pruning off")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-sized device buffer 3 times (scales the HBM byte counters of a "
                         "rocprofv3 --pmc run against a kernel of known traffic)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration << 20
        a = torch.empty(n, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n, "copies": 3}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}): this build runs the join on a "
              "single GPU; sharding over several GPUs is a design, not a run", file=sys.stderr)
        return 2

    if args.workload == "batched_uneven":
        return bench_batched(args)

    import torch
    from tensorjoin_variants_amd import _lib
    from tensorjoin_variants_amd.join import Plan, _as_device

    grid = args.workload == "grid_rxs"
    prec = args.precision or ("fp8" if args.workload == "heavytail_fp8" else "fp16")
    method = "grid" if grid else "certified"
    log(f"generating {args.workload} (scale {args.scale})")
    t0 = time.perf_counter()
    w = make_workload(args.workload, args.scale)
    log(f"generated in {time.perf_counter() - t0:.1f}s, eps={w.eps:.6g} T={w.T!r}")
    R = _as_device(w.R)
    S = None if w.S is None else _as_device(w.S)
    nR, d = w.R.shape
    nS = None if w.S is None else w.S.shape[0]
    pairs_total = w.pairs_evaluated
    fmt = "int32" if w.out_dtype == np.int32 else "int64"
    tdt = torch.int32 if fmt == "int32" else torch.int64

    guess = int(pairs_total * 2e-4) + 4096 if not grid else 16 * nR
    plan = Plan(nR, nS, d, method, prec if not grid else "none", out_capacity=guess,
                amb_capacity=int(pairs_total * 2e-4) + 4096, timing=True)
    log("plan", json.dumps(plan.describe()))
    res = plan.run(R, S, w.T, w.out_dtype)  # sizes the capacities
    log("first run stats", {k: v for k, v in res.stats.items() if k != "plan"})
    out = torch.empty((plan.out_capacity, 2), dtype=tdt, device="cuda")

    tile = None
    if not grid and not args.no_tune_tile:
        # screen tile: each candidate (row blocks per wave x column split
        # width) timed on this join; the headline runs the fastest
        from tensorjoin_variants_amd.autotune import Autotuner
        d0 = plan.describe()
        rep = Autotuner(cache=None).tune_tile(plan, R, S, w.T, out=out, fmt=fmt)
        tile = {"default": {"row_blocks": d0["row_blocks"], "cols_per_split": d0["cols_per_split"]},
                "tuned": rep["winner"], "candidates_ms": rep["ms"]}
        log("tile", json.dumps(tile["tuned"]), "of", len(rep["ms"]), "candidates")

    # ---- timed region: K joins back to back, no host synchronisation inside
    # (the whole join replayed as one hipGraph; eager launches timed beside it)
    ms_per_step = timed(plan, R, S, w.T, out, fmt, args.steps, args.warmup, graph=True)
    eager_ms = timed(plan, R, S, w.T, out, fmt, args.steps, 1)
    stats = _lib.decode_stats(plan._stats.cpu().numpy().view(np.uint64))
    assert stats["out"] <= plan.out_capacity and stats["candidates"] <= plan.amb_capacity
    assert stats["violations"] == 0
    value = pairs_total / (ms_per_step / 1e3)

    # ---- stage breakdown (separate instrumented loop)
    stages = []
    for _ in range(max(3, min(args.steps, 5))):
        plan.enqueue(R, S, w.T, out, fmt)
        stages.append(plan.stage_times())
    st_ms = {k: float(np.mean([s[k] for s in stages])) for k in stages[0]}
    log("stages ms", {k: round(v, 4) for k, v in st_ms.items()})
    if grid:
        # dominant work: the grid probe; algorithmic bytes = inputs read once + pairs written once
        alg_bytes = (nR + (nS or 0)) * d * 4 + stats["out"] * 2 * (4 if fmt == "int32" else 8)
        achieved = alg_bytes / (st_ms["screen"] / 1e3)
        roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM / 1e9, "unit": "GB/s",
                    "frac": achieved / HBM, "traffic": None, "kernel": "grid join (index build + probe + write)",
                    "algorithmic_bytes_per_launch": alg_bytes}
    else:
        flops = 2.0 * d * pairs_total
        achieved = flops / (st_ms["screen"] / 1e3)
        peak = PEAK[prec]
        roofline = {"bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                    "frac": achieved / peak, "traffic": None, "kernel": "screen_kernel",
                    "algorithmic_flops_per_launch": flops}
    roofline["traffic"], roofline["traffic_source"] = measured_traffic(args.workload)

    result = {
        "metric": METRIC, "value": value, "unit": "pair-evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": prec if not grid else "f64", "data": "synthetic",
        "config": {"workload": args.workload, "nR": nR, "nS": nS if nS is not None else nR, "d": d,
                   "self_join": w.S is None, "eps": w.eps, "T": w.T, "output": f"sorted {fmt} (i, j)",
                   "method": method, "precision": prec if not grid else None, "scale": args.scale},
        "launch": "hipGraph replay", "eager_ms_per_step": eager_ms,
        "stages_ms": st_ms, "roofline": roofline,
        "pairs_out": stats["out"], "ambiguous": stats["ambiguous"], "candidates": stats["candidates"],
        "ambiguous_fraction": stats["ambiguous"] / pairs_total,
        "fp32_decided": stats["fp32_accept"] + stats["fp32_reject"], "fp64_refined": stats["fp64_refined"],
        "max_abs_d2_err_vs_fp64": stats["max_err"], "max_err_over_halfwidth": stats["max_ratio"],
        "certificate_violations": stats["violations"],
    }
    if grid:
        result["grid_candidate_evals"] = stats["grid_candidates"]
    if tile is not None:
        result["screen_tile"] = tile

    if args.variants is None:
        args.variants = "fp16x3,bf16x3" if args.workload == "certified_d128" else ""
    if not grid and args.variants:
        # the other screen precisions on the same join: the hi/lo splits
        # issue three MFMA terms per product for a band ~2^11 narrower
        result["precision_variants"] = {}
        for vp in [p for p in args.variants.split(",") if p and p != prec]:
            log(f"precision variant {vp}")
            vplan = Plan(nR, nS, d, "certified", vp, out_capacity=plan.out_capacity,
                         amb_capacity=max(4096, plan.amb_capacity), timing=True)
            vres = vplan.run(R, S, w.T, w.out_dtype)
            vout = torch.empty((vplan.out_capacity, 2), dtype=tdt, device="cuda")
            vms = timed(vplan, R, S, w.T, vout, fmt, max(2, min(args.steps, 3)), 1)
            vst = _lib.decode_stats(vplan._stats.cpu().numpy().view(np.uint64))
            vstages = vplan.stage_times()
            n1 = stats["out"]
            same = vst["out"] == n1 and bool(torch.equal(vout[:n1], out[:n1]))
            vflops = 2.0 * d * pairs_total
            result["precision_variants"][vp] = {
                "ms_per_step": vms, "value": pairs_total / (vms / 1e3), "stages_ms": vstages,
                "screen_mfma_frac_algorithmic": vflops / (vstages["screen"] / 1e3) / PEAK[vp],
                "screen_mfma_frac_issued": 3 * vflops / (vstages["screen"] / 1e3) / PEAK[vp]
                if vp.endswith("x3") else None,
                "candidates": vst["candidates"], "ambiguous": vst["ambiguous"],
                "ambiguous_fraction": vst["ambiguous"] / pairs_total, "fp64_refined": vst["fp64_refined"],
                "max_abs_d2_err_vs_fp64": vst["max_err"], "certificate_violations": vst["violations"],
                "identical_to_headline": same}
            log(f"  {vp}: {vms:.1f} ms, ambiguous {vst['ambiguous']}, identical={same}")
            del vout, vplan

    if not args.no_plain and not grid:
        # the plain GPU methods on the same join, one run each: the FP64 brute
        # force (the metric's denominator) and the FP32-first control (FP32
        # GEMM-form tile with a rigorous interval, FP64 for what it leaves)
        for key, pm in (("plain_gpu", "plain_fp64"), ("plain_gpu_fp32", "plain_fp32")):
            log(f"{pm} GPU method (one run)")
            pplan = Plan(nR, nS, d, pm, "none", out_capacity=plan.out_capacity, amb_capacity=0, timing=True)
            pout = torch.empty((pplan.out_capacity, 2), dtype=tdt, device="cuda")
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            pplan.enqueue(R, S, w.T, pout, fmt)
            torch.cuda.synchronize()
            pwall = time.perf_counter() - t0
            pst = _lib.decode_stats(pplan._stats.cpu().numpy().view(np.uint64))
            n1, n2 = stats["out"], pst["out"]
            same = n1 == n2 and bool(torch.equal(out[:n1], pout[:n2]))
            diffs = 0
            if not same:
                a = out[:n1].cpu().numpy().astype(np.int64)
                b = pout[:n2].cpu().numpy().astype(np.int64)
                nn = nS or nR
                ka, kb = a[:, 0] * nn + a[:, 1], b[:, 0] * nn + b[:, 1]
                diffs = int(np.setdiff1d(ka, kb).size + np.setdiff1d(kb, ka).size)
            rec = {"method": pm, "value": pairs_total / pwall, "ms": pwall * 1e3, "pairs_out": n2,
                   "stages_ms": pplan.stage_times(), "pair_diffs_vs_certified": diffs}
            if pm == "plain_fp32":
                rec.update(certified_by_interval=pst["cert_accept"], fp64_refined=pst["fp64_refined"],
                           fp64_fraction=pst["fp64_refined"] / pairs_total)
            result[key] = rec
            result["speedup_vs_" + key] = value / (pairs_total / pwall)
            if pm == "plain_fp64":
                result["speedup_vs_plain"] = value / (pairs_total / pwall)
                result["pair_diffs_vs_plain_fp64"] = diffs
            log(f"{pm} {pwall*1e3:.1f} ms, diffs {diffs}")
            del pout, pplan

    if not args.no_host:
        # the same join from pageable host arrays to a sorted host array
        # (PCIe and host copies included); never the headline value
        from tensorjoin_variants_amd.host import HostPlan
        log("host-to-host pipeline")
        hp = HostPlan(nR, nS, d, method, prec if not grid else "none")
        hres, hst = hp.run(w.R, w.S, w.T, fmt)  # sizes the capacities
        walls, parts = [], []
        for _ in range(3):
            del hres
            t0 = time.perf_counter()
            hres, hst = hp.run(w.R, w.S, w.T, fmt)
            walls.append(time.perf_counter() - t0)
            parts.append(hst["host_ms"])
        hp.close()
        k = int(np.argsort(walls)[len(walls) // 2])
        n1 = stats["out"]
        same = hres.shape[0] == n1 and bool(np.array_equal(hres, out[:n1].cpu().numpy()))
        result["host_to_host"] = {
            "value": pairs_total / walls[k], "unit": "pair-evals/s", "ms": walls[k] * 1e3,
            "parts_ms": parts[k], "inputs": "pageable NumPy float32", "output": f"sorted {fmt} (i, j), pageable",
            "input_bytes": int(w.R.nbytes + (0 if w.S is None else w.S.nbytes)),
            "output_bytes": int(hres.nbytes), "identical_to_device_result": same}
        log(f"host-to-host {walls[k]*1e3:.1f} ms (device-resident {ms_per_step:.1f} ms), identical={same}")
        del hres
        if w.S is None:
            # the reference's output format: sorted directed u64 ids incl. self pairs
            from tensorjoin_variants_amd.vecio import validate_canonical
            hp = HostPlan(nR, None, d, method, prec if not grid else "none")
            ids, _ = hp.run(w.R, None, w.T, "canonical")
            cw = []
            for _ in range(3):
                del ids
                t0 = time.perf_counter()
                ids, cst = hp.run(w.R, None, w.T, "canonical")
                cw.append(time.perf_counter() - t0)
            hp.close()
            rep = validate_canonical(ids, nR, expected_count=2 * n1 + nR)
            result["host_to_host_canonical"] = {
                "ms": float(np.median(cw)) * 1e3, "ids": int(ids.size), "sha256": rep["canonical_raw_u64_sha256"],
                "contract_checks_pass": rep["exact_contract_pass"],
                "output": "sorted directed uint64 i*n+j incl. self pairs, pageable"}
            log(f"host-to-host canonical {np.median(cw)*1e3:.1f} ms, {ids.size} ids, checks {rep['exact_contract_pass']}")
            del ids

    if not args.no_cpu:
        log("cpu baseline")
        result["cpu_baseline"] = cpu_baseline(w)
    print(json.dumps(result), flush=True)
    return 0


def bench_batched(args) -> int:
    """4,096 independent self-joins in one pipeline (hipGraph replay)."""
    import torch
    from tensorjoin_variants_amd import workloads
    from tensorjoin_variants_amd.autotune import BatchTuner, width_class
    from tensorjoin_variants_amd.batch import BatchPlan, initial_capacities
    from tensorjoin_variants_amd.join import Plan, _as_device

    joins = max(1, int(4096 * args.scale))
    log(f"generating batched_uneven ({joins} joins)")
    t0 = time.perf_counter()
    ws = workloads.batched_uneven(joins=joins)
    sets = [w.R for w in ws]
    Ts = np.array([w.T for w in ws])
    ns = [x.shape[0] for x in sets]
    ds = [x.shape[1] for x in sets]
    pairs_total = sum(n * (n - 1) // 2 for n in ns)
    flops = float(sum(2.0 * d * (n * (n - 1) // 2) for n, d in zip(ns, ds)))
    log(f"generated in {time.perf_counter() - t0:.1f}s: rows {sum(ns)}, pairs {pairs_total:.4e}")

    # per-width precision from the batch tuner (timed on the largest joins of each width)
    if args.precision:
        precs = [args.precision] * joins
        tune = None
    else:
        tuner = BatchTuner()
        tune = tuner.tune(sets, Ts)
        precs = tuner.choose(ds)
        log("tuner", json.dumps(tune))
    # tile pruning on or off: measured (--prune / --no-prune force it)
    if args.prune is None:
        ptune = BatchTuner().tune_prune(sets, Ts, precs)
        prune = ptune["prune"]
        log("prune tuner", json.dumps(ptune))
    else:
        prune, ptune = args.prune, None
    oc, cc = initial_capacities(ns)
    plan = BatchPlan(ns, ds, precs, oc, cc, timing=True, prune=prune)
    res, st = plan.run(sets, Ts)  # sizes the capacities
    log("first run", {k: v for k, v in st.items() if k != "plan"})
    xs, ptrs, Td = plan.inputs(sets, Ts)
    plan.capture(ptrs, Td)

    def run_steps(k, graph):
        for _ in range(k):
            plan.replay() if graph else plan.enqueue(ptrs, Td)

    run_steps(args.warmup, True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run_steps(args.steps, True)
    torch.cuda.synchronize()
    ms_per_step = (time.perf_counter() - t0) * 1e3 / args.steps
    got, stats = plan.results()
    assert stats["out"] <= plan.out_capacity and stats["candidates"] <= plan.cand_capacity
    assert stats["violations"] == 0
    value = pairs_total / (ms_per_step / 1e3)

    t0 = time.perf_counter()
    run_steps(args.steps, False)
    torch.cuda.synchronize()
    eager_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    stages = []
    for _ in range(3):
        plan.enqueue(ptrs, Td)
        stages.append(plan.stage_times())
    st_ms = {k: float(np.mean([s[k] for s in stages])) for k in stages[0]}
    log("stages ms", st_ms)
    # the screen's algorithmic work: 2 d per pair it sweeps (pairs of panel
    # pairs certified apart by the tile bounds are never swept)
    flops_screened = flops - 2.0 * stats["pruned_work"]
    achieved = flops_screened / (st_ms["screen"] / 1e3)
    roofline = {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK["fp16"] / 1e12, "unit": "TFLOP/s",
                "frac": achieved / PEAK["fp16"], "traffic": None,
                "kernel": "batch_screen_kernel (one persistent launch per precision family)",
                "algorithmic_flops_per_launch": flops_screened}
    roofline["traffic"], roofline["traffic_source"] = measured_traffic("batched_uneven")
    result = {
        "metric": METRIC, "value": value, "unit": "pair-evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp16" if all(p == "fp16" for p in precs) else "fp16+fp8",
        "data": "synthetic",
        "config": {"workload": "batched_uneven", "joins": joins, "rows": sum(ns), "pairs": pairs_total,
                   "d": sorted(set(ds)), "n_range": [min(ns), max(ns)], "percentile": 0.5,
                   "output": "per join sorted int32 (i, j), i < j, plus offsets", "launch": "hipGraph replay",
                   "precision_by_width": {str(width_class(d)): p for d, p in zip(ds, precs)}},
        "stages_ms": st_ms, "eager_ms_per_step": eager_ms, "roofline": roofline, "tuner": tune,
        "tile_pruning": {"on": prune, "tuner": ptune},
        "plan": plan.describe(),
        "pairs_out": stats["out"], "ambiguous": stats["ambiguous"], "candidates": stats["candidates"],
        "ambiguous_fraction": stats["ambiguous"] / pairs_total,
        "fp32_decided": stats["fp32_accept"] + stats["fp32_reject"], "fp64_refined": stats["fp64_refined"],
        "pruned_pairs": stats["pruned"], "screened_pairs": pairs_total - stats["pruned"],
        "max_abs_d2_err_vs_fp64": stats["max_err"], "max_err_over_halfwidth": stats["max_ratio"],
        "certificate_violations": stats["violations"],
    }

    if not args.no_plain:
        # the plain FP64 GPU method, one join at a time, on the same device inputs
        log("plain FP64 GPU method (per join)")
        diffs = 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plain_ms = 0.0
        for k, (x, n, d) in enumerate(zip(xs, ns, ds)):
            if n < 2:
                continue
            pp = Plan(n, None, d, "plain_fp64", "none", out_capacity=max(1024, got[k].shape[0] + 1024),
                      amb_capacity=0)
            a = time.perf_counter()
            r = pp.run(x, None, float(Ts[k]), np.int32)
            plain_ms += (time.perf_counter() - a) * 1e3
            pk = r.numpy()
            if pk.shape != got[k].shape or not np.array_equal(pk, got[k]):
                ka = got[k][:, 0].astype(np.int64) * n + got[k][:, 1]
                kb = pk[:, 0].astype(np.int64) * n + pk[:, 1]
                diffs += int(np.setdiff1d(ka, kb).size + np.setdiff1d(kb, ka).size)
            pp.close()
        result["plain_gpu"] = {"value": pairs_total / (plain_ms / 1e3), "ms": plain_ms,
                               "note": "one plain FP64 join per set, each synchronised (includes its readback)"}
        result["speedup_vs_plain"] = value / (pairs_total / (plain_ms / 1e3))
        result["pair_diffs_vs_plain_fp64"] = diffs
        log(f"plain {plain_ms:.1f} ms, diffs {diffs}")

    if not args.no_cpu:
        from oracle import oracle
        oracle.build()
        threads = len(os.sched_getaffinity(0))
        pick = list(range(0, joins, max(1, joins // 64)))[:64]
        sp = sum(ns[k] * (ns[k] - 1) // 2 for k in pick)
        t0 = time.perf_counter()
        cnt = 0
        for k in pick:
            cnt += oracle.join(sets[k], None, float(Ts[k]), threads=threads, count_only=True)
        dt = time.perf_counter() - t0
        result["cpu_baseline"] = {"value": sp / dt, "unit": "pair-evals/s", "cores": threads, "kind": "port",
                                  "sample": f"{len(pick)} whole joins (every {max(1, joins // 64)}th), "
                                            f"{sp:.3e} pairs, {cnt} inside, {dt:.2f} s"}
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
