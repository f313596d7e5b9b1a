#!/usr/bin/env python3
"""Throughput of the TM-score superposition search on one MI355X.

Workload: 4096 synthetic pairs, seeds 1..4096, N~U{40,300}, sigma~U(0.5,14) A,
Haar rotation + U(-50,50) A translation, all masks valid, Lnorm = N, default
NativeConfig, plain FP64 method, batch-entry semantics (serial seed bank).

A step is one search of the whole batch with the inputs already resident in
device memory (indialign_batch_load once, indialign_batch_run per step; the
call returns after the device is idle).  `value` is aligned pairs per second
over the timed steps.  The roofline object covers the dominant kernel (the
NW chain kernel): algorithmic FP64 operations of the fused score matrix + NW
fill (15 per cell) divided by its device time from HIP events recorded on
the stream the kernel runs on, against the FP64 vector peak.  The CPU
baseline is the reference library compiled from source (oracle/_ref), batch
entry with OpenMP over pairs on the host cores, on a bounded sample of the
same workload.

The `learned` object times method 2 on the same batch (the fast method, and
of the local-fragment rescue work items only the top-64 per pair by the
packaged trained ranker are refined), the `fast` object method 1 (one step);
both give their TM loss against the plain FP64 scores of the same pairs,
overall and per pipeline stage.

`--workload long` runs the long-pair configuration instead (512 pairs, N 400-1500,
masks and domain motion; use --steps 1 --warmup 0 --cpu-sample 2: one plain search
takes minutes).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--pairs P] [--workload short|long]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "aligned pairs/sec on 1 MI355X (40-300 nt) + TM loss vs plain FP64 search"
FP64_VALU_PEAK_TFLOPS = 78.6
OPS_PER_NW_CELL = 15  # d2 (8) + cut, divide, add (3) + diag add, gap add, 2 compares (4)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(pairs, sample):
    """Reference library (built from source here) over the first `sample`
    pairs with OpenMP on all host cores."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import libs
    if not os.path.exists(libs.REF_SO):
        return None
    cores = os.cpu_count() or 1
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    sub = pairs[:sample]
    t0 = time.perf_counter()
    libs.ref_batch(sub)
    dt = time.perf_counter() - t0
    return {"value": round(len(sub) / dt, 4), "unit": "pairs/s",
            "cores": int(os.environ.get("OMP_NUM_THREADS", cores)), "kind": "reference",
            "sample": f"first {len(sub)} pairs of the same workload, reference batch entry, "
                      f"{dt:.1f} s"}


def workload_name(kind, n):
    if kind == "short":
        return (f"{n} pairs, seeds 1..{n}, N~U{{40,300}}, sigma~U(0.5,14) A, plain FP64 method, "
                "batch semantics")
    return (f"{n} pairs, seeds 20001..{20000 + n}, N~U{{400,1500}}, sigma~U(2,10) A, 20% of helix "
            "segments moved as rigid domains, 10% of pred and native residues masked, Lnorm = "
            "native valid count, plain FP64 method, batch semantics")


def stage_delta(ia, plain_stage, stage):
    """max / mean |dTM| of the running best after each pipeline stage,
    over pairs where the stage ran in both searches."""
    out = {}
    for k, name in enumerate(ia.STAGES):
        a, b = plain_stage[:, k], stage[:, k]
        ok = np.isfinite(a) & np.isfinite(b)
        if ok.any():
            d = np.abs(a[ok] - b[ok])
            out[name] = {"max": float(d.max()), "mean": float(d.mean()), "pairs": int(ok.sum())}
    return out


def method_leg(ia, cfg, plain_scores, plain_stage, args, plain_s_per_step, method, steps, warmup):
    """Another method on the same resident batch; TM loss against the plain
    FP64 scores of the same pairs, overall and per stage."""
    label = {ia.METHOD_FAST: "fast", ia.METHOD_LEARNED: "learned"}[method]
    if method == ia.METHOD_LEARNED:
        ia.load_default_ranker(top_m=args.top_m)
    for w in range(warmup):
        t = time.perf_counter()
        ia.batch_run(cfg, serial_bank=1, method=method)
        log(f"[bench] {label} warmup {w}: {time.perf_counter() - t:.2f} s")
    t0 = time.perf_counter()
    for k in range(steps):
        t = time.perf_counter()
        s, _, _ = ia.batch_run(cfg, serial_bank=1, method=method)
        log(f"[bench] {label} step {k}: {time.perf_counter() - t:.2f} s")
    el = (time.perf_counter() - t0) / steps
    st = ia.last_stats()
    stage = ia.last_stage_scores(len(s))
    ia.set_ranker(None)
    d = plain_scores - s
    n = len(s)
    desc = {ia.METHOD_FAST: "FP32 NW fill, single-pass FP64 moments, closed-form quaternion rotation",
            ia.METHOD_LEARNED: f"fast method + trained ranker, top-{args.top_m} local-fragment "
                               "work items per pair and pass"}[method]
    return {"method": desc,
            "value": round(n / el, 3), "unit": "pairs/s", "ms_per_step": round(1000 * el, 2),
            "steps": steps,
            "speedup_vs_plain": round(plain_s_per_step / el, 3),
            "tm_loss_vs_plain": {"mean": float(d.mean()), "max": float(d.max()),
                                 "pairs_over_1e-3": int((d > 1e-3).sum()),
                                 "pairs_over_1e-4": int((d > 1e-4).sum())},
            "stage_abs_dtm_vs_plain": stage_delta(ia, plain_stage, stage),
            "nw_cells": st["nw_cells"], "chain_ms": round(st["chain_ms"], 2),
            "bank_ms": round(st["bank_ms"], 2)}


def measured_traffic():
    """HBM bytes per chain-kernel launch from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (profiles/chain_kernel_short4096.md)."""
    p = os.path.join(ROOT, "profiles", "chain_kernel_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pairs", type=int, default=None,
                    help="pairs in the batch (default 4096 short, 512 long)")
    ap.add_argument("--workload", choices=["short", "long"], default="short",
                    help="short: N~U{40,300} (the headline); long: N~U{400,1500} with masks "
                         "and domain motion")
    ap.add_argument("--cpu-sample", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--legs", default="learned,fast",
                    help="comma list of the methods timed after the plain one: learned, fast "
                         "(empty: plain only)")
    ap.add_argument("--top-m", type=int, default=64)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a device-to-device copy kernel of this size (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe_mib:
        import torch
        n = args.copy_probe_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 4, "launches": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(pairs are independent: a multi-GPU run would split the batch by pair cost, one process per GPU)", file=sys.stderr)
        sys.exit(2)

    import indialign_amd as ia
    from indialign_amd import synth

    if args.pairs is None:
        args.pairs = 4096 if args.workload == "short" else 512
    t0 = time.perf_counter()
    pairs = synth.short_pairs(args.pairs) if args.workload == "short" else synth.long_pairs(args.pairs)
    log(f"[bench] generated {len(pairs)} pairs in {time.perf_counter() - t0:.1f} s")
    if not ia.gpu_available():
        sys.exit("bench.py: no GPU visible")
    cfg = ia.default_config()
    ia.batch_load(pairs)
    for w in range(args.warmup):
        t = time.perf_counter()
        ia.batch_run(cfg, serial_bank=1)
        log(f"[bench] warmup {w}: {time.perf_counter() - t:.2f} s")
    step_stats = []
    t_start = time.perf_counter()
    for k in range(args.steps):
        t = time.perf_counter()
        scores, R, T = ia.batch_run(cfg, serial_bank=1)
        step_stats.append(ia.last_stats())
        plain_stage = ia.last_stage_scores(len(pairs))
        log(f"[bench] step {k}: {time.perf_counter() - t:.2f} s")
    elapsed = time.perf_counter() - t_start
    ms_per_step = 1000.0 * elapsed / args.steps
    value = len(pairs) * args.steps / elapsed

    chain_ms = sum(s["chain_ms"] for s in step_stats)
    cells = sum(s["nw_cells"] for s in step_stats)
    achieved = cells * OPS_PER_NW_CELL / (chain_ms / 1e3) / 1e12 if chain_ms > 0 else 0.0
    st = step_stats[-1]
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": workload_name(args.workload, len(pairs)),
                   "pairs": len(pairs), "method": "plain-fp64"},
        "tm_loss_vs_plain": {"mean": 0.0, "max": 0.0},
        "candidate_evals_per_s": round(st["candidates"] / (st["wall_ms"] / 1e3), 1),
        "nw_cells_per_s": round(st["nw_cells"] / (st["wall_ms"] / 1e3), 1),
        "roofline": {"kernel": "chain_kernel (fused TM score matrix + NW + traceback + Kabsch + lite search)",
                     "bound": "valu", "achieved": round(achieved, 3), "peak": FP64_VALU_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP64_VALU_PEAK_TFLOPS, 4),
                     "traffic": measured_traffic()},
        "mean_score": round(float(scores.mean()), 6),
    }
    legs = [x for x in args.legs.split(",") if x]
    if "learned" in legs:
        out["learned"] = method_leg(ia, cfg, scores, plain_stage, args, elapsed / args.steps,
                                    ia.METHOD_LEARNED, args.steps, args.warmup)
    if "fast" in legs:
        out["fast"] = method_leg(ia, cfg, scores, plain_stage, args, elapsed / args.steps,
                                 ia.METHOD_FAST, 1, 0)
    if not args.no_cpu:
        log(f"[bench] cpu baseline on {args.cpu_sample} pairs ...")
        out["cpu_baseline"] = cpu_baseline(pairs, args.cpu_sample)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
