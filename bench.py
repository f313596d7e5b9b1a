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

The `learned` object times method 2 (the fast method; of the local-fragment
rescue work items only the packaged trained ranker's top ones per pair are
refined, with the wider fallback budget for pairs predicted to stay below
the rescue thresholds) on its 4096 held-out evaluation pairs (seeds
60001..64096, same distribution, used neither for training nor for choosing
budgets), with their own plain FP64 search as the reference; the `fast`
object times method 1 on the headline batch (one step).  Both give their TM
loss against the plain FP64 scores of the same pairs, overall and per
pipeline stage.

`candidate_roofline` covers the seed banks (K1: bank_kernel and its
continuation kernel): their algorithmic FP64 operations, counted on the
device per refinement pass (26 n + 31 n_sel + 3400) and TM-score (29 n),
over their device time.  `single_pair` times the entry the reference's
callers use, indialign_tmscore_search, on the single-pair workload and on a
sample of the headline pairs.  The CPU baseline also reports the single-pair
workload's latency on one host core and on all of them.

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
HELD_OUT_FIRST_SEED = 60001  # the learned method's evaluation pairs


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(pairs, sample):
    """Reference library (built from source here) on a systematic sample of
    `sample` pairs (every len/sample-th pair, so lengths and noise levels are
    spread like the whole workload) through its batch entry with OpenMP on
    all host cores, and the single-pair workload's latency through its
    single-pair entry on one core and on all cores."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import ctypes
    import libs
    if not os.path.exists(libs.REF_SO) or not os.path.exists(libs.REF_HARNESS_SO):
        return None
    from indialign_amd import synth
    env = os.environ.get("OMP_NUM_THREADS", "").split(",")[0]
    cores = int(env) if env.isdigit() and int(env) > 0 else len(os.sched_getaffinity(0))
    h = libs.harness()
    h.ref_set_threads(cores)
    step = max(1, len(pairs) // max(1, sample))
    sub = pairs[::step][:sample]
    t0 = time.perf_counter()
    libs.ref_batch(sub)
    dt = time.perf_counter() - t0
    lib = libs.ref()
    cfg = libs.Cfg()
    lib.indialign_default_config(ctypes.byref(cfg))
    c0 = libs.PackedPairs([synth.single_pair()])
    lat = {}
    for threads in (1, cores):
        h.ref_set_threads(threads)
        best = None
        for _ in range(3):
            r = libs.Res()
            t1 = time.perf_counter()
            lib.indialign_tmscore_search(ctypes.byref(c0.inps[0]), ctypes.byref(cfg), ctypes.byref(r))
            e = time.perf_counter() - t1
            best = e if best is None else min(best, e)
        lat["1_core" if threads == 1 else f"{cores}_cores"] = round(1000 * best, 2)
    h.ref_set_threads(cores)
    return {"value": round(len(sub) / dt, 4), "unit": "pairs/s", "cores": cores,
            "kind": "reference", "cpu": cpu_model(),
            "sample": f"{len(sub)} pairs (every {step}th pair of the same workload), reference "
                      f"batch entry, OpenMP over pairs, {dt:.1f} s",
            "single_pair_ms": lat,
            "single_pair_workload": "N=64, seed 0, sigma 3 A, reference single-pair entry, best of 3"}


def single_pair_latency(ia, pairs, k=16):
    """indialign_tmscore_search (the reference callers' entry) on the
    single-pair workload and on k pairs of the headline workload: ms per call,
    each call a fresh search of one pair (the first call's set-up excluded)."""
    from indialign_amd import synth
    c0 = synth.single_pair()
    ia.tmscore_search(c0.pred, c0.native)
    ts = []
    for _ in range(5):
        t = time.perf_counter()
        ia.tmscore_search(c0.pred, c0.native)
        ts.append(time.perf_counter() - t)
    step = max(1, len(pairs) // k)
    tp = []
    for p in pairs[::step][:k]:
        t = time.perf_counter()
        ia.tmscore_search(p.pred, p.native, p.valid, p.pvalid, p.nvalid, p.lnorm)
        tp.append(time.perf_counter() - t)
    return {"entry": "indialign_tmscore_search", "single_pair_ms": round(1000 * float(np.median(ts)), 3),
            "sample_pairs": len(tp), "sample_mean_ms": round(1000 * float(np.mean(tp)), 3),
            "sample_max_ms": round(1000 * float(np.max(tp)), 3)}


def workload_name(kind, n):
    if kind == "short":
        return (f"{n} pairs, seeds 1..{n}, N~U{{40,300}}, sigma~U(0.5,14) A, plain FP64 method, "
                "batch semantics")
    if kind == "mixed":
        return (f"{n} pairs, seeds 50001..{50000 + n}, N log-uniform over [27, 1500], "
                "sigma~U(0.5,14) A, all masks valid, plain FP64 method, batch semantics")
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


def method_leg(ia, cfg, plain_scores, plain_stage, plain_s_per_step, method, steps, warmup,
               label_pairs):
    """Another method on the resident batch; TM loss against the plain FP64
    scores of the same pairs, overall and per stage."""
    label = {ia.METHOD_FAST: "fast", ia.METHOD_LEARNED: "learned"}[method]
    if method == ia.METHOD_LEARNED:
        ia.load_default_ranker()
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
            ia.METHOD_LEARNED: "fast method + trained ranker: top 64 dense / 16 first-pass "
                               "local-fragment work items per pair, 512 / 64 for pairs predicted "
                               "to stay below the rescue thresholds"}[method]
    return {"method": desc, "pairs": label_pairs,
            "value": round(n / el, 3), "unit": "pairs/s", "ms_per_step": round(1000 * el, 2),
            "steps": steps,
            "speedup_vs_plain": round(plain_s_per_step / el, 3),
            "tm_loss_vs_plain": {"mean": float(d.mean()), "max": float(d.max()),
                                 "pairs_over_1e-3": int((d > 1e-3).sum()),
                                 "pairs_over_1e-4": int((d > 1e-4).sum())},
            "stage_abs_dtm_vs_plain": stage_delta(ia, plain_stage, stage),
            "nw_cells": st["nw_cells"], "chain_ms": round(st["chain_ms"], 2),
            "bank_ms": round(st["bank_ms"], 2),
            "fallback_budget_passes": st["fallback_budget_passes"]}


def plain_reference(ia, cfg, pairs):
    """Plain FP64 search of a batch (after one warm-up): scores, stage
    scores, seconds per search."""
    ia.batch_load(pairs)
    ia.batch_run(cfg, serial_bank=1)
    t = time.perf_counter()
    s, _, _ = ia.batch_run(cfg, serial_bank=1)
    el = time.perf_counter() - t
    return s, ia.last_stage_scores(len(pairs)), el


def measured_traffic(workload: str, n_pairs: int):
    """HBM bytes per plain chain-kernel launch from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summaries (profiles/chain_kernel_traffic.json,
    one entry per profiled workload); None for a workload not profiled."""
    p = os.path.join(ROOT, "profiles", "chain_kernel_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        entries = json.load(f).get("entries", [])
    for e in entries:
        if e.get("workload") == workload and e.get("pairs") == n_pairs:
            return e.get("bytes_per_launch")
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pairs", type=int, default=None,
                    help="pairs in the batch (default 4096 short, 512 long)")
    ap.add_argument("--workload", choices=["short", "long", "mixed"], default="short",
                    help="short: N~U{40,300} (the headline); long: N~U{400,1500} with masks "
                         "and domain motion; mixed: N log-uniform over [27, 1500]")
    ap.add_argument("--cpu-sample", type=int, default=None,
                    help="pairs in the CPU baseline's sample (default 64 short, 16 long)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-single", action="store_true",
                    help="skip the single-pair latency of indialign_tmscore_search")
    ap.add_argument("--legs", default="learned,fast",
                    help="comma list of the methods timed after the plain one: learned, fast "
                         "(empty: plain only)")
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
        args.pairs = {"short": 4096, "long": 512, "mixed": 1024}[args.workload]
    t0 = time.perf_counter()
    pairs = {"short": synth.short_pairs, "long": synth.long_pairs,
             "mixed": synth.mixed_pairs}[args.workload](args.pairs)
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

    # the chain kernel's algorithmic FP64 work: 15 ops per NW cell plus the
    # superpositions and scores of its chains (Kabsch on each aligned set,
    # its TM-score, the lite detailed search's passes), counted on the device
    chain_ms = sum(s["chain_ms"] for s in step_stats)
    cells = sum(s["nw_cells"] for s in step_stats)
    nw_ops = cells * OPS_PER_NW_CELL
    sup_ops = sum(s["chain_ops"] for s in step_stats)
    per_s = 1.0 / (chain_ms / 1e3) / 1e12 if chain_ms > 0 else 0.0
    achieved = (nw_ops + sup_ops) * per_s
    achieved_nw = nw_ops * per_s
    st = step_stats[-1]
    bank_tflops = st["bank_ops"] / (st["bank_ms"] / 1e3) / 1e12 if st["bank_ms"] > 0 else 0.0
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
        "refine_passes_per_s": round(st["refine_passes"] / (st["wall_ms"] / 1e3), 1),
        "nw_cells_per_s": round(st["nw_cells"] / (st["wall_ms"] / 1e3), 1),
        "roofline": {"kernel": "chain_kernel (fused TM score matrix + NW + traceback + Kabsch + lite search)",
                     "bound": "valu", "achieved": round(achieved, 3), "peak": FP64_VALU_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP64_VALU_PEAK_TFLOPS, 4),
                     "traffic": measured_traffic(args.workload, len(pairs)),
                     "ops_per_step": {"nw_cells_x15": nw_ops / args.steps,
                                      "superposition": sup_ops / args.steps},
                     "nw_cells_only": {"achieved": round(achieved_nw, 3),
                                       "frac": round(achieved_nw / FP64_VALU_PEAK_TFLOPS, 4)}},
        "candidate_roofline": {"kernel": "bank_kernel + bank_cont_kernel (seed-bank refine + TM-score)",
                               "bound": "valu", "ops": st["bank_ops"],
                               "device_ms": round(st["bank_ms"], 2),
                               "achieved": round(bank_tflops, 3), "peak": FP64_VALU_PEAK_TFLOPS,
                               "unit": "TFLOP/s",
                               "frac": round(bank_tflops / FP64_VALU_PEAK_TFLOPS, 4)},
        "mean_score": round(float(scores.mean()), 6),
    }
    legs = [x for x in args.legs.split(",") if x]
    if "fast" in legs:
        out["fast"] = method_leg(ia, cfg, scores, plain_stage, elapsed / args.steps,
                                 ia.METHOD_FAST, 1, 0, out["config"]["workload"])
    if "learned" in legs:
        if args.workload == "short":
            held = synth.short_pairs(len(pairs), first_seed=HELD_OUT_FIRST_SEED)
            hname = (f"{len(held)} held-out pairs, seeds {HELD_OUT_FIRST_SEED}..."
                     f"{HELD_OUT_FIRST_SEED + len(held) - 1}, same distribution")
        else:
            held, hname = pairs, out["config"]["workload"]
        hs, hstage, hel = plain_reference(ia, cfg, held)
        log(f"[bench] plain search of the learned method's pairs: {hel:.2f} s")
        out["learned"] = method_leg(ia, cfg, hs, hstage, hel, ia.METHOD_LEARNED, args.steps,
                                    args.warmup, hname)
    if not args.no_single:
        out["single_pair"] = single_pair_latency(ia, pairs)
    if not args.no_cpu:
        sample = args.cpu_sample or {"short": 64, "long": 16, "mixed": 32}[args.workload]
        log(f"[bench] cpu baseline on {sample} pairs ...")
        out["cpu_baseline"] = cpu_baseline(pairs, sample)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
