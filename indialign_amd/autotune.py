"""Launch-shape autotuner for the NW chain kernel and the seed banks.

The chain kernel and the seed-bank kernel carry the search's device time.
Their launch shapes are runtime choices (indialign_set_launch_tuning): the
chain kernel variant's wavefronts per SIMD (2 or 3, i.e. a 256- or 168-VGPR
budget), its persistent wavefronts per CU and where the NW trace lives (LDS
or a global slot); and the banks' pass cap, after which candidates still
refining continue in a second launch that packs them densely.  Results do not
depend on the shape, so the tuner times each kernel's device time
(indialign_last_stats: NW-chain and seed-bank milliseconds) on a sample of
the caller's workload for every candidate, checks that the scores are
identical, and writes the fastest shape to indialign_tuning.txt beside the
library, which the library reads at its first search.

Usage: python -m indialign_amd.autotune [--workload short|long] [--pairs P]
       [--method 0|1] [--out PATH] [--dry-run]
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "indialign_tuning.txt")

WAVES = (3, 2)
SLOTS = (0, 8, 4)   # 0: every resident wavefront
TRACE = (0, 1, 2)   # auto / LDS when it fits / global
BANK_CAPS = (0, 2, 3, 4, 5, 6, 8)  # 0: one bank launch; 5 is the default


def candidates():
    """Every (chain_waves, chain_slots_per_cu, chain_trace) shape, the
    library default first."""
    return list(itertools.product(WAVES, SLOTS, TRACE))


def time_shape(ia, cfg, shape, method, reps, key):
    ia.set_launch_tuning(*shape)
    best = float("inf")
    scores = None
    for _ in range(reps):
        s, R, t = ia.batch_run(cfg, serial_bank=1, method=method)
        st = ia.last_stats()
        best = min(best, st[key])
        scores = (s, R, t) if scores is None else scores
    return best, scores


def tune(pairs, method: int = 0, reps: int = 2, log=print):
    """Time every candidate chain shape, then every bank pass cap with the
    best chain shape, on `pairs` (kept resident); return (best_shape,
    table) with best_shape = (chain_waves, chain_slots_per_cu, chain_trace,
    bank_pass_cap).  Raises if any shape changes a result."""
    import indialign_amd as ia
    cfg = ia.default_config()
    ia.batch_load(pairs)
    ia.set_launch_tuning(3, 0, 0, 5)
    ia.batch_run(cfg, serial_bank=1, method=method)  # warm-up (buffers, code objects)
    ref = None
    table = []

    def check(shape, out):
        nonlocal ref
        if ref is None:
            ref = out
            return
        for a, b in zip(out, ref):
            if not np.array_equal(a, b):
                raise RuntimeError(f"launch shape {shape} changed the results")

    for shape in candidates():
        ms, out = time_shape(ia, cfg, shape + (5,), method, reps, "chain_ms")
        check(shape, out)
        table.append({"chain_waves": shape[0], "chain_slots_per_cu": shape[1],
                      "chain_trace": shape[2], "chain_ms": round(ms, 2)})
        log(f"[autotune] waves {shape[0]} slots/CU {shape[1] or 'all'} trace "
            f"{['auto', 'lds', 'global'][shape[2]]}: chain kernel {ms:.1f} ms")
    best = min(table, key=lambda r: r["chain_ms"])
    chain = (best["chain_waves"], best["chain_slots_per_cu"], best["chain_trace"])
    caps = []
    for cap in BANK_CAPS:
        ms, out = time_shape(ia, cfg, chain + (cap,), method, reps, "bank_ms")
        check(chain + (cap,), out)
        caps.append({"bank_pass_cap": cap, "bank_ms": round(ms, 2)})
        log(f"[autotune] bank pass cap {cap or 'none'}: seed banks {ms:.1f} ms")
    table += caps
    cap = min(caps, key=lambda r: r["bank_ms"])["bank_pass_cap"]
    return chain + (cap,), table


def write_tuning(path, shape, note):
    with open(path, "w") as f:
        f.write(f"# written by indialign_amd/autotune.py: {note}\n")
        f.write(f"chain_waves {shape[0]}\nchain_slots_per_cu {shape[1]}\nchain_trace {shape[2]}\n"
                f"bank_pass_cap {shape[3]}\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["short", "long"], default="short")
    ap.add_argument("--pairs", type=int, default=None)
    ap.add_argument("--method", type=int, default=0, choices=[0, 1])
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--dry-run", action="store_true", help="print the result, write nothing")
    args = ap.parse_args()
    sys.path.insert(0, os.path.dirname(HERE))
    import indialign_amd as ia
    from indialign_amd import synth
    if not ia.gpu_available():
        sys.exit("autotune: no GPU visible")
    n = args.pairs or (512 if args.workload == "short" else 64)
    pairs = synth.short_pairs(n) if args.workload == "short" else synth.long_pairs(n)
    t0 = time.perf_counter()
    shape, table = tune(pairs, args.method, args.reps, log=lambda m: print(m, file=sys.stderr, flush=True))
    note = (f"{args.workload} workload, {n} pairs, method {args.method}, "
            f"{time.perf_counter() - t0:.0f} s")
    print(json.dumps({"best": {"chain_waves": shape[0], "chain_slots_per_cu": shape[1],
                               "chain_trace": shape[2], "bank_pass_cap": shape[3]},
                      "note": note, "table": table}))
    if not args.dry_run:
        write_tuning(args.out, shape, note)


if __name__ == "__main__":
    main()
