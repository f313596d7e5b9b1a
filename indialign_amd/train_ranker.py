"""Train the local-fragment work-item ranker on the GPU.

    python -m indialign_amd.train_ranker [--pairs 2048] [--first-seed 10000]
        [--epochs 4,4,2] [--out indialign_amd/ranker_params.npy]

1. Training pairs: the synthetic generator (synth.short_pairs: N~U{40,300},
   sigma~U(0.5,14) A) with seeds first_seed.., disjoint from the benchmark
   pairs (seeds 1..4096).
2. Records: the plain FP64 search with recording on
   (indialign_collect_records): one row per local-fragment work item, its
   96-value descriptor and its refined score (best of its two NW chains).
3. Split by pair (10% validation), Adam on the GPU (indialign_mlp_train:
   BF16-operand matrix-core forward and backward kernels) in phases of
   decreasing learning rate, MSE on the refined score.
4. Validation: per pair and pass, whether the best work item is among the
   top-m by predicted score, and how much refined score the top-m miss,
   next to the same numbers for ranking by single descriptor columns (11:
   same-index TM-score of the fragment superposition, 12: register pairing,
   30/31: after one/two cutoff refinements; 31 is the network's skip input).

Writes the parameters (float32, .npy) and a JSON summary beside them.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import indialign_amd as ia  # noqa: E402
from indialign_amd import synth  # noqa: E402

DEFAULT_OUT = os.path.join(HERE, "ranker_params.npy")


def log(msg):
    print(msg, flush=True)


def collect(n_pairs, first_seed, chunk=256):
    Xs, ys, ps = [], [], []
    for c0 in range(0, n_pairs, chunk):
        m = min(chunk, n_pairs - c0)
        pairs = synth.short_pairs(m, first_seed=first_seed + c0)
        _, _, _, X, y, pi = ia.collect_records(pairs, cap=4_000_000)
        Xs.append(X)
        ys.append(y)
        ps.append(pi + c0)
        log(f"[train] records: pairs {c0 + m}/{n_pairs}, rows {sum(len(v) for v in ys)}")
    return np.concatenate(Xs), np.concatenate(ys), np.concatenate(ps)


def topm_quality(score, y, pair, dense, m):
    """Per (pair, pass): recall of the best item in the top-m by `score`, and
    the refined score lost (best item - best of the top-m)."""
    order = np.lexsort((np.arange(len(y)), -score, dense, pair))
    hit, miss = [], []
    keys = pair.astype(np.int64) * 2 + dense.astype(np.int64)
    ks = keys[order]
    starts = np.flatnonzero(np.r_[True, ks[1:] != ks[:-1]])
    ends = np.r_[starts[1:], len(order)]
    for a, b in zip(starts, ends):
        grp = order[a:b]
        best = y[grp].max()
        top = y[grp[:m]].max()
        hit.append(top >= best)
        miss.append(best - top)
    miss = np.asarray(miss)
    return {"groups": len(hit), "recall_best": float(np.mean(hit)),
            "miss_mean": float(miss.mean()), "miss_max": float(miss.max())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=2048)
    ap.add_argument("--first-seed", type=int, default=10000)
    ap.add_argument("--epochs", default="4,4,2", help="epochs per phase (lr 1e-3, 3e-4, 1e-4)")
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--top-m", type=int, default=64)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=DEFAULT_OUT)
    args = ap.parse_args()
    if not ia.gpu_available():
        sys.exit("train_ranker: no GPU visible")

    t0 = time.time()
    X, y, pair = collect(args.pairs, args.first_seed)
    log(f"[train] {len(y)} records from {args.pairs} pairs in {time.time() - t0:.1f} s")
    ok = np.isfinite(y) & (y > -1.0)
    X, y, pair = X[ok], y[ok], pair[ok]
    rng = np.random.default_rng(args.seed)
    val_pairs = rng.permutation(args.pairs)[: max(1, args.pairs // 10)]
    is_val = np.isin(pair, val_pairs)
    Xt, yt = X[~is_val], y[~is_val]
    Xv, yv, pv = X[is_val], y[is_val], pair[is_val]

    params = ia.init_params(args.seed)
    history = []
    phases = [int(e) for e in args.epochs.split(",")]
    for k, (ep, lr) in enumerate(zip(phases, [1e-3, 3e-4, 1e-4])):
        if ep <= 0:
            continue
        t1 = time.time()
        params, el = ia.mlp_train(params, Xt, yt, epochs=ep, batch=args.batch, lr=lr,
                                  seed=args.seed + 101 * k)
        history += [float(v) for v in el]
        log(f"[train] phase {k}: lr {lr:g}, {ep} epochs, loss {el.tolist()} "
            f"({time.time() - t1:.1f} s)")

    pred = ia.mlp_forward(params, Xv)
    val_mse = float(np.mean((pred - yv) ** 2))
    dense = Xv[:, 8] > 0.5
    ms = sorted({16, args.top_m})
    learned = {m: topm_quality(pred, yv, pv, dense, m) for m in ms}
    heuristic = {f"col{c}": {m: topm_quality(Xv[:, c], yv, pv, dense, m) for m in ms}
                 for c in (11, 12, 30, 31)}
    summary = {
        "pairs": args.pairs, "first_seed": args.first_seed, "records": int(len(y)),
        "train_rows": int(len(yt)), "val_rows": int(len(yv)), "val_pairs": int(len(val_pairs)),
        "epoch_loss": history, "val_mse": val_mse, "top_m": args.top_m,
        "val_learned": learned, "val_descriptor_columns": heuristic,
        "seconds": round(time.time() - t0, 1),
    }
    log("[train] " + json.dumps(summary))
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.save(args.out, params.astype(np.float32))
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(summary, f, indent=1)
    log(f"[train] wrote {args.out}")


if __name__ == "__main__":
    main()
