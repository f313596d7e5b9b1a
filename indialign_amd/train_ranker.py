"""Train the local-fragment work-item ranker on the GPU.

    python -m indialign_amd.train_ranker [--pairs 200000] [--first-seed 100000]
        [--top-per-pass 4] [--rand-per-pass 10] [--epochs 4,4,2]
        [--out indialign_amd/ranker_params.npy]

1. Training pairs: the synthetic generator (synth.short_pairs: N~U{40,300},
   sigma~U(0.5,14) A) with seeds first_seed.., disjoint from the benchmark
   pairs (seeds 1..4096), from the held-out evaluation pairs (seeds
   60001..64096) and from the validation pairs (seeds val_seed..).
2. Records: the plain FP64 search with recording on
   (indialign_collect_records): one row per local-fragment work item, its
   96-value descriptor and its refined score (best of its two NW chains).
   Of each pair's pass (first or dense local-fragment pass) the rows kept are
   the top_per_pass items by refined score and rand_per_pass others drawn
   uniformly, so that many pairs contribute a few rows each (about 5M rows
   from 200k pairs) instead of few pairs contributing every row.
3. Adam on the GPU (indialign_mlp_train: BF16-operand matrix-core forward and
   backward kernels) in phases of decreasing learning rate, MSE on the
   refined score.
4. Validation on separate pairs with every row kept: per pair and pass,
   whether the best work item is among the top-m by predicted score, and how
   much refined score the top-m miss, next to the same numbers for ranking by
   single descriptor columns (11: same-index TM-score of the fragment
   superposition, 12: register pairing, 30/31: after one/two cutoff
   refinements; 31 is the network's skip input).

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


def _make_chunk(args):
    first, m = args
    return synth.short_pairs(m, first_seed=first)


def _subsample(X, y, pi, top_k, rand_k, rng):
    """Rows of each (pair, pass) group: its top_k by refined score and
    rand_k of the others drawn uniformly."""
    if top_k < 0:
        return X, y, pi
    key = pi.astype(np.int64) * 2 + (X[:, 8] > 0.5)
    order = np.lexsort((-y, key))
    ks = key[order]
    start = np.flatnonzero(np.r_[True, ks[1:] != ks[:-1]])
    size = np.diff(np.r_[start, len(order)])
    rank = np.arange(len(order)) - np.repeat(start, size)
    rest = np.maximum(np.repeat(size, size) - top_k, 1)
    keep = (rank < top_k) | (rng.uniform(size=len(order)) < rand_k / rest)
    idx = np.sort(order[keep])
    return X[idx], y[idx], pi[idx]


def collect(pool, n_pairs, first_seed, chunk=1024, top_k=-1, rand_k=0, seed=0):
    """Records of the plain search over n_pairs pairs from first_seed (the
    pairs are generated by `pool`); top_k < 0 keeps every row."""
    rng = np.random.default_rng(seed)
    Xs, ys, ps = [], [], []
    jobs = [(first_seed + c0, min(chunk, n_pairs - c0)) for c0 in range(0, n_pairs, chunk)]
    for k, pairs in enumerate(pool.imap(_make_chunk, jobs)):
        c0 = k * chunk
        _, _, _, X, y, pi = ia.collect_records(pairs, cap=8_000_000)
        X, y, pi = _subsample(X, y, pi, top_k, rand_k, rng)
        Xs.append(X)
        ys.append(y)
        ps.append(pi + c0)
        log(f"[train] records: pairs {c0 + len(pairs)}/{n_pairs}, rows {sum(len(v) for v in ys)}")
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
    ap.add_argument("--pairs", type=int, default=200000)
    ap.add_argument("--first-seed", type=int, default=100000)
    ap.add_argument("--top-per-pass", type=int, default=4,
                    help="rows kept per (pair, pass) from its best items (-1: keep every row)")
    ap.add_argument("--rand-per-pass", type=int, default=10,
                    help="rows kept per (pair, pass) drawn uniformly from the others")
    ap.add_argument("--val-pairs", type=int, default=256)
    ap.add_argument("--val-seed", type=int, default=300000)
    ap.add_argument("--epochs", default="4,4,2", help="epochs per phase (lr 1e-3, 3e-4, 1e-4)")
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--top-m", type=int, default=64)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=DEFAULT_OUT)
    args = ap.parse_args()
    import multiprocessing as mp
    # the generator workers are forked before this process touches the GPU
    pool = mp.get_context("fork").Pool(8)
    if not ia.gpu_available():
        sys.exit("train_ranker: no GPU visible")
    lo, hi = args.first_seed, args.first_seed + args.pairs
    for name, a, b in (("benchmark", 1, 4097), ("held-out", 60001, 64097),
                       ("validation", args.val_seed, args.val_seed + args.val_pairs)):
        if lo < b and a < hi:
            sys.exit(f"train_ranker: training seeds {lo}..{hi - 1} overlap the {name} seeds")

    t0 = time.time()
    Xt, yt, _ = collect(pool, args.pairs, args.first_seed, top_k=args.top_per_pass,
                        rand_k=args.rand_per_pass, seed=args.seed)
    ok = np.isfinite(yt) & (yt > -1.0)
    Xt, yt = Xt[ok], yt[ok]
    log(f"[train] {len(yt)} training rows from {args.pairs} pairs in {time.time() - t0:.1f} s")
    Xv, yv, pv = collect(pool, args.val_pairs, args.val_seed, chunk=256)
    pool.close()
    ok = np.isfinite(yv) & (yv > -1.0)
    Xv, yv, pv = Xv[ok], yv[ok], pv[ok]

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
        "pairs": args.pairs, "first_seed": args.first_seed,
        "top_per_pass": args.top_per_pass, "rand_per_pass": args.rand_per_pass,
        "train_rows": int(len(yt)), "val_rows": int(len(yv)), "val_pairs": args.val_pairs,
        "val_seed": args.val_seed, "epoch_loss": history, "val_mse": val_mse,
        "top_m": args.top_m, "val_learned": learned, "val_descriptor_columns": heuristic,
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
