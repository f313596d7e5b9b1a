"""Fair comparison harness: one common TM-score applied to every tool's (R, t).

Restates the reference's benchmark harness over this library
(benchmark/fair_benchmark.py:49-61 scorer and d0, :64-85 caller,
:117-123 generator; benchmark/pdb_benchmark.py:116-141 C1' extraction) so a
README-style comparison can be reproduced on one GPU:

- `d0_rna`, `common_tm`: the harness's own d0 (no 0.3 floor, unlike the
  search's core.h:12-25) and TM-score of `pred @ R.T + t` against native,
  the same formula for every tool, so no tool grades itself;
- `extract_c1_models`: C1' coordinates per MODEL from PDB text (local files:
  there is no download here);
- `synthetic_pairs`: the harness's random-walk chains (numpy RandomState,
  as the reference draws them);
- `run_indialign`: this library's batch search, returning (score, R, t) in
  the caller convention (R = rotation.reshape(3, 3).T, applied as
  pred @ R.T + t);
- `run_usalign`: the external USalign binary when one is on PATH or in
  $USALIGN (absent here; the comparison then covers this library alone);
- `compare`: wins / ties on the common score, self-reported vs common,
  sign and Wilcoxon tests, strata by difficulty.

    python -m indialign_amd.fair_benchmark [n_pairs] [--pdb-dir DIR] [--json OUT]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import indialign_amd as ia  # noqa: E402


def d0_rna(L: float) -> float:
    """d0 of the common scorer (benchmark/fair_benchmark.py:49-56)."""
    if L <= 11:
        return 0.3
    if L <= 15:
        return 0.4
    if L <= 19:
        return 0.5
    if L <= 23:
        return 0.6
    if L < 30:
        return 0.7
    return 0.6 * math.sqrt(L - 0.5) - 2.5


def common_tm(pred, native, R, t, d0, Lnorm) -> float:
    """sum 1 / (1 + d^2 / d0^2) / Lnorm over all residues of pred @ R.T + t
    against native (benchmark/fair_benchmark.py:58-61)."""
    moved = np.asarray(pred, np.float64) @ np.asarray(R, np.float64).T + np.asarray(t, np.float64)
    d2 = np.sum((moved - np.asarray(native, np.float64)) ** 2, axis=1)
    return float(np.sum(1.0 / (1.0 + d2 / (d0 * d0))) / Lnorm)


def extract_c1_models(pdb_text: str) -> list[np.ndarray]:
    """C1' coordinates of each MODEL (or of the whole file when it has no
    MODEL records), in file order (benchmark/pdb_benchmark.py:116-141)."""
    models, current = [], []
    for line in pdb_text.splitlines():
        rec = line[:6].strip()
        if rec == "MODEL":
            current = []
        elif rec == "ENDMDL":
            if current:
                models.append(np.array(current))
            current = []
        elif rec == "ATOM" and line[12:16].strip() == "C1'":
            current.append([float(line[30:38]), float(line[38:46]), float(line[46:54])])
    if not models and current:
        models.append(np.array(current))
    return models


def make_rna(N: int, rng: np.random.RandomState) -> np.ndarray:
    """Random-walk chain of the harness (benchmark/fair_benchmark.py:117-123):
    5.9 A steps (sd 0.5) whose direction drifts by N(0, 0.3^2) per step, centred."""
    coords = np.zeros((N, 3))
    d = rng.randn(3)
    d /= np.linalg.norm(d)
    for i in range(1, N):
        d = d + rng.randn(3) * 0.3
        d /= np.linalg.norm(d)
        coords[i] = coords[i - 1] + d * (5.9 + rng.randn() * 0.5)
    return coords - coords.mean(axis=0)


def synthetic_pairs(n_pairs: int, seed: int = 42):
    """The harness's workload (benchmark/fair_benchmark.py:126-134): N~U{40,300},
    sigma~U(0.5,14), Haar rotation, N(0, 20^2) A translation per axis."""
    from scipy.spatial.transform import Rotation
    rng = np.random.RandomState(seed)
    out = []
    for _ in range(n_pairs):
        N = rng.randint(40, 301)
        sigma = rng.uniform(0.5, 14.0)
        native = make_rna(N, rng)
        pred = (Rotation.random(random_state=rng).apply(native) + rng.randn(N, 3) * sigma
                + rng.randn(3) * 20)
        out.append((pred, native, sigma))
    return out


def run_indialign(pairs, method: int = ia.METHOD_PLAIN):
    """(score, R, t) per (pred, native) pair, all masks valid, Lnorm = N,
    through one batched search; R in the caller convention (the ABI's
    rotation transposed, applied as pred @ R.T + t, fair_benchmark.py:82-85).
    Returns (list of (score, R, t), seconds for the whole batch)."""
    tup = [(p, q, None, None, None, float(len(q))) for p, q in pairs]
    t0 = time.perf_counter()
    s, R, t = ia.tmscore_search_batch(tup, method=method)
    dt = time.perf_counter() - t0
    return [(float(s[k]), R[k].reshape(3, 3).T.copy(), t[k].copy()) for k in range(len(tup))], dt


def usalign_path() -> str | None:
    p = os.environ.get("USALIGN", "USalign")
    return p if os.path.isfile(p) and os.access(p, os.X_OK) else shutil.which(p)


def _write_pdb(coords, path):
    with open(path, "w") as f:
        for i, (x, y, z) in enumerate(coords):
            f.write(f"ATOM  {i + 1:>5d}  C1'   A A{i + 1:>4d}    "
                    f"{x:>8.3f}{y:>8.3f}{z:>8.3f}  1.00  0.00           C\n")


def run_usalign(pred, native, exe: str):
    """(score, R, t, seconds) from the USalign binary, C1' atoms, TM-score
    normalised by the native (benchmark/fair_benchmark.py:88-115)."""
    with tempfile.TemporaryDirectory() as d:
        pp, nn, mf = (os.path.join(d, x) for x in ("p.pdb", "n.pdb", "m.txt"))
        _write_pdb(pred, pp)
        _write_pdb(native, nn)
        t0 = time.perf_counter()
        out = subprocess.run([exe, pp, nn, "-atom", " C1'", "-TMscore", "1", "-m", mf],
                             capture_output=True, text=True, timeout=60)
        dt = time.perf_counter() - t0
        m = re.findall(r"TM-score=\s+([\d.]+)", out.stdout)
        score = float(m[1]) if len(m) >= 2 else 0.0
        R, t = np.zeros((3, 3)), np.zeros(3)
        for line in open(mf):
            parts = line.split()
            if len(parts) == 5 and parts[0] in ("0", "1", "2"):
                k = int(parts[0])
                t[k] = float(parts[1])
                R[k] = [float(x) for x in parts[2:]]
    return score, R, t, dt


STRATA = [(0, 0.17, "very hard (TM<0.17)"), (0.17, 0.3, "hard (0.17-0.3)"),
          (0.3, 0.5, "medium (0.3-0.5)"), (0.5, 0.7, "moderate (0.5-0.7)"),
          (0.7, 1.01, "easy (TM>0.7)")]


def compare(rows) -> dict:
    """Summary of rows with keys self_i, common_i and, when a second tool
    ran, self_u, common_u (benchmark/fair_benchmark.py:160-228)."""
    ci = np.array([r["common_i"] for r in rows])
    si = np.array([r["self_i"] for r in rows])
    out = {"pairs": len(rows), "mean_common_indialign": float(ci.mean()),
           "mean_self_indialign": float(si.mean()),
           "max_abs_self_minus_common_indialign": float(np.abs(si - ci).max())}
    if len(rows) > 1 and si.std() > 0 and ci.std() > 0:
        out["self_common_corr_indialign"] = float(np.corrcoef(si, ci)[0, 1])
    if rows and "common_u" in rows[0]:
        cu = np.array([r["common_u"] for r in rows])
        d = ci - cu
        wi, wu = int((d > 1e-6).sum()), int((d < -1e-6).sum())
        out.update({"mean_common_usalign": float(cu.mean()), "indialign_wins": wi,
                    "usalign_wins": wu, "ties": len(rows) - wi - wu,
                    "mean_delta": float(d.mean()), "median_delta": float(np.median(d))})
        nontie = d[np.abs(d) > 1e-6]
        if len(nontie):
            from scipy.stats import binomtest
            pos = int((nontie > 0).sum())
            out["sign_test_p"] = float(binomtest(pos, len(nontie), 0.5, alternative="greater").pvalue)
        if len(nontie) >= 10:
            from scipy.stats import wilcoxon
            out["wilcoxon_p"] = float(wilcoxon(nontie, zero_method="wilcox",
                                               alternative="greater").pvalue)
        avg = (ci + cu) / 2
        out["strata"] = {name: {"pairs": int(m.sum()), "mean_delta": float(d[m].mean())}
                         for lo, hi, name in STRATA if (m := (avg >= lo) & (avg < hi)).any()}
    return out


def pdb_pairs(pdb_dir: str, max_models: int = 5):
    """Model 0 as native against models spread over the ensemble
    (benchmark/pdb_benchmark.py:215-232), from local *.pdb files whose
    models all have the same C1' count (20..4096)."""
    pairs, names = [], []
    for fn in sorted(os.listdir(pdb_dir)):
        if not fn.lower().endswith(".pdb"):
            continue
        models = extract_c1_models(open(os.path.join(pdb_dir, fn)).read())
        if len(models) < 2 or len({len(m) for m in models}) != 1 or not 20 <= len(models[0]) <= 4096:
            continue
        n = len(models)
        test = list(range(1, n)) if n <= max_models else sorted(
            {1, n // 4, n // 2, 3 * n // 4, n - 1})
        for mi in test:
            pairs.append((models[mi], models[0]))
            names.append(f"{fn}:{mi}")
    return pairs, names


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("n_pairs", nargs="?", type=int, default=50)
    ap.add_argument("--pdb-dir", default=None, help="local PDB files instead of synthetic pairs")
    ap.add_argument("--method", type=int, default=ia.METHOD_PLAIN)
    ap.add_argument("--json", default=None)
    args = ap.parse_args(argv)
    if args.pdb_dir:
        pairs, _ = pdb_pairs(args.pdb_dir)
    else:
        pairs = [(p, q) for p, q, _ in synthetic_pairs(args.n_pairs)]
    res, dt = run_indialign(pairs, method=args.method)
    exe = usalign_path()
    rows = []
    for (pred, native), (s, R, t) in zip(pairs, res):
        N = len(native)
        d0 = d0_rna(N)
        row = {"N": N, "self_i": s, "common_i": common_tm(pred, native, R, t, d0, N)}
        if exe:
            su, Ru, tu, dtu = run_usalign(pred, native, exe)
            row.update(self_u=su, common_u=common_tm(pred, native, Ru, tu, d0, N), dt_u=dtu)
        rows.append(row)
    summary = compare(rows)
    summary["indialign_ms_per_pair"] = 1000 * dt / max(1, len(pairs))
    summary["usalign"] = exe or "not found (set USALIGN to compare)"
    print(json.dumps(summary, indent=1))
    if args.json:
        with open(args.json, "w") as f:
            json.dump({"summary": summary, "rows": rows}, f, indent=1)
    return summary


if __name__ == "__main__":
    main()
