"""Vectors of the reference's PyTorch engine for the torch-semantics search.

The reference's scoring/ drivers call its PyTorch engine
(scoring/_tmscore_gpu/search.py:492-761) with dp_iter = 1 and same-index
masks by default (scoring/_local_validation_gpu/runtime.py:16,
scoring.py:52-64).  The reference holds one snapshot of that engine (the
5-residue toy at dp_iter = 0, tests/test_gpu_scoring_refactor.py:176-208) and
nothing for dp_iter > 0, so this script imports the engine from the
reference checkout, runs it on the CPU in float64 on seeded synthetic pairs
(one pair per call, B = 1), and stores the final scores for dp_iter = 0, 1
and 2 in torch_engine.npz.  The pairs are regenerated from their seeds by
indialign_amd.synth; the file keeps an input checksum per pair.

Usage (in a container with the reference at /root/reference):
    python tests/golden/make_torch_fixtures.py
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REFERENCE = "/root/reference"

# (kind, seed): "plain" all-valid pairs; "drop" pairs with 10% of residues
# dropped, their pred/native masks merged into one same-index mask (the
# drivers pass valid = pred_valid & native_valid and Lnorm =
# native_valid.sum()); "indel" pairs whose correspondence is shifted by an
# indel (synth.indel_pair), where the DP refinement changes the alignment
CASES = ([("plain", 70000 + k) for k in range(16)] + [("drop", 70100 + k) for k in range(12)]
         + [("indel", 71000 + k) for k in range(16)])
N_RANGE = (40, 160)
DP_ITERS = (0, 1, 2)


def case_pair(kind: str, seed: int):
    sys.path.insert(0, ROOT)
    from indialign_amd import synth
    if kind == "indel":
        return synth.indel_pair(seed, n_range=N_RANGE)
    p = synth.make_pair(seed, n_range=N_RANGE, drop_frac=0.1 if kind == "drop" else 0.0)
    v = p.valid.copy()
    return synth.Pair(p.pred, p.native, v, v.copy(), v.copy(), p.lnorm, p.seed, p.sigma)


def checksum(p) -> list[float]:
    return [float(np.abs(p.pred).sum() + 2.0 * np.abs(p.native).sum()), float(p.valid.sum()),
            float(len(p.valid)), float(p.lnorm)]


def main() -> None:
    sys.path.insert(0, REFERENCE)
    import torch
    from scoring._tmscore_gpu import runtime as rt
    from scoring._tmscore_gpu import search as S

    kinds, seeds, sums, scores = [], [], [], []
    t0 = time.time()
    for kind, seed in CASES:
        p = case_pair(kind, seed)
        pred = torch.from_numpy(p.pred)[None]
        native = torch.from_numpy(p.native)[None]
        valid = torch.from_numpy(p.valid.astype(bool))[None]
        L = torch.tensor([p.lnorm], dtype=torch.float64)
        d0, d0s, d8 = rt.d0_from_length(L, mol="rna", dtype=torch.float64)
        row = []
        for dp in DP_ITERS:
            s = S.tmscore_search(pred, native, valid, d0, d0s, d8, L, max_iter=20,
                                 use_fragment_search=True, max_mem_gb=4.0, dp_iter=dp)
            row.append(float(s[0]))
        kinds.append(kind)
        seeds.append(seed)
        sums.append(checksum(p))
        scores.append(row)
        print(f"{kind} {seed} N {len(p.valid)}: {row} ({time.time() - t0:.0f} s)",
              flush=True)
    np.savez(os.path.join(HERE, "torch_engine.npz"), kinds=np.array(kinds), seeds=np.array(seeds),
             checksum=np.array(sums), dp_iters=np.array(DP_ITERS), scores=np.array(scores),
             n_range=np.array(N_RANGE))


if __name__ == "__main__":
    main()
