"""Multimer scoring by chain permutation over the batched GPU search.

A multimer target has several chains, some of them copies of the same
molecule.  A prediction may name its copies in a different order, so the
score of a (pred frame, native frame) pair is the best over candidate chain
permutations of the prediction.  The candidates follow the reference's
scorer (scoring/_local_validation_gpu/multimer.py:12-298):

- residues are grouped by (chain, copy) in first-seen order (:12-26);
- per chain type with several copies, a cost matrix of per-copy-pair Kabsch
  mean squared deviations (:199-218); if every row is flat (max row range
  < 1e-4: the copies are interchangeable) all permutations are candidates
  when there are at most 720 of them (:220-231), else ICP on the copies'
  centroids from a distance-profile seed, the identity and 8 n - 2 random
  restarts (:50-130); otherwise the Hungarian assignment alone (:236-239);
- the cross product over chain types, capped at 720 by collapsing the
  largest candidate list to the identity (:241-253);
- every candidate of every frame pair is scored in one batched search
  (valid = pred-valid & native-valid after the permutation, Lnorm = native
  valid count, Lnorm <= 2 scores 0) and the maximum is returned (:256-298).

The search itself is this library's (the C path's semantics); the
reference's torch engine scores candidates with its own semantics.
"""
from __future__ import annotations

import math
from collections import defaultdict
from itertools import permutations, product

import numpy as np

import indialign_amd as ia

DEGEN_EPS = 1e-4
MAX_CANDIDATES = 720


def group_indices(chains, copies):
    """(blocks, chain_labels): residue index arrays of each (chain, copy)
    group in first-seen order, and each group's chain name (:12-26)."""
    by_group: dict = {}
    for idx, g in enumerate(zip((str(c) for c in chains), (int(c) for c in copies))):
        by_group.setdefault(g, []).append(idx)
    blocks = [np.asarray(v, dtype=np.int64) for v in by_group.values()]
    return blocks, [k[0] for k in by_group.keys()]


def group_targets_by_length(targets, lengths, max_ratio: float = 2.0):
    """Targets sorted by length, cut into groups whose max/min <= max_ratio
    (:29-47)."""
    items = sorted(targets, key=lambda t: lengths[t])
    groups, group, gmin = [], [], 0
    for t in items:
        n = lengths[t]
        if group and n > gmin * max_ratio:
            groups.append(group)
            group = []
        if not group:
            gmin = n
        group.append(t)
    if group:
        groups.append(group)
    return groups


def _kabsch(P, Q):
    """Proper rotation R and t minimising sum |R p + t - q|^2 (SVD)."""
    cp, cq = P.mean(axis=0), Q.mean(axis=0)
    H = (P - cp).T @ (Q - cq)
    U, _, Vt = np.linalg.svd(H)
    d = np.sign(np.linalg.det(Vt.T @ U.T))
    R = Vt.T @ np.diag([1.0, 1.0, d]) @ U.T
    return R, cq - cp @ R.T


def _assign(cost):
    from scipy.optimize import linear_sum_assignment
    return linear_sum_assignment(cost)[1]


def icp_candidates(p_full, n_full, pv, nv, blocks, gis, max_rounds: int = 20, seed: int = 0):
    """Candidate permutations of the copies `gis` from ICP on their
    centroids (:50-130): seeds are the rotation-invariant distance-profile
    assignment, the identity and 8 n - 2 random permutations; each round is
    a Kabsch of the centroids under the current matching, then a Hungarian
    assignment of the rotated centroids.  Returned as reorder indices."""
    n = len(gis)
    pc = np.zeros((n, 3))
    nc = np.zeros((n, 3))
    for ci, gi in enumerate(gis):
        b = blocks[gi]
        if pv[b].any():
            pc[ci] = p_full[b][pv[b].astype(bool)].mean(axis=0)
        if nv[b].any():
            nc[ci] = n_full[b][nv[b].astype(bool)].mean(axis=0)
    pc -= pc.mean(axis=0)
    nc -= nc.mean(axis=0)
    dist = lambda A, B: np.linalg.norm(A[:, None, :] - B[None, :, :], axis=2)
    prof_p = np.sort(dist(pc, pc), axis=1)[:, 1:]
    prof_n = np.sort(dist(nc, nc), axis=1)[:, 1:]
    seed_dp = _assign(dist(prof_p, prof_n))
    rng = np.random.RandomState(seed)
    seeds = [seed_dp, np.arange(n)] + [rng.permutation(n) for _ in range(max(0, 8 * n - 2))]
    found = set()
    for s in seeds:
        cur = np.array(s)
        visited = set()
        for _ in range(max_rounds):
            key = tuple(int(x) for x in cur)
            if key in visited:
                break
            visited.add(key)
            H = pc.T @ nc[cur]
            U, _, Vt = np.linalg.svd(H)
            d = np.sign(np.linalg.det(Vt.T @ U.T))
            R = Vt.T @ np.diag([1.0, 1.0, d]) @ U.T
            cur = _assign(dist(pc @ R.T, nc))
        found.add(tuple(int(x) for x in cur))
    found.add(tuple(range(n)))
    # ICP maps pred copies to native copies; the caller reorders, so invert
    return [tuple(int(x) for x in np.argsort(p)) for p in sorted(found)]


def frame_candidates(p_full, n_full, pv, nv, blocks, labels):
    """Full-length residue orders of the prediction to score for one
    (pred frame, native frame) pair."""
    G = len(blocks)
    by_type = defaultdict(list)
    for gi, ch in enumerate(labels):
        by_type[ch].append(gi)
    cands = {}
    for ch, gis in by_type.items():
        n = len(gis)
        if n <= 1:
            cands[ch] = [(0,)]
            continue
        cost = np.zeros((n, n))
        for ci, gi in enumerate(gis):
            for cj, gj in enumerate(gis):
                bi, bj = blocks[gi], blocks[gj]
                m = (pv[bi] & nv[bj]).astype(bool)
                if m.sum() < 3:
                    cost[ci, cj] = 1e12
                    continue
                R, t = _kabsch(p_full[bi][m], n_full[bj][m])
                d2 = np.sum((p_full[bi][m] @ R.T + t - n_full[bj][m]) ** 2, axis=1)
                cost[ci, cj] = d2.sum() / m.sum()
        span = 0.0
        for r in range(n):
            row = cost[r][cost[r] < 1e11]
            if len(row) > 1:
                span = max(span, row.max() - row.min())
        if span < DEGEN_EPS and math.factorial(n) <= MAX_CANDIDATES:
            cands[ch] = list(permutations(range(n)))
        elif span < DEGEN_EPS:
            cands[ch] = icp_candidates(p_full, n_full, pv, nv, blocks, gis)
        else:
            cands[ch] = [tuple(int(x) for x in np.argsort(_assign(cost)))]
    order = sorted(cands)
    total = math.prod(len(cands[ch]) for ch in order)
    while total > MAX_CANDIDATES:
        big = max(order, key=lambda ch: len(cands[ch]))
        cands[big] = [tuple(range(len(by_type[big])))]
        total = math.prod(len(cands[ch]) for ch in order)
    out = []
    for combo in product(*[cands[ch] for ch in order]):
        perm = list(range(G))
        for k, ch in enumerate(order):
            gis = by_type[ch]
            for ci, gi in enumerate(gis):
                perm[gi] = gis[combo[k][ci]]
        out.append(np.concatenate([blocks[perm[j]] for j in range(G)]))
    return out


def candidate_pairs(pred_coords, pred_valid, native_coords, native_valid, blocks, labels):
    """Every (permuted pred frame, native frame) candidate as a search input
    tuple (pred, native, valid, pred_valid, native_valid, Lnorm), with the
    (pred frame, native frame) index of each."""
    pred_coords = np.asarray(pred_coords, np.float64)
    native_coords = np.asarray(native_coords, np.float64)
    pred_valid = np.asarray(pred_valid).astype(np.uint8)
    native_valid = np.asarray(native_valid).astype(np.uint8)
    tuples, owner = [], []
    for pi in range(pred_coords.shape[0]):
        for ni in range(native_coords.shape[0]):
            for idx in frame_candidates(pred_coords[pi], native_coords[ni], pred_valid[pi],
                                        native_valid[ni], blocks, labels):
                p = pred_coords[pi][idx]
                v = (pred_valid[pi][idx] & native_valid[ni]).astype(np.uint8)
                tuples.append((p, native_coords[ni], v, v, v, float(native_valid[ni].sum())))
                owner.append((pi, ni))
    return tuples, owner


def score_multimer(pred_coords, pred_valid, native_coords, native_valid, blocks, labels,
                   cfg=None, method: int = ia.METHOD_PLAIN, search=None) -> float:
    """Best TM-score over every candidate chain permutation of every
    (pred frame, native frame) pair; `search` (default: this library's
    batched search) maps a list of input tuples to (scores, R, t)."""
    tuples, _ = candidate_pairs(pred_coords, pred_valid, native_coords, native_valid, blocks,
                                labels)
    if not tuples:
        return 0.0
    if search is None:
        search = lambda tup: ia.tmscore_search_batch(tup, cfg=cfg, method=method)
    scores, _, _ = search(tuples)
    return float(np.max(scores))
