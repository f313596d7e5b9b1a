"""torch-tensor entry for the scoring drivers.

`tmscore_search` has the signature and argument meaning of the reference's
batched torch search (scoring/_tmscore_gpu/search.py:492-541): padded
(B, N, 3) coordinate tensors, a (B, N) same-index validity mask, per-pair
d0 / d0_search / score_d8 / Lnorm tensors, the refinement and DP knobs, and
optional separate pred/native masks.  It returns the (B,) best TM-scores on
pred's device and dtype, like the reference function.

Two semantics are selectable.  semantics="c" (default) runs this library's
C-ABI search (the reference's indialign_tmscore_search semantics, method 0
by default, FP64): its stage list, seed schedule and refinement differ from
the torch engine's, so scores differ from that engine's on the same inputs.
semantics="torch" runs the torch engine's own search with same-index masks
(scoring/_tmscore_gpu/search.py:492-761; library method 3): window seeds
without anchor seeds, cutoffs d0_search x {2, 1.5, 1, .75, .5, .25} only,
exactly max_iter refinement passes per candidate with no convergence break,
candidates ranked by the TM-score with the score_d8 cut (the engine's
_SEARCH_SELECTION_USES_SCORE_D8 default), the top-k distance seed; with
dp_iter > 0 its DP refinement (dp.py:186-252: NW over all N x N residue
pairs with masked scores, no gap penalty and up before left on ties, the
detailed search over the list, dp_iter iterations) from the best seed; the
reported score is the maximum of the DP score, the TM-scores without the cut
of the DP and seed transforms, and the ranking score (search.py:690-757).
Each pair is searched as the engine searches a batch of one.  The engine's
cross-index rescue (separate pred/native masks) is not rebuilt:
semantics="torch" with separate masks raises NotImplementedError.  The C search derives d0, d0_search and score_d8 from Lnorm with the
closed forms of core.h:12-25, which are also the reference's
d0_from_length (scoring/_tmscore_gpu/runtime.py:91-134); the given tensors
are checked against them and a different rule raises ValueError, since the
C ABI cannot take other values.  Padded rows must be marked invalid; like the
C search, window seeds are laid over index positions, so padding can change
which seeds run.

The coordinates cross to the host (the C ABI takes host arrays) and the
scores come back as a tensor; there is no CPU search anywhere.
"""
from __future__ import annotations

import math

import numpy as np

from . import METHOD_PLAIN, METHOD_TORCH_SEED, default_config, tmscore_search_batch

# Python reference MAX_ITER default (scoring/_tmscore_gpu/runtime.py: _MAX_ITER)
DEFAULT_MAX_ITER = 20


def d0_params(L: float) -> tuple[float, float, float]:
    """(d0, d0_search, score_d8) for an effective length L (core.h:12-25)."""
    Lc = max(float(L), 1.0)
    if Lc >= 30.0:
        d0 = 0.6 * math.sqrt(max(Lc - 0.5, 1e-12)) - 2.5
    elif Lc > 23.0:
        d0 = 0.7
    elif Lc > 19.0:
        d0 = 0.6
    elif Lc > 15.0:
        d0 = 0.5
    elif Lc > 11.0:
        d0 = 0.4
    else:
        d0 = 0.3
    d0 = max(d0, 0.3)
    d0_search = min(max(1.24 * max(Lc - 15.0, 0.0) ** (1.0 / 3.0) - 1.8, 4.5), 8.0)
    score_d8 = 1.5 * Lc ** 0.3 + 3.5
    return d0, d0_search, score_d8


def _host(x, dtype):
    return np.ascontiguousarray(x.detach().to("cpu").numpy(), dtype=dtype)


def tmscore_search(pred, native, valid, d0, d0_search, score_d8, Lnorm, max_iter=None,
                   use_fragment_search: bool = True, max_mem_gb: float = 20.0, dp_iter: int = 0,
                   pred_valid=None, native_valid=None, method: int = METHOD_PLAIN,
                   rtol: float = 1e-5, semantics: str = "c"):
    """Best TM-score per pair of a padded batch (see the module docstring).

    pred, native: (B, N, 3); valid: (B, N) bool, pred/native same-index
    validity; d0, d0_search, score_d8, Lnorm: (B,); max_iter: refinement
    passes per seed (None = 20); dp_iter: DP refinement iterations;
    pred_valid / native_valid: (B, N) separate masks (default: valid).
    max_mem_gb is accepted and unused (device memory is sized by the batch).
    method: 0 plain FP64 (default), 1 fast, 2 learned (needs a ranker).
    rtol: relative tolerance of the d0 / d0_search / score_d8 check (the
    reference computes them in float32 by default).
    semantics: "c" (the C path, `method` applies) or "torch" (the torch
    engine's search, FP64; no separate masks)."""
    import torch

    if semantics not in ("c", "torch"):
        raise ValueError('semantics must be "c" or "torch"')
    if semantics == "torch":
        for m in (pred_valid, native_valid):
            if m is not None and not bool(torch.equal(m.bool(), valid.bool())):
                raise NotImplementedError("semantics='torch': the torch engine's cross-index "
                                          "rescue (separate masks) is not rebuilt")
        method = METHOD_TORCH_SEED

    if max_iter is None:
        max_iter = DEFAULT_MAX_ITER
    if pred.ndim != 3 or native.ndim != 3 or pred.shape != native.shape or pred.shape[2] != 3:
        raise ValueError("pred and native must be shape (B, N, 3) and match.")
    if valid.shape != pred.shape[:2]:
        raise ValueError("valid mask must be shape (B, N).")
    B, N, _ = pred.shape
    for name, m in (("pred_valid", pred_valid), ("native_valid", native_valid)):
        if m is not None and m.shape != pred.shape[:2]:
            raise ValueError(f"{name} must be shape (B, N).")
    for name, v in (("d0", d0), ("d0_search", d0_search), ("score_d8", score_d8), ("Lnorm", Lnorm)):
        if tuple(v.shape) != (B,):
            raise ValueError(f"{name} must be shape (B,).")
    if B == 0:
        return torch.zeros((0,), dtype=pred.dtype, device=pred.device)

    P = _host(pred, np.float64)
    Q = _host(native, np.float64)
    vm = _host(valid, np.uint8)
    pv = vm if pred_valid is None else _host(pred_valid, np.uint8)
    nv = vm if native_valid is None else _host(native_valid, np.uint8)
    L = _host(Lnorm, np.float64)
    given = np.stack([_host(d0, np.float64), _host(d0_search, np.float64),
                      _host(score_d8, np.float64)], axis=1)
    for b in range(B):
        want = np.array(d0_params(L[b]))
        if not np.allclose(given[b], want, rtol=rtol, atol=0.0):
            raise ValueError(
                f"pair {b}: (d0, d0_search, score_d8) = {tuple(given[b])} is not the RNA rule of "
                f"Lnorm = {L[b]} {tuple(want)}; the C search derives them from Lnorm")
    pairs = [(P[b], Q[b], vm[b], pv[b], nv[b], float(L[b])) for b in range(B)]
    cfg = default_config(max_iter=int(max_iter), use_fragment_search=int(bool(use_fragment_search)),
                         max_mem_gb=float(max_mem_gb), dp_iter=int(dp_iter))
    scores, _, _ = tmscore_search_batch(pairs, cfg, method=int(method))
    return torch.as_tensor(scores, dtype=pred.dtype, device=pred.device)


def score_pair_batches(pred_coords, pred_valid, native_coords, native_valid, pair_chunk: int = 256,
                       max_iter=None, use_fragment_search: bool = True, max_mem_gb: float = 20.0,
                       dp_iter: int = 1, method: int = METHOD_PLAIN) -> float:
    """Best TM-score over every (pred frame, native frame) pair, with the
    argument meaning of the reference's multi-frame scorer
    (scoring/_local_validation_gpu/scoring.py:15-70): pred_coords (P, N, 3)
    with pred_valid (P, N), native_coords (F, N, 3) with native_valid (F, N);
    a frame pair is scored on the same-index mask pred_valid & native_valid
    with Lnorm = its native frame's valid count, and pairs with Lnorm <= 2
    score 0.  Defaults follow that driver's (dp_iter 1,
    scoring/_local_validation_gpu/runtime.py:13-16).  All P x F pairs go to
    the device as one batch (pair_chunk is accepted for signature
    compatibility; device memory holds the batch).  Returns a float, 0.0 for
    no frames."""
    import torch

    P, F = int(pred_coords.shape[0]), int(native_coords.shape[0])
    if P == 0 or F == 0:
        return 0.0
    if pred_coords.shape[1:] != native_coords.shape[1:] or pred_coords.shape[2] != 3:
        raise ValueError("pred and native frames must be shape (P, N, 3) and (F, N, 3).")
    if pred_valid.shape != pred_coords.shape[:2] or native_valid.shape != native_coords.shape[:2]:
        raise ValueError("valid masks must be shape (P, N) and (F, N).")
    pidx = torch.arange(P * F, device=pred_coords.device) // F
    nidx = torch.arange(P * F, device=pred_coords.device) % F
    p = pred_coords.index_select(0, pidx)
    n = native_coords.index_select(0, nidx)
    nv = native_valid.index_select(0, nidx).bool()
    v = pred_valid.index_select(0, pidx).bool() & nv
    L = nv.sum(dim=1).to(torch.float64)
    good = L > 2.0
    if not bool(good.any()):
        return 0.0
    Lg = L[good]
    d = torch.tensor([d0_params(float(x)) for x in Lg], dtype=torch.float64)
    s = tmscore_search(p[good], n[good], v[good], d[:, 0], d[:, 1], d[:, 2], Lg, max_iter=max_iter,
                       use_fragment_search=use_fragment_search, max_mem_gb=max_mem_gb,
                       dp_iter=dp_iter, method=method)
    return max(0.0, float(s.max()))
