"""indialign_amd: MI355X (gfx950) TM-score superposition search.

The product is the shared library ``libindialign.so`` in this package; it
exports the reference C ABI (indialign_c/api.h:38-46) unchanged, so callers
written against the reference (ctypes, README.md:17-65) load it as a drop-in.
This module is the Python mirror of that interface: the same three ctypes
structures (api.h:10-36) and thin wrappers that pack numpy arrays.

There is no CPU fallback: if the library is missing or no GPU is present the
calls raise.
"""
from __future__ import annotations

import ctypes
import os
import warnings

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libindialign.so")

c_double_p = ctypes.POINTER(ctypes.c_double)
c_u8_p = ctypes.POINTER(ctypes.c_uint8)


class NativeConfig(ctypes.Structure):
    """api.h:10-20"""
    _fields_ = [("backend_mode", ctypes.c_int), ("use_fragment_search", ctypes.c_int),
                ("max_iter", ctypes.c_int), ("dp_iter", ctypes.c_int),
                ("max_mem_gb", ctypes.c_double), ("request_cuda", ctypes.c_int),
                ("rescue_score_1", ctypes.c_double), ("rescue_score_2", ctypes.c_double),
                ("early_term_score", ctypes.c_double)]


class NativePairInput(ctypes.Structure):
    """api.h:22-30"""
    _fields_ = [("length", ctypes.c_int), ("pred_xyz", c_double_p), ("native_xyz", c_double_p),
                ("valid_mask", c_u8_p), ("pred_valid_mask", c_u8_p),
                ("native_valid_mask", c_u8_p), ("lnorm", ctypes.c_double)]


class NativePairResult(ctypes.Structure):
    """api.h:32-36"""
    _fields_ = [("score", ctypes.c_double), ("rotation", ctypes.c_double * 9),
                ("translation", ctypes.c_double * 3)]


class IndialignError(RuntimeError):
    pass


_lib = None


def load_library(path: str | None = None):
    """Load libindialign.so; raises if it was not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise IndialignError(f"{p} not found: build it with `python -m indialign_amd.build`")
    lib = ctypes.CDLL(p)
    lib.indialign_default_config.argtypes = [ctypes.POINTER(NativeConfig)]
    lib.indialign_default_config.restype = None
    lib.indialign_tmscore_search.argtypes = [ctypes.POINTER(NativePairInput),
                                             ctypes.POINTER(NativeConfig),
                                             ctypes.POINTER(NativePairResult)]
    lib.indialign_tmscore_search.restype = ctypes.c_int
    lib.indialign_tmscore_search_batch.argtypes = [ctypes.c_int, ctypes.POINTER(NativePairInput),
                                                   ctypes.POINTER(NativeConfig),
                                                   ctypes.POINTER(NativePairResult)]
    lib.indialign_tmscore_search_batch.restype = ctypes.c_int
    lib.indialign_search_ex.argtypes = [ctypes.c_int, ctypes.POINTER(NativePairInput),
                                        ctypes.POINTER(NativeConfig),
                                        ctypes.POINTER(NativePairResult), ctypes.c_int,
                                        ctypes.c_int]
    lib.indialign_search_ex.restype = ctypes.c_int
    lib.indialign_cuda_available.argtypes = []
    lib.indialign_cuda_available.restype = ctypes.c_int
    lib.indialign_batch_load.argtypes = [ctypes.c_int, ctypes.POINTER(NativePairInput)]
    lib.indialign_batch_load.restype = ctypes.c_int
    lib.indialign_batch_run.argtypes = [ctypes.c_int, ctypes.POINTER(NativeConfig),
                                        ctypes.POINTER(NativePairResult), ctypes.c_int,
                                        ctypes.c_int]
    lib.indialign_batch_run.restype = ctypes.c_int
    lib.indialign_release.argtypes = []
    lib.indialign_release.restype = ctypes.c_int
    lib.indialign_set_learned_options.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.indialign_set_learned_options.restype = ctypes.c_int
    lib.indialign_last_stats.argtypes = [c_double_p, ctypes.c_int]
    lib.indialign_last_stats.restype = ctypes.c_int
    c_float_p = ctypes.POINTER(ctypes.c_float)
    lib.indialign_last_stage_scores.argtypes = [c_double_p, ctypes.c_int]
    lib.indialign_last_stage_scores.restype = ctypes.c_int
    lib.indialign_set_ranker.argtypes = [c_float_p]
    lib.indialign_set_ranker.restype = ctypes.c_int
    lib.indialign_set_learned_budgets.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                  ctypes.c_int, ctypes.c_int, ctypes.c_double]
    lib.indialign_set_learned_budgets.restype = ctypes.c_int
    lib.indialign_set_launch_tuning.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_int]
    lib.indialign_set_launch_tuning.restype = ctypes.c_int
    lib.indialign_set_chain_buckets.argtypes = [ctypes.c_int]
    lib.indialign_set_chain_buckets.restype = ctypes.c_int
    lib.indialign_get_launch_tuning.argtypes = [ctypes.POINTER(ctypes.c_int)]
    lib.indialign_get_launch_tuning.restype = ctypes.c_int
    lib.indialign_collect_records.argtypes = [
        ctypes.c_int, ctypes.POINTER(NativePairInput), ctypes.POINTER(NativeConfig),
        ctypes.POINTER(NativePairResult), c_float_p, c_float_p, ctypes.POINTER(ctypes.c_int),
        ctypes.c_longlong, ctypes.POINTER(ctypes.c_longlong)]
    lib.indialign_collect_records.restype = ctypes.c_int
    lib.indialign_mlp_forward.argtypes = [c_float_p, c_float_p, ctypes.c_longlong, c_float_p]
    lib.indialign_mlp_forward.restype = ctypes.c_int
    lib.indialign_mlp_grad.argtypes = [c_float_p, c_float_p, c_float_p, ctypes.c_int, c_float_p,
                                       c_float_p]
    lib.indialign_mlp_grad.restype = ctypes.c_int
    lib.indialign_mlp_train.argtypes = [c_float_p, c_float_p, c_float_p, ctypes.c_longlong,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                        ctypes.c_uint, c_float_p]
    lib.indialign_mlp_train.restype = ctypes.c_int
    if path is None:
        _lib = lib
    return lib


def default_config(**overrides) -> NativeConfig:
    cfg = NativeConfig()
    load_library().indialign_default_config(ctypes.byref(cfg))
    for k, v in overrides.items():
        setattr(cfg, k, v)
    return cfg


def gpu_available() -> bool:
    return bool(load_library().indialign_cuda_available())


class _Packed:
    """Owns contiguous copies of the pair arrays behind NativePairInput structs."""

    def __init__(self, pairs):
        self.keep = []
        self.inps = (NativePairInput * max(1, len(pairs)))()
        for k, p in enumerate(pairs):
            pred, native, valid, pv, nv, lnorm = p
            pf = np.ascontiguousarray(np.asarray(pred, np.float64).reshape(-1))
            nf = np.ascontiguousarray(np.asarray(native, np.float64).reshape(-1))
            n = pf.size // 3
            v = np.ascontiguousarray(valid if valid is not None else np.ones(n), dtype=np.uint8)
            pvm = np.ascontiguousarray(pv if pv is not None else v, dtype=np.uint8)
            nvm = np.ascontiguousarray(nv if nv is not None else v, dtype=np.uint8)
            self.keep += [pf, nf, v, pvm, nvm]
            self.inps[k] = NativePairInput(
                n, pf.ctypes.data_as(c_double_p), nf.ctypes.data_as(c_double_p),
                v.ctypes.data_as(c_u8_p), pvm.ctypes.data_as(c_u8_p), nvm.ctypes.data_as(c_u8_p),
                float(lnorm if lnorm is not None else n))
        self.n = len(pairs)


def _as_tuple(p):
    if isinstance(p, tuple):
        return p + (None,) * (6 - len(p))
    return (p.pred, p.native, p.valid, p.pvalid, p.nvalid, p.lnorm)


def _results(res, n):
    score = np.array([res[k].score for k in range(n)])
    R = np.array([list(res[k].rotation) for k in range(n)]).reshape(n, 3, 3)
    t = np.array([list(res[k].translation) for k in range(n)])
    return score, R, t


def tmscore_search(pred, native, valid=None, pred_valid=None, native_valid=None, lnorm=None,
                   cfg: NativeConfig | None = None):
    """One pair through indialign_tmscore_search (indialign.cpp:284-303).

    Returns (score, R, t) with R as stored by the C ABI (out_d = sum_k P_k R[k,d] + t_d),
    i.e. ``aligned = pred @ R + t``; the reference's callers use R.T with pred @ R.T.
    """
    lib = load_library()
    cfg = cfg or default_config()
    pk = _Packed([(pred, native, valid, pred_valid, native_valid, lnorm)])
    res = NativePairResult()
    rc = lib.indialign_tmscore_search(ctypes.byref(pk.inps[0]), ctypes.byref(cfg),
                                      ctypes.byref(res))
    if rc != 0:
        raise IndialignError(f"indialign_tmscore_search returned {rc}")
    s, R, t = _results([res], 1)
    return float(s[0]), R[0], t[0]


METHOD_PLAIN, METHOD_FAST, METHOD_LEARNED = 0, 1, 2
# the search of the reference's separate PyTorch engine (same-index masks;
# indialign_search_ex method 3)
METHOD_TORCH_SEED = 3
# return code of a search with a pair whose compact native length exceeds
# the NW chain kernel's limit (that pair's score is NaN, the others are set)
TOO_LONG = -6
UNSUPPORTED = -7


def tmscore_search_batch(pairs, cfg: NativeConfig | None = None, serial_bank: int | None = None,
                         method: int = METHOD_PLAIN):
    """Many pairs in one batched call.  `pairs` holds synth.Pair objects or
    (pred, native, valid, pred_valid, native_valid, lnorm) tuples.  Returns
    (scores[n], R[n,3,3], t[n,3]).  serial_bank None uses the batch entry's
    convention (early termination applies); 0/1 selects it explicitly.
    method: METHOD_PLAIN (FP64, reference arithmetic), METHOD_FAST or
    METHOD_LEARNED (needs set_ranker)."""
    lib = load_library()
    cfg = cfg or default_config()
    pk = _Packed([_as_tuple(p) for p in pairs])
    res = (NativePairResult * max(1, pk.n))()
    if serial_bank is None and method == METHOD_PLAIN:
        rc = lib.indialign_tmscore_search_batch(pk.n, pk.inps, ctypes.byref(cfg), res)
    else:
        sb = 1 if serial_bank is None else int(serial_bank)
        rc = lib.indialign_search_ex(pk.n, pk.inps, ctypes.byref(cfg), res, sb, int(method))
    if rc == TOO_LONG:
        warnings.warn("some pairs have more native-valid residues than the NW chain kernel "
                      "holds; their scores are NaN", RuntimeWarning)
    elif rc == UNSUPPORTED:
        raise IndialignError("method 3 (the PyTorch engine's search) takes same-index masks only")
    elif rc != 0:
        raise IndialignError(f"batch search returned {rc}")
    return _results(res, pk.n)


_resident_n = 0


def batch_load(pairs) -> int:
    """Copy a batch into device memory once (indialign_batch_load)."""
    global _resident_n
    pk = _Packed([_as_tuple(p) for p in pairs])
    rc = load_library().indialign_batch_load(pk.n, pk.inps)
    if rc != 0:
        raise IndialignError(f"indialign_batch_load returned {rc}")
    _resident_n = pk.n
    return pk.n


def batch_run(cfg: NativeConfig | None = None, serial_bank: int = 1, method: int = METHOD_PLAIN):
    """Search the resident batch (indialign_batch_run); (scores, R, t).
    Raises if another search replaced the resident batch since batch_load."""
    cfg = cfg or default_config()
    res = (NativePairResult * max(1, _resident_n))()
    rc = load_library().indialign_batch_run(_resident_n, ctypes.byref(cfg), res, int(serial_bank),
                                            int(method))
    if rc == -5:
        raise IndialignError("indialign_batch_run: no resident batch of this size (another "
                             "search replaced it; call batch_load again)")
    if rc != 0:
        raise IndialignError(f"indialign_batch_run returned {rc}")
    return _results(res, _resident_n)


def release() -> None:
    """Free the library's device memory (indialign_release)."""
    load_library().indialign_release()


def set_learned_options(reduce_rescored: bool = False, fallback_top_m: int = 512,
                        fallback_top_m_first: int = 64) -> None:
    """Method 2's options (indialign_set_learned_options): the pass reduction
    (keep the work item with the best same-index TM-score, True, or the plain
    search's best chain score, False) and the fallback budgets of pairs whose
    predicted best stays below the pass's rescue threshold (0: off)."""
    rc = load_library().indialign_set_learned_options(int(bool(reduce_rescored)),
                                                      int(fallback_top_m),
                                                      int(fallback_top_m_first))
    if rc != 0:
        raise IndialignError(f"indialign_set_learned_options returned {rc}")


def last_stats() -> dict:
    """Work counters of the last call (see indialign_last_stats in the header)."""
    buf = (ctypes.c_double * 12)()
    load_library().indialign_last_stats(buf, 12)
    keys = ["wall_ms", "candidates", "nw_cells", "chains", "launches", "chain_ms", "bank_ms",
            "fallbacks", "fallback_budget_passes", "refine_passes", "bank_ops", "chain_ops"]
    return {k: buf[i] for i, k in enumerate(keys)}


STAGES = ["seed_bank", "ss", "gapless", "local_fragment", "threading_dp", "dense_local_fragment",
          "topk_seed", "dp_refine", "offset_dp", "weighted_refine"]


def last_stage_scores(n: int) -> np.ndarray:
    """Running best score of each pair after each pipeline stage of the last
    search, shape (n, 10), NaN where a stage did not run (order: STAGES)."""
    out = np.full((max(n, 1), len(STAGES)), np.nan)
    m = load_library().indialign_last_stage_scores(out.ctypes.data_as(c_double_p), n)
    return out[:m]


# ---- learned ranking of the local-fragment work items -------------------------
NFEAT = 96
NPARAM = 90881
_f32p = ctypes.POINTER(ctypes.c_float)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def set_ranker(params, top_m: int = 64, fallback_margin: float = -1.0,
               top_offsets: int = 16, top_seeds: int = 0, top_m_first: int = 16,
               top_offset_banks: int = 16) -> None:
    """Install the ranker (NPARAM float32 values) used by METHOD_LEARNED, or
    remove it with params=None, and set the method's budgets
    (indialign_set_learned_budgets): top_m dense and top_m_first first-pass
    local-fragment work items per pair, top_offsets threading-DP offsets per
    pair that start NW chains, top_offset_banks threading-DP offsets per pair
    that run their seed bank (0 = all), top_seeds seeds refined per SS-stage
    and threading bank (0 = all), and the fallback margin (< 0: off)."""
    lib = load_library()
    if params is None:
        lib.indialign_set_ranker(None)
        return
    p = _f32(params).reshape(-1)
    if p.size != NPARAM:
        raise ValueError(f"ranker needs {NPARAM} parameters, got {p.size}")
    rc = lib.indialign_set_learned_budgets(int(top_m), int(top_m_first), int(top_offsets),
                                           int(top_offset_banks), int(top_seeds),
                                           float(fallback_margin))
    if rc != 0:
        raise IndialignError(f"indialign_set_learned_budgets returned {rc}")
    rc = lib.indialign_set_ranker(p.ctypes.data_as(_f32p))
    if rc != 0:
        raise IndialignError(f"indialign_set_ranker returned {rc}")


def set_launch_tuning(chain_waves: int = 3, chain_slots_per_cu: int = 0,
                      chain_trace: int = 0, bank_pass_cap: int = 5) -> None:
    """Launch shape (indialign_set_launch_tuning): the NW chain kernel's
    wavefronts per SIMD (2 or 3), persistent slots per CU (0 = all
    resident) and trace placement (0 auto, 1 LDS when it fits, 2 global),
    and the seed banks' pass cap (0 = one launch; else candidates still
    refining after that many passes continue in a compacted launch).
    Results do not depend on it."""
    rc = load_library().indialign_set_launch_tuning(int(chain_waves), int(chain_slots_per_cu),
                                                    int(chain_trace), int(bank_pass_cap))
    if rc != 0:
        raise IndialignError(f"indialign_set_launch_tuning({chain_waves}, {chain_slots_per_cu}, "
                             f"{chain_trace}, {bank_pass_cap}) returned {rc}")


def set_chain_buckets(n_classes: int = 1) -> None:
    """Length classes of the NW chain launches (indialign_set_chain_buckets);
    results do not depend on it."""
    rc = load_library().indialign_set_chain_buckets(int(n_classes))
    if rc != 0:
        raise IndialignError(f"indialign_set_chain_buckets({n_classes}) returned {rc}")


def get_launch_tuning() -> tuple[int, int, int, int]:
    """Current (chain_waves, chain_slots_per_cu, chain_trace, bank_pass_cap)."""
    out = (ctypes.c_int * 4)()
    load_library().indialign_get_launch_tuning(out)
    return int(out[0]), int(out[1]), int(out[2]), int(out[3])


def collect_records(pairs, cap: int, cfg: NativeConfig | None = None):
    """Plain search that also returns every local-fragment work item's
    descriptor X (rows x 96), refined score y and pair index.  Returns
    (scores, R, t, X, y, pair_idx)."""
    lib = load_library()
    cfg = cfg or default_config()
    pk = _Packed([_as_tuple(p) for p in pairs])
    res = (NativePairResult * max(1, pk.n))()
    X = np.empty((max(cap, 1), NFEAT), np.float32)
    y = np.empty(max(cap, 1), np.float32)
    pi = np.empty(max(cap, 1), np.int32)
    nrec = ctypes.c_longlong(0)
    rc = lib.indialign_collect_records(pk.n, pk.inps, ctypes.byref(cfg), res,
                                       X.ctypes.data_as(_f32p), y.ctypes.data_as(_f32p),
                                       pi.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                       int(cap), ctypes.byref(nrec))
    if rc != 0:
        raise IndialignError(f"indialign_collect_records returned {rc}")
    m = nrec.value
    s, R, t = _results(res, pk.n)
    return s, R, t, X[:m].copy(), y[:m].copy(), pi[:m].copy()


def mlp_forward(params, X) -> np.ndarray:
    """Ranker forward pass on the GPU."""
    p, Xf = _f32(params).reshape(-1), _f32(X).reshape(-1, NFEAT)
    y = np.empty(Xf.shape[0], np.float32)
    rc = load_library().indialign_mlp_forward(p.ctypes.data_as(_f32p), Xf.ctypes.data_as(_f32p),
                                              Xf.shape[0], y.ctypes.data_as(_f32p))
    if rc != 0:
        raise IndialignError(f"indialign_mlp_forward returned {rc}")
    return y


def mlp_grad(params, X, t):
    """(grad[NPARAM], loss) of the MSE from the GPU backward kernels."""
    p, Xf, tf = _f32(params).reshape(-1), _f32(X).reshape(-1, NFEAT), _f32(t).reshape(-1)
    g = np.empty(NPARAM, np.float32)
    loss = np.zeros(1, np.float32)
    rc = load_library().indialign_mlp_grad(p.ctypes.data_as(_f32p), Xf.ctypes.data_as(_f32p),
                                           tf.ctypes.data_as(_f32p), Xf.shape[0],
                                           g.ctypes.data_as(_f32p), loss.ctypes.data_as(_f32p))
    if rc != 0:
        raise IndialignError(f"indialign_mlp_grad returned {rc}")
    return g, float(loss[0])


def mlp_train(params, X, t, epochs: int, batch: int = 4096, lr: float = 1e-3, seed: int = 0):
    """Adam training on the GPU; returns (new params, per-epoch loss)."""
    p = _f32(params).reshape(-1).copy()
    Xf, tf = _f32(X).reshape(-1, NFEAT), _f32(t).reshape(-1)
    el = np.zeros(max(epochs, 1), np.float32)
    rc = load_library().indialign_mlp_train(p.ctypes.data_as(_f32p), Xf.ctypes.data_as(_f32p),
                                            tf.ctypes.data_as(_f32p), Xf.shape[0], int(epochs),
                                            int(batch), float(lr), int(seed),
                                            el.ctypes.data_as(_f32p))
    if rc != 0:
        raise IndialignError(f"indialign_mlp_train returned {rc}")
    return p, el[:epochs]


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform initial ranker parameters (biases zero)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    w1, w2 = 256 * 96, 256 * 256
    p[:w1] = rng.uniform(-1, 1, w1) * np.sqrt(6.0 / 96)
    o = w1 + 256
    p[o:o + w2] = rng.uniform(-1, 1, w2) * np.sqrt(6.0 / 256)
    o += w2 + 256
    # small output weights: the untrained ranker is close to its skip input
    # (the refined same-index TM-score of each work item)
    p[o:o + 256] = rng.uniform(-1, 1, 256) * 0.01
    return p


DEFAULT_RANKER = os.path.join(HERE, "ranker_params.npy")


def load_default_ranker(top_m: int = 64, fallback_margin: float = -1.0,
                        top_offsets: int = 16, top_seeds: int = 0, top_m_first: int = 16,
                        top_offset_banks: int = 16, path: str | None = None) -> np.ndarray:
    """Install the packaged trained ranker (train_ranker.py) for METHOD_LEARNED."""
    p = np.load(path or DEFAULT_RANKER, allow_pickle=False)
    set_ranker(p, top_m, fallback_margin, top_offsets, top_seeds, top_m_first, top_offset_banks)
    return p
