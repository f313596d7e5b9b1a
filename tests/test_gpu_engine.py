"""Library state across calls on the GPU: the resident batch, re-creating
the engine, and the reference's thread switch of the single-pair entry."""
import ctypes
import os

import numpy as np
import pytest

import indialign_amd as ia
import libs
from indialign_amd import synth

pytestmark = pytest.mark.gpu


def _gpu_or_skip():
    if not ia.gpu_available():
        pytest.fail("no GPU visible to the HIP runtime (GPU tests must run on the MI355X)")


def test_resident_batch_is_invalidated_by_other_searches():
    """batch_run after another search (which replaces the device batch) or
    with a different pair count is an error, not a silent overrun."""
    _gpu_or_skip()
    lib = ia.load_library()
    pairs = synth.short_pairs(6, first_seed=700)
    ia.batch_load(pairs)
    s0, _, _ = ia.batch_run()
    cfg = ia.default_config()
    res = (ia.NativePairResult * 16)()
    assert lib.indialign_batch_run(5, ctypes.byref(cfg), res, 1, 0) == -5
    assert lib.indialign_batch_run(16, ctypes.byref(cfg), res, 1, 0) == -5
    ia.tmscore_search_batch(synth.short_pairs(3, first_seed=800))
    with pytest.raises(ia.IndialignError):
        ia.batch_run()
    ia.batch_load(pairs)
    s1, _, _ = ia.batch_run()
    np.testing.assert_array_equal(s0, s1)


def test_engine_recreated_keeps_ranker_and_budgets():
    """After the engine is dropped (indialign_release; a device error takes
    the same path) the next engine gets the installed ranker and budgets: with
    budgets above every pair's item counts method 2 still equals method 1."""
    _gpu_or_skip()
    pairs = synth.short_pairs(12, first_seed=501)
    sf, Rf, tf = ia.tmscore_search_batch(pairs, serial_bank=1, method=ia.METHOD_FAST)
    ia.set_ranker(ia.init_params(0), top_m=1 << 30, top_offsets=1 << 30, top_seeds=0,
                  top_m_first=1 << 30, top_offset_banks=0)
    ia.set_learned_options(reduce_rescored=False)
    try:
        for _ in range(2):
            ia.release()
            sl, Rl, tl = ia.tmscore_search_batch(pairs, serial_bank=1, method=ia.METHOD_LEARNED)
            np.testing.assert_array_equal(sl, sf)
            np.testing.assert_array_equal(Rl, Rf)
            np.testing.assert_array_equal(tl, tf)
    finally:
        ia.set_ranker(None)
        ia.set_learned_options(reduce_rescored=False)


def test_single_pair_entry_thread_switch():
    """indialign_tmscore_search picks the main seed bank's convention as the
    reference does (search.h:14-24): the serial path (early termination at
    early_term_score) with one OpenMP thread or fewer than 4 candidates per
    thread, the parallel path (none) otherwise.  With early_term_score 0.5
    this pair's two conventions give different results."""
    _gpu_or_skip()
    p = synth.make_pair(52003, n=120, sigma=2.0)
    ocfg = libs.oracle_config(early_term_score=0.5)
    ser = libs.oracle_search([p], cfg=ocfg, serial_bank=1)[0]
    par = libs.oracle_search([p], cfg=ocfg, serial_bank=0)[0]
    assert ser[12] != par[12]
    cfg = ia.default_config(early_term_score=0.5)
    old = os.environ.get("OMP_NUM_THREADS")
    try:
        for threads, want in (("1", ser), ("8", par), ("100000", ser)):
            os.environ["OMP_NUM_THREADS"] = threads
            sc, R, t = ia.tmscore_search(p.pred, p.native, cfg=cfg)
            got = np.r_[R.reshape(-1), t, sc]
            print(f"OMP_NUM_THREADS={threads}: {sc:.15f} (serial {ser[12]:.15f}, parallel "
                  f"{par[12]:.15f})", flush=True)
            assert np.abs(got - want).max() <= 1e-12
    finally:
        if old is None:
            os.environ.pop("OMP_NUM_THREADS", None)
        else:
            os.environ["OMP_NUM_THREADS"] = old


def _long_native_pair(seed, n_native_valid, n_pred_valid=300):
    """A pair of N = n_native_valid residues, every native residue valid
    (compact native length = n_native_valid) and n_pred_valid pred residues
    valid, Lnorm = the same-index valid count, low noise: the pipeline's NW
    chains run at n_pred_valid x n_native_valid, and the score stays above the
    rescue thresholds, so the CPU oracle finishes in seconds."""
    p = synth.make_pair(seed, n=n_native_valid, sigma=1.0)
    rng = np.random.default_rng(seed)
    pv = np.zeros(p.n, np.uint8)
    start = int(rng.integers(0, p.n - n_pred_valid))
    pv[start:start + n_pred_valid] = 1
    p.pvalid = pv
    p.nvalid = np.ones(p.n, np.uint8)
    p.valid = (pv & p.nvalid).astype(np.uint8)
    p.lnorm = float(p.valid.sum())
    return p


def test_native_length_limit():
    """The NW chain kernel holds one FP64 (H, G) pair per compact native
    residue in LDS: 10239 native-valid residues are searched (checked against
    the oracle), 10240 give NaN and return code -6 while the other pairs of
    the batch are searched normally."""
    _gpu_or_skip()
    at = _long_native_pair(60500, 10239)
    over = _long_native_pair(60501, 10240)
    small = synth.make_pair(60502)
    lib = ia.load_library()
    pk = ia._Packed([ia._as_tuple(p) for p in (small, at, over)])
    res = (ia.NativePairResult * 3)()
    cfg = ia.default_config()
    rc = lib.indialign_tmscore_search_batch(3, pk.inps, ctypes.byref(cfg), res)
    assert rc == ia.TOO_LONG
    s, R, t = ia._results(res, 3)
    assert np.isnan(s[2]) and np.array_equal(R[2], np.eye(3)) and not t[2].any()
    ref = libs.oracle_search([small, at], serial_bank=1)
    for k in range(2):
        print(f"pair {k}: N={len(s) and [small, at][k].n} gpu {s[k]:.15f} oracle {ref[k, 12]:.15f}",
              flush=True)
        assert abs(s[k] - ref[k, 12]) <= 1e-6
        assert np.abs(R[k].reshape(-1) - ref[k, :9]).max() <= 1e-5
        assert np.abs(t[k] - ref[k, 9:12]).max() <= 1e-5
    # the single-pair entry refuses the long pair with the same code
    r1 = ia.NativePairResult()
    assert lib.indialign_tmscore_search(ctypes.byref(pk.inps[2]), ctypes.byref(cfg),
                                        ctypes.byref(r1)) == ia.TOO_LONG
    assert np.isnan(r1.score)
