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
