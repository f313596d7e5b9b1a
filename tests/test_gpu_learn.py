"""Learned ranking of the local-fragment work items on the GPU.

- the ranker's forward and backward kernels (matrix-core BF16 GEMMs) against
  the CPU oracle (oracle/mlp_oracle.py), and one Adam step;
- method 2 with an unlimited budget is the fast method exactly (selection
  path, item remapping and result layout);
- the packaged trained ranker on the 4096 held-out pairs (seeds
  60001..64096, disjoint from the training seeds 100000.. and not used to
  choose budgets): TM loss against the plain FP64 search within mean 1e-4 /
  max 1e-3, with a fraction of the NW work;
- ablation, end to end at equal budgets on the same pairs: the trained
  network against ranking by its skip input alone (descriptor column 31, the
  twice-refined same-index TM-score of each work item) and an untrained
  network.
"""
import os
import sys

import numpy as np
import pytest

import indialign_amd as ia
from indialign_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as mo  # noqa: E402

pytestmark = pytest.mark.gpu


def _gpu_or_skip():
    if not ia.gpu_available():
        pytest.fail("no GPU visible to the HIP runtime (GPU tests must run on the MI355X)")


def _rand_problem(n, seed):
    rng = np.random.default_rng(seed)
    p = ia.init_params(seed) * 0.5
    p[mo.OFF_B1:mo.OFF_W2] = rng.normal(0, 0.1, 256)
    p[mo.OFF_B2:mo.OFF_W3] = rng.normal(0, 0.1, 256)
    X = rng.uniform(0, 1, (n, mo.NFEAT)).astype(np.float32)
    t = rng.uniform(0, 1, n).astype(np.float32)
    return p.astype(np.float32), X, t


def test_mlp_forward_matches_oracle():
    _gpu_or_skip()
    for n in (1, 37, 1000, 70000):
        p, X, _ = _rand_problem(n, n)
        y = ia.mlp_forward(p, X)
        yo, _ = mo.forward(p, X)
        scale = max(float(np.abs(yo).max()), 1e-3)
        err = float(np.abs(y - yo).max())
        print(f"forward n={n}: max |gpu - oracle| {err:.2e} (scale {scale:.2e})", flush=True)
        # BF16 operands: a hidden value that rounds differently moves y by
        # at most ~2^-8 of one term
        assert err <= 5e-3 * scale


def test_mlp_grad_matches_oracle():
    _gpu_or_skip()
    for n in (5, 1000, 4096):
        p, X, t = _rand_problem(n, 100 + n)
        g, loss = ia.mlp_grad(p, X, t)
        go, lo = mo.backward(p, X, t)
        assert abs(loss - lo) <= 1e-5 * lo
        blocks = [("W1", mo.OFF_W1, mo.OFF_B1), ("b1", mo.OFF_B1, mo.OFF_W2),
                  ("W2", mo.OFF_W2, mo.OFF_B2), ("b2", mo.OFF_B2, mo.OFF_W3),
                  ("w3", mo.OFF_W3, mo.OFF_B3), ("b3", mo.OFF_B3, mo.NPARAM)]
        for name, a, b in blocks:
            rel = np.linalg.norm(g[a:b] - go[a:b]) / max(np.linalg.norm(go[a:b]), 1e-30)
            print(f"grad n={n} {name}: relative error {rel:.2e}", flush=True)
            assert rel <= 1e-3, name


def test_mlp_adam_step_matches_oracle():
    _gpu_or_skip()
    p, X, t = _rand_problem(2048, 7)
    lr = 1e-3
    p1, el = ia.mlp_train(p, X, t, epochs=1, batch=2048, lr=lr, seed=1)
    go, lo = mo.backward(p, X, t)
    po, _, _ = mo.adam(p.astype(np.float64), go, 0.0, 0.0, 1, lr)
    sig = np.abs(go) > 1e-3 * np.abs(go).max()
    assert sig.sum() > 1000
    assert np.abs(p1[sig] - po[sig]).max() <= 1e-6
    assert abs(el[0] - lo) <= 1e-5 * lo


def test_mlp_training_fits_a_known_function():
    _gpu_or_skip()
    rng = np.random.default_rng(11)
    X = rng.uniform(0, 1, (40000, mo.NFEAT)).astype(np.float32)
    t = (0.6 * X[:, 11] + 0.3 * X[:, 3] * X[:, 40] - 0.2 * X[:, 0]).astype(np.float32)
    p, el = ia.mlp_train(ia.init_params(2), X, t, epochs=12, batch=1024, lr=1e-3, seed=3)
    print("epoch loss", el.tolist(), flush=True)
    assert el[-1] < 0.05 * el[0]
    y = ia.mlp_forward(p, X[:2000])
    assert float(np.mean((y - t[:2000]) ** 2)) < 2e-3


def test_learned_unlimited_budget_is_fast_method():
    """Budgets above every pair's item and offset counts: the selections keep
    everything in order, so method 2 with the plain search's pass reduction
    (best chain score, then rescored) must reproduce method 1 bit for bit."""
    _gpu_or_skip()
    pairs = synth.short_pairs(24, first_seed=501)
    sf, Rf, tf = ia.tmscore_search_batch(pairs, serial_bank=1, method=ia.METHOD_FAST)
    ia.set_ranker(ia.init_params(0), top_m=1 << 30, top_offsets=1 << 30, top_seeds=0,
                  top_m_first=1 << 30, top_offset_banks=0)
    ia.set_learned_options(reduce_rescored=False)
    try:
        sl, Rl, tl = ia.tmscore_search_batch(pairs, serial_bank=1, method=ia.METHOD_LEARNED)
    finally:
        ia.set_learned_options(reduce_rescored=False)
        ia.set_ranker(None)
    np.testing.assert_array_equal(sl, sf)
    np.testing.assert_array_equal(Rl, Rf)
    np.testing.assert_array_equal(tl, tf)


def test_seed_screening_with_full_budget_changes_nothing():
    """Seed screening keeping at least every seed of every bank selects them
    all in order: the same results as no screening, bit for bit."""
    _gpu_or_skip()
    pairs = synth.short_pairs(16, first_seed=801)
    p = ia.init_params(0)
    ia.set_ranker(p, top_m=64, top_offsets=16, top_seeds=0)
    s0, R0, t0 = ia.tmscore_search_batch(pairs, serial_bank=1, method=ia.METHOD_LEARNED)
    ia.set_ranker(p, top_m=64, top_offsets=16, top_seeds=1 << 20)
    s1, R1, t1 = ia.tmscore_search_batch(pairs, serial_bank=1, method=ia.METHOD_LEARNED)
    ia.set_ranker(None)
    np.testing.assert_array_equal(s1, s0)
    np.testing.assert_array_equal(R1, R0)
    np.testing.assert_array_equal(t1, t0)


def test_learned_without_ranker_is_an_error():
    _gpu_or_skip()
    ia.set_ranker(None)
    with pytest.raises(ia.IndialignError):
        ia.tmscore_search_batch(synth.short_pairs(2), serial_bank=1, method=ia.METHOD_LEARNED)


HELD_OUT_FIRST_SEED, HELD_OUT_PAIRS = 60001, 4096


@pytest.fixture(scope="module")
def held_out():
    """The learned method's evaluation set: 4096 pairs of the short-pair
    distribution with seeds 60001..64096, used neither for training (seeds
    100000..) nor for choosing budgets, with their plain FP64 scores."""
    _gpu_or_skip()
    pairs = synth.short_pairs(HELD_OUT_PAIRS, first_seed=HELD_OUT_FIRST_SEED)
    cfg = ia.default_config()
    ia.batch_load(pairs)
    ia.batch_run(cfg, serial_bank=1)
    import time
    t0 = time.perf_counter()
    sp, _, _ = ia.batch_run(cfg, serial_bank=1)
    tp = time.perf_counter() - t0
    return {"pairs": pairs, "plain": sp, "plain_s": tp, "cells": ia.last_stats()["nw_cells"]}


def _learned_on(held, params, **budgets):
    import time
    cfg = ia.default_config()
    ia.set_ranker(params, **budgets)
    try:
        ia.batch_load(held["pairs"])
        ia.batch_run(cfg, serial_bank=1, method=ia.METHOD_LEARNED)
        t0 = time.perf_counter()
        sl, _, _ = ia.batch_run(cfg, serial_bank=1, method=ia.METHOD_LEARNED)
        tl = time.perf_counter() - t0
        st = ia.last_stats()
    finally:
        ia.set_ranker(None)
    d = held["plain"] - sl
    return {"loss": d, "s": tl, "speedup": held["plain_s"] / tl,
            "cell_frac": st["nw_cells"] / held["cells"], "fallback_passes": st["fallback_budget_passes"]}


def test_learned_tm_loss_on_held_out_pairs(held_out):
    """The north star's bars for the learned method on its 4096 held-out
    pairs: mean TM loss against the plain FP64 search <= 1e-4, no pair worse
    by more than 1e-3, and a fraction of the plain search's NW cells, with the
    packaged ranker and the default budgets (top 64 / 16 items, fallback
    budgets 512 / 64 for pairs predicted to stay below the rescue
    thresholds)."""
    r = _learned_on(held_out, np.load(ia.DEFAULT_RANKER))
    d = r["loss"]
    print(f"held-out learned: mean {d.mean():.3e} max {d.max():.3e} (pair {int(d.argmax())}), "
          f"pairs > 1e-4: {int((d > 1e-4).sum())}; {r['speedup']:.2f}x the plain search, "
          f"NW cells {r['cell_frac']:.3f} of plain, fallback-budget passes {r['fallback_passes']:.0f}",
          flush=True)
    assert d.mean() <= 1e-4
    assert d.max() <= 1e-3
    assert r["cell_frac"] < 0.2


def test_ablation_trained_ranker_vs_without(held_out):
    """What training gains, end to end at equal budgets on the held-out
    pairs: the packaged trained ranker against ranking by descriptor column
    31 alone (the network's skip input: every weight zero, i.e. no network)
    and an untrained network.  The trained ranker has the smallest worst-pair
    loss and refines the fewest NW cells (its predicted scores send fewer
    pairs to the fallback budget)."""
    zero = np.zeros(ia.NPARAM, np.float32)
    res = {"trained": _learned_on(held_out, np.load(ia.DEFAULT_RANKER)),
           "column 31": _learned_on(held_out, zero),
           "untrained": _learned_on(held_out, ia.init_params(5))}
    for k, r in res.items():
        d = r["loss"]
        print(f"{k}: TM loss mean {d.mean():.3e} max {d.max():.3e}, pairs > 1e-4 "
              f"{int((d > 1e-4).sum())}, {r['speedup']:.2f}x plain, NW cells {r['cell_frac']:.4f}, "
              f"fallback-budget passes {r['fallback_passes']:.0f}", flush=True)
    tr = res["trained"]
    for k in ("column 31", "untrained"):
        assert tr["loss"].max() <= res[k]["loss"].max(), k
        assert tr["cell_frac"] < res[k]["cell_frac"], k
    assert tr["loss"].max() <= 1e-3


