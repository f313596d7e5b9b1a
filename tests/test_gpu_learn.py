"""Learned ranking of the local-fragment work items on the GPU.

- the ranker's forward and backward kernels (matrix-core BF16 GEMMs) against
  the CPU oracle (oracle/mlp_oracle.py), and one Adam step;
- method 2 with an unlimited budget is the fast method exactly (selection
  path, item remapping and result layout);
- the packaged trained ranker on held-out pairs (seeds 1.., disjoint from the
  training seeds 10000..): TM loss against the plain FP64 search within
  mean 1e-4 / max 1e-3, with a fraction of the NW work;
- ablation: at the same budget the trained network finds the best work item
  far more often than the same search without it (the first m items), and
  training does not make it worse than its skip input alone (an untrained
  network with near-zero output weights ranks by the twice-refined same-index
  TM-score of each work item).
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
    everything in order, so method 2 must reproduce method 1 bit for bit."""
    _gpu_or_skip()
    pairs = synth.short_pairs(24, first_seed=501)
    sf, Rf, tf = ia.tmscore_search_batch(pairs, serial_bank=1, method=ia.METHOD_FAST)
    ia.set_ranker(ia.init_params(0), top_m=1 << 30, top_offsets=1 << 30, top_seeds=0,
                  top_m_first=1 << 30, top_offset_banks=0)
    sl, Rl, tl = ia.tmscore_search_batch(pairs, serial_bank=1, method=ia.METHOD_LEARNED)
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


def test_learned_tm_loss_on_held_out_pairs():
    _gpu_or_skip()
    pairs = synth.short_pairs(256)
    sp, _, _ = ia.tmscore_search_batch(pairs, serial_bank=1)
    cells_plain = ia.last_stats()["nw_cells"]
    ia.load_default_ranker(top_m=64)
    sl, _, _ = ia.tmscore_search_batch(pairs, serial_bank=1, method=ia.METHOD_LEARNED)
    cells_learned = ia.last_stats()["nw_cells"]
    ia.set_ranker(None)
    d = sp - sl
    print(f"TM loss vs plain: mean {d.mean():.2e} max {d.max():.2e}; NW cells "
          f"{cells_learned:.3e} vs {cells_plain:.3e}", flush=True)
    assert d.mean() <= 1e-4
    assert d.max() <= 1e-3
    assert cells_learned < 0.2 * cells_plain


def _recall(score, y, pair, dense, m):
    hit, miss = [], []
    for k in np.unique(pair):
        for dflag in (False, True):
            idx = np.where((pair == k) & (dense == dflag))[0]
            if len(idx) == 0:
                continue
            top = idx[np.lexsort((idx, -score[idx]))[:m]]
            hit.append(y[top].max() >= y[idx].max())
            miss.append(y[idx].max() - y[top].max())
    return float(np.mean(hit)), float(np.mean(miss))


def test_ablation_trained_ranker_vs_without():
    """Records of held-out pairs: how often the top-m work items contain the
    pair's best one (the item the full pass would keep), for the trained
    network, the search without a network at the same budget (the first m
    items) and an untrained network."""
    _gpu_or_skip()
    pairs = synth.short_pairs(48, first_seed=3001)
    _, _, _, X, y, pair = ia.collect_records(pairs, cap=2_000_000)
    assert len(y) > 10000
    dense = X[:, 8] > 0.5
    trained = np.load(ia.DEFAULT_RANKER)
    m = 16
    r_tr, miss_tr = _recall(ia.mlp_forward(trained, X), y, pair, dense, m)
    r_none, miss_none = _recall(np.zeros(len(y), np.float32), y, pair, dense, m)
    r_rand, miss_rand = _recall(ia.mlp_forward(ia.init_params(5), X), y, pair, dense, m)
    r_heur, miss_heur = _recall(X[:, 11], y, pair, dense, m)
    print(f"top-{m} recall of the best item / mean refined-score miss: trained {r_tr:.3f}/{miss_tr:.2e}, "
          f"first-m {r_none:.3f}/{miss_none:.2e}, untrained {r_rand:.3f}/{miss_rand:.2e}, "
          f"same-index TM {r_heur:.3f}/{miss_heur:.2e}", flush=True)
    assert r_tr >= r_none + 0.3
    assert miss_tr <= 0.5 * miss_none
    assert r_tr >= r_rand - 0.05
    assert miss_tr <= 1.2 * miss_rand
