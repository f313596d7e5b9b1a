"""GPU parity: the HIP path through the C ABI against the CPU oracle and the
golden fixtures written from the reference build.  Bar of the plain FP64 path:
score within 1e-6 absolute, R and t within 1e-5, per pair."""
import ctypes
import os
import sys

import numpy as np
import pytest

import indialign_amd as ia
import libs
from indialign_amd import synth

pytestmark = pytest.mark.gpu

SCORE_TOL, RT_TOL = 1e-6, 1e-5


def _gpu_or_skip():
    if not ia.gpu_available():
        pytest.fail("no GPU visible to the HIP runtime (GPU tests must run on the MI355X)")


def _check(scores, R, t, ref, label):
    bad = []
    for k in range(len(scores)):
        ds = abs(scores[k] - ref[k, 12])
        dR = np.abs(R[k].reshape(-1) - ref[k, :9]).max()
        dt = np.abs(t[k] - ref[k, 9:12]).max()
        print(f"{label} pair {k}: gpu {scores[k]:.15f} ref {ref[k,12]:.15f} ds={ds:.1e} dR={dR:.1e} dt={dt:.1e}",
              flush=True)
        if not (ds <= SCORE_TOL and dR <= RT_TOL and dt <= RT_TOL):
            bad.append(k)
    assert not bad, f"{label}: pairs outside tolerance: {bad}"


def golden_pairs(gs):
    pairs, pos = [], 0
    for k, n in enumerate(gs["n"]):
        n = int(n)
        sl = slice(pos, pos + n)
        pairs.append(synth.Pair(gs["pred"][sl], gs["native"][sl], gs["valid"][sl], gs["pvalid"][sl],
                                gs["nvalid"][sl], float(gs["lnorm"][k]), int(gs["seeds"][k]), 0.0))
        pos += n
    return pairs


def test_golden_search_batch(golden_search):
    """31 fixture pairs (the single-pair workload, 24 short pairs, 4 pairs with
    masks and domain motion, 2 near-exact pairs) through the batch entry vs
    the reference batch entry's outputs."""
    _gpu_or_skip()
    pairs = golden_pairs(golden_search)
    s, R, t = ia.tmscore_search_batch(pairs)
    _check(s, R, t, golden_search["batch"], "golden-batch")


def test_golden_search_single_entry(golden_search):
    """Single-pair entry (multithreaded reference semantics: no early
    termination of the main seed bank) vs the reference single-pair entry."""
    _gpu_or_skip()
    pairs = golden_pairs(golden_search)
    idx = [0, 1, 5, 26, 29, 30]
    out_s, out_R, out_t = [], [], []
    for k in idx:
        p = pairs[k]
        sc, R, t = ia.tmscore_search(p.pred, p.native, p.valid, p.pvalid, p.nvalid, p.lnorm)
        out_s.append(sc); out_R.append(R); out_t.append(t)
    _check(out_s, out_R, out_t, golden_search["single"][idx], "golden-single")


def test_short_pairs_vs_oracle():
    """40 more short pairs (seeds 1001..1040, N~U{40,300}) vs the oracle."""
    _gpu_or_skip()
    pairs = synth.short_pairs(40, first_seed=1001)
    s, R, t = ia.tmscore_search_batch(pairs)
    ref = libs.oracle_search(pairs, serial_bank=1)
    _check(s, R, t, ref, "short")


def test_long_partial_pairs_vs_oracle():
    """Long pairs with masked residues and domain motion (N 400-520, 10%
    pred/native residues dropped independently, Lnorm = native count)."""
    _gpu_or_skip()
    pairs = [synth.make_pair(s, n_range=(400, 520), sigma_range=(2.0, 10.0), domain_motion=True,
                             drop_frac=0.10) for s in (20001, 20002)]
    s, R, t = ia.tmscore_search_batch(pairs)
    ref = libs.oracle_search(pairs, serial_bank=1)
    _check(s, R, t, ref, "long")


def test_long_pairs_global_trace_vs_oracle():
    """The long-pair configuration (N 400-1500, domain motion, masks): six
    pairs of N 541-1122, whose NW traces exceed LDS and live in per-wavefront
    global slots, and one noisy N = 600 pair that runs every rescue pass,
    dense included, at that size."""
    _gpu_or_skip()
    pairs = synth.long_pairs(6) + [synth.make_pair(20999, n_range=(600, 600),
                                                   sigma_range=(12.0, 12.0), domain_motion=True,
                                                   drop_frac=0.10)]
    s, R, t = ia.tmscore_search_batch(pairs)
    ref = libs.oracle_search(pairs, serial_bank=1)
    _check(s, R, t, ref, "long-global-trace")
    assert ref[-1, 12] < 0.3  # the dense rescue pass ran


def test_fill_round_boundaries_vs_oracle():
    """Chain sizes at the NW fill's round boundaries (a lane computes rows
    l, l+64, ...; the period is max(nlen, 65)): N = 63-66, 127-130, 192, 193,
    and masked pairs whose compact row and column counts differ.  The noise
    keeps every pair below the rescue thresholds, so the dense rescue runs
    its chains at exactly these sizes."""
    _gpu_or_skip()
    pairs = [synth.make_pair(30000 + n, n=n, sigma=10.0)
             for n in (63, 64, 65, 66, 127, 128, 129, 130, 192, 193)]
    pairs += [synth.make_pair(30500 + n, n=n, sigma=9.0, drop_frac=0.15) for n in (70, 140, 200)]
    s, R, t = ia.tmscore_search_batch(pairs)
    ref = libs.oracle_search(pairs, serial_bank=1)
    _check(s, R, t, ref, "fill-boundaries")
    assert (ref[:, 12] < 0.3).all()  # every pair ran the dense rescue
    for p in pairs[-3:]:
        assert p.pvalid.sum() != p.nvalid.sum() or p.pvalid.sum() != len(p.pvalid)


def test_selection_cache_boundaries_vs_oracle():
    """Point sets at the 512-point limits of the plain refinement's selection
    caches (seed banks: bits of points 0..511 kept in LDS, later points
    recomputed; lite search: register mask for aligned sets up to 512):
    N = 511, 512, 513 with every residue valid, one noisy N = 512 pair whose
    rescue chains run at that size, and a masked N = 600 pair (538 residues
    valid on both sides)."""
    _gpu_or_skip()
    pairs = [synth.make_pair(40000 + n + 3, n=n, sigma=3.0) for n in (511, 512, 513)]
    pairs.append(synth.make_pair(40522, n=512, sigma=10.0))
    pairs.append(synth.make_pair(40600, n=600, sigma=4.0, drop_frac=0.05))  # 538 same-index
    s, R, t = ia.tmscore_search_batch(pairs)
    ref = libs.oracle_search(pairs, serial_bank=1)
    _check(s, R, t, ref, "selection-cache")
    assert [int(p.valid.sum()) for p in pairs[:3]] == [511, 512, 513]


def test_anchor_global_path_vs_oracle():
    """A masked N = 1600 pair, beyond the 1536 residues whose anchor-seed
    counts, candidate list, neighbourhood mask and coordinates the anchor
    kernel keeps in LDS: its global-memory path (149 same-index valid
    residues, 485 / 517 pred / native valid) against the oracle."""
    _gpu_or_skip()
    p = synth.make_pair(41600, n=1600, sigma=3.0, drop_frac=0.7)
    assert len(p.valid) == 1600 and int(p.valid.sum()) == 149
    s, R, t = ia.tmscore_search_batch([p])
    _check(s, R, t, libs.oracle_search([p], serial_bank=1), "anchor-global")


def test_reference_toy_pair():
    """The 5-residue toy of the reference tests (tests/test_gpu_scoring_refactor.py:179-190):
    the reference C path scores it 0.77317584653223026 for dp_iter 0 and 1."""
    _gpu_or_skip()
    pred = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [0.5, 0.2, 0.3]], float)
    native = np.array([[2, 3, 4], [2.8, 3.4, 4.1], [1.7, 4.1, 4.2], [2.5, 4.5, 4.3], [2.1, 3.6, 4.6]])
    for dp_iter in (0, 1):
        sc, R, t = ia.tmscore_search(pred, native, cfg=ia.default_config(dp_iter=dp_iter))
        assert abs(sc - 0.77317584653223026) < 1e-12, sc


def test_edge_cases():
    _gpu_or_skip()
    lib = ia.load_library()
    cfg = ia.default_config()
    # Lnorm <= 2 -> score 0, identity (indialign.cpp:47-51)
    p = synth.make_pair(3, n=30, sigma=1.0)
    sc, R, t = ia.tmscore_search(p.pred, p.native, lnorm=2.0)
    assert sc == 0.0 and np.array_equal(R, np.eye(3)) and np.array_equal(t, np.zeros(3))
    # no valid residue -> no seed -> score 0, identity (indialign.cpp:77-81)
    z = np.zeros(30, np.uint8)
    sc, R, t = ia.tmscore_search(p.pred, p.native, z, z, z, 30.0)
    assert sc == 0.0 and np.array_equal(R, np.eye(3))
    # N <= 0 -> -1 without touching the output
    inp = ia.NativePairInput()
    inp.length = -3
    res = ia.NativePairResult()
    res.score = 42.0
    assert lib.indialign_tmscore_search(ctypes.byref(inp), ctypes.byref(cfg), ctypes.byref(res)) == -1
    assert res.score == 42.0
    # exact rigid copy -> 1 with the generating transform
    q = synth.make_pair(17, n=80, sigma=0.0)
    sc, R, t = ia.tmscore_search(q.pred, q.native)
    assert abs(sc - 1.0) < 1e-12
    np.testing.assert_allclose(q.pred @ R + t, q.native, atol=1e-8)
    # tiny chains: N = 3, 4, 5 vs oracle; mixed batch with an empty pair
    tiny = [synth.make_pair(50 + n, n=n, sigma=1.0) for n in (3, 4, 5, 7)]
    s, R, t = ia.tmscore_search_batch(tiny)
    _check(s, R, t, libs.oracle_search(tiny, serial_bank=1), "tiny")


def _smoke_inputs():
    """Exact inputs of the reference's tests/test_cpp_smoke.py:34-105."""
    from scipy.spatial.transform import Rotation
    rng = np.random.RandomState(42)

    def make_chain(N):
        coords = np.zeros((N, 3))
        d = rng.randn(3)
        d /= np.linalg.norm(d)
        for i in range(1, N):
            d = d + rng.randn(3) * 0.3
            d /= np.linalg.norm(d)
            coords[i] = coords[i - 1] + d * (5.9 + rng.randn() * 0.5)
        return coords - coords.mean(axis=0)

    cases = []
    native = make_chain(50)
    cases.append((Rotation.random(random_state=rng).apply(native) + rng.randn(50, 3) * 1.0, native, 0.5))
    native = make_chain(200)
    cases.append((Rotation.random(random_state=rng).apply(native) + rng.randn(200, 3) * 3.0, native, 0.2))
    native = make_chain(100)
    cases.append((native.copy(), native, 0.95))
    native = make_chain(6000)
    cases.append((Rotation.random(random_state=rng).apply(native) + rng.randn(6000, 3) * 2.0, native, 0.3))
    native = make_chain(8192)
    cases.append((Rotation.random(random_state=rng).apply(native) + rng.randn(8192, 3) * 2.0, native, 0.3))
    return cases


def test_reference_smoke_thresholds():
    """The reference's own ctypes smoke test (tests/test_cpp_smoke.py:63-105):
    rc == 0 and the same score thresholds, N up to 8192, single-pair entry."""
    _gpu_or_skip()
    for k, (pred, native, thr) in enumerate(_smoke_inputs()):
        sc, R, t = ia.tmscore_search(pred, native)
        print(f"smoke case {k}: N={len(native)} score={sc:.6f} threshold={thr}", flush=True)
        assert sc > thr
