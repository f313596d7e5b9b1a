"""The torch-tensor entry for the scoring drivers (indialign_amd.torch_ops):
argument checks and the d0 rule run on the CPU (they precede any device
call); on the GPU the adapter returns the C search's scores for the same
padded inputs."""
import ctypes

import numpy as np
import pytest
import torch

import indialign_amd as ia
import libs
from indialign_amd import synth, torch_ops


def oracle_d0(L):
    o = np.zeros(3)
    lib = libs.oracle()
    lib.oracle_d0(ctypes.c_double(L), o.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return o


def test_d0_rule_matches_the_search():
    """The adapter's check uses the search's d0 / d0_search / score_d8 rule
    (pinned against the oracle's, which the closed-form tests pin)."""
    for L in [1, 3, 11, 11.5, 12, 15, 16, 19, 20, 23, 24, 29, 29.9, 30, 47, 64, 150, 300, 1500]:
        np.testing.assert_allclose(torch_ops.d0_params(L), oracle_d0(L), rtol=1e-14, atol=0)


def padded_batch(pairs, N):
    B = len(pairs)
    pred = torch.zeros(B, N, 3, dtype=torch.float64)
    native = torch.zeros(B, N, 3, dtype=torch.float64)
    pv = torch.zeros(B, N, dtype=torch.bool)
    nv = torch.zeros(B, N, dtype=torch.bool)
    for b, p in enumerate(pairs):
        n = len(p.pred)
        pred[b, :n] = torch.from_numpy(p.pred)
        native[b, :n] = torch.from_numpy(p.native)
        pv[b, :n] = torch.from_numpy(p.pvalid.astype(bool))
        nv[b, :n] = torch.from_numpy(p.nvalid.astype(bool))
    L = torch.tensor([p.lnorm for p in pairs], dtype=torch.float64)
    d = torch.tensor([torch_ops.d0_params(float(x)) for x in L], dtype=torch.float32)
    return pred, native, pv & nv, d[:, 0], d[:, 1], d[:, 2], L.float(), pv, nv


def test_argument_checks_without_device():
    pairs = [synth.make_pair(5, n=40, sigma=1.0)]
    pred, native, valid, d0, d0s, d8, L, pv, nv = padded_batch(pairs, 48)
    with pytest.raises(ValueError):
        torch_ops.tmscore_search(pred[0], native[0], valid, d0, d0s, d8, L)
    with pytest.raises(ValueError):
        torch_ops.tmscore_search(pred, native[:, :40], valid, d0, d0s, d8, L)
    with pytest.raises(ValueError):
        torch_ops.tmscore_search(pred, native, valid[:, :40], d0, d0s, d8, L)
    with pytest.raises(ValueError):
        torch_ops.tmscore_search(pred, native, valid, d0, d0s, d8, L, pred_valid=pv[:, :3])
    with pytest.raises(ValueError):
        torch_ops.tmscore_search(pred, native, valid, d0[:0], d0s, d8, L)
    # a d0 rule other than the search's is refused
    with pytest.raises(ValueError, match="RNA rule"):
        torch_ops.tmscore_search(pred, native, valid, d0 * 1.1, d0s, d8, L)
    # empty batch
    out = torch_ops.tmscore_search(pred[:0], native[:0], valid[:0], d0[:0], d0s[:0], d8[:0], L[:0])
    assert out.shape == (0,) and out.dtype == pred.dtype
    # the torch engine's semantics: its cross-index rescue (separate masks)
    # is not rebuilt, and an unknown semantics is refused
    pv2 = pv.clone()
    pv2[0, 0] = ~pv2[0, 0]
    with pytest.raises(NotImplementedError):
        torch_ops.tmscore_search(pred, native, valid, d0, d0s, d8, L, pred_valid=pv2,
                                 native_valid=nv, semantics="torch")
    with pytest.raises(ValueError):
        torch_ops.tmscore_search(pred, native, valid, d0, d0s, d8, L, semantics="numpy")


TOY_PRED = [[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [1.0, 1.0, 0.0], [0.5, 0.2, 0.3]]
TOY_NATIVE = [[2.0, 3.0, 4.0], [2.8, 3.4, 4.1], [1.7, 4.1, 4.2], [2.5, 4.5, 4.3], [2.1, 3.6, 4.6]]


@pytest.mark.gpu
def test_torch_semantics_reference_snapshot():
    """The reference's own snapshot of its torch engine
    (tests/test_gpu_scoring_refactor.py:176-208: the 5-residue toy, float64,
    max_iter 20, fragment search, dp_iter 0) scores 0.7666330261384652;
    semantics="torch" reproduces it within the reference test's 1e-12, and
    the default C semantics give the C path's 0.77317584653223026 on the
    same inputs."""
    if not ia.gpu_available():
        pytest.fail("no GPU visible to the HIP runtime (GPU tests must run on the MI355X)")
    pred = torch.tensor([TOY_PRED], dtype=torch.float64)
    native = torch.tensor([TOY_NATIVE], dtype=torch.float64)
    valid = torch.ones((1, 5), dtype=torch.bool)
    L = valid.sum(dim=1).to(torch.float64)
    d = torch.tensor([torch_ops.d0_params(5.0)], dtype=torch.float64)
    args = (pred, native, valid, d[:, 0], d[:, 1], d[:, 2], L)
    kw = dict(max_iter=20, use_fragment_search=True, max_mem_gb=4.0, dp_iter=0)
    st = torch_ops.tmscore_search(*args, semantics="torch", **kw)
    sc = torch_ops.tmscore_search(*args, **kw)
    print(f"torch semantics {float(st[0])!r}, C semantics {float(sc[0])!r}", flush=True)
    assert abs(float(st[0]) - 0.7666330261384652) <= 1e-12
    assert abs(float(sc[0]) - 0.77317584653223026) <= 1e-12


@pytest.mark.gpu
def test_adapter_matches_batch_entry_and_oracle():
    """Padded batch (N 40-120 in a 128-row tensor, masks with dropped
    residues, float32 d0 tensors as the reference computes them): the
    adapter's scores equal the batch entry's on the same padded arrays and
    agree with the oracle within the parity bar."""
    if not ia.gpu_available():
        pytest.fail("no GPU visible to the HIP runtime")
    pairs = [synth.make_pair(40000 + k, n_range=(40, 120), drop_frac=0.1 if k % 2 else 0.0)
             for k in range(6)]
    args = padded_batch(pairs, 128)
    pred, native, valid, d0, d0s, d8, L, pv, nv = args
    for dp_iter in (0, 1):
        got = torch_ops.tmscore_search(pred, native, valid, d0, d0s, d8, L, dp_iter=dp_iter,
                                       pred_valid=pv, native_valid=nv)
        assert got.dtype == pred.dtype and got.shape == (6,)
        tup = [(pred[b].numpy(), native[b].numpy(), valid[b].numpy().astype(np.uint8),
                pv[b].numpy().astype(np.uint8), nv[b].numpy().astype(np.uint8), float(L[b]))
               for b in range(6)]
        cfg = ia.default_config(dp_iter=dp_iter)
        want, _, _ = ia.tmscore_search_batch(tup, cfg)
        assert np.array_equal(got.numpy(), want)
        opairs = [synth.Pair(*t, 0, 0.0) for t in tup]
        ref = libs.oracle_search(opairs, cfg=libs.oracle_config(dp_iter=dp_iter), serial_bank=1)
        print(f"dp_iter {dp_iter}: adapter {got.numpy()} oracle {ref[:, 12]}", flush=True)
        assert np.abs(got.numpy() - ref[:, 12]).max() <= 1e-6


def test_pair_batches_trivial_cases():
    """No frames -> 0.0; frame pairs whose native frame has <= 2 valid
    residues score 0 without a search."""
    z3 = torch.zeros(0, 10, 3, dtype=torch.float64)
    z = torch.zeros(0, 10, dtype=torch.bool)
    p = torch.randn(2, 10, 3, dtype=torch.float64)
    v = torch.ones(2, 10, dtype=torch.bool)
    assert torch_ops.score_pair_batches(z3, z, p, v) == 0.0
    assert torch_ops.score_pair_batches(p, v, z3, z) == 0.0
    nv = torch.zeros(2, 10, dtype=torch.bool)
    nv[:, :2] = True
    assert torch_ops.score_pair_batches(p, v, p, nv) == 0.0
    with pytest.raises(ValueError):
        torch_ops.score_pair_batches(p, v[:, :5], p, v)


@pytest.mark.gpu
def test_pair_batches_is_the_best_frame_pair():
    """3 pred frames x 2 native frames of one target (noisy copies, one
    native frame with dropped residues): the result is the maximum of the
    six single-pair searches, each checked against the oracle."""
    if not ia.gpu_available():
        pytest.fail("no GPU visible to the HIP runtime")
    base = synth.make_pair(41000, n=90, sigma=2.0)
    rng = np.random.default_rng(7)
    preds = np.stack([base.pred + rng.normal(0, s, base.pred.shape) for s in (0.5, 3.0, 8.0)])
    natives = np.stack([base.native, base.native + rng.normal(0, 1.0, base.native.shape)])
    pv = np.ones((3, 90), bool)
    pv[1, 10:20] = False
    nv = np.ones((2, 90), bool)
    nv[1, 40:50] = False
    got = torch_ops.score_pair_batches(torch.from_numpy(preds), torch.from_numpy(pv),
                                       torch.from_numpy(natives), torch.from_numpy(nv))
    singles = []
    for i in range(3):
        for j in range(2):
            m = (pv[i] & nv[j]).astype(np.uint8)
            singles.append(synth.Pair(preds[i], natives[j], m, m, m, float(nv[j].sum()), 0, 0.0))
    ref = libs.oracle_search(singles, serial_bank=1)
    s, _, _ = ia.tmscore_search_batch(singles)
    print(f"frames: best {got} singles {s} oracle {ref[:, 12]}", flush=True)
    assert got == float(s.max())
    assert np.abs(s - ref[:, 12]).max() <= 1e-6


@pytest.mark.gpu
def test_torch_semantics_vs_engine_vectors():
    """semantics="torch" (method 3) against the reference's PyTorch engine
    itself: tests/golden/make_torch_fixtures.py ran that engine (float64, CPU,
    one pair per call, max_iter 20, fragment search) on 44 seeded pairs of
    40-160 nt -- all-valid pairs, pairs with dropped residues under one
    same-index mask, and pairs whose correspondence is shifted by an indel --
    and stored its scores for dp_iter = 0, 1 and 2.  Each pair's score is
    checked within 1e-6 (the C parity bar; the engine's matrix product and
    eigen solver round differently from this library's FP64 sums)."""
    if not ia.gpu_available():
        pytest.fail("no GPU visible to the HIP runtime")
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    import make_torch_fixtures as mtf
    fx = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                              "torch_engine.npz"))
    pairs = [mtf.case_pair(str(k), int(s)) for k, s in zip(fx["kinds"], fx["seeds"])]
    np.testing.assert_allclose(np.array([mtf.checksum(p) for p in pairs]), fx["checksum"],
                               rtol=1e-12)
    for j, dp in enumerate(fx["dp_iters"]):
        cfg = ia.default_config(max_iter=20, dp_iter=int(dp))
        s, _, _ = ia.tmscore_search_batch(pairs, cfg, method=ia.METHOD_TORCH_SEED)
        d = np.abs(s - fx["scores"][:, j])
        print(f"dp_iter {dp}: max |dscore| {d.max():.2e}, mean {d.mean():.2e}, "
              f"identical {int((d == 0).sum())}/{len(d)}", flush=True)
        bad = np.where(d > 1e-6)[0]
        assert len(bad) == 0, [(str(fx["kinds"][k]), int(fx["seeds"][k]), float(s[k]),
                                float(fx["scores"][k, j])) for k in bad]
