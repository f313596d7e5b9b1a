"""Full-size parity of the plain FP64 path against the reference build.

The fixtures (tests/golden/make_full_fixtures.py) hold the reference
library's (score, R, t) for every pair of the two batch workloads and for
the two largest cases of the reference's own smoke test; the inputs are
regenerated here and checked against the stored checksums.  Bar, per pair:
score within 1e-6, R and t within 1e-5.
"""
import os

import numpy as np
import pytest

import indialign_amd as ia
from indialign_amd import synth

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SCORE_TOL, RT_TOL = 1e-6, 1e-5


def _gpu_or_skip():
    if not ia.gpu_available():
        pytest.fail("no GPU visible to the HIP runtime (GPU tests must run on the MI355X)")


def _checksum(pairs):
    out = np.zeros((len(pairs), 4))
    for k, p in enumerate(pairs):
        out[k] = [np.abs(p.pred).sum() + 2.0 * np.abs(p.native).sum(), p.valid.sum(),
                  p.pvalid.sum(), p.nvalid.sum()]
    return out


def _compare(label, fx, s, R, t):
    ds = np.abs(s - fx["score"])
    dR = np.abs(R.reshape(len(s), 9) - fx["R"]).max(axis=1)
    dt = np.abs(t - fx["t"]).max(axis=1)
    bad = np.where((ds > SCORE_TOL) | (dR > RT_TOL) | (dt > RT_TOL))[0]
    print(f"{label}: {len(s)} pairs, max |dscore| {ds.max():.2e}, max |dR| {dR.max():.2e}, "
          f"max |dt| {dt.max():.2e}, bit-identical scores {int((ds == 0).sum())}; "
          f"(oracle vs reference build: {float(fx['oracle_max_dscore']):.2e})", flush=True)
    assert len(bad) == 0, f"{label}: pairs outside the bar: {bad[:20].tolist()}"


@pytest.mark.parametrize("name,make", [("short", lambda: synth.short_pairs(4096)),
                                       ("long", lambda: synth.long_pairs(512))])
def test_batch_workload_vs_reference(name, make):
    """Every pair of the batch workload through the batch entry (serial
    seed-bank semantics, as the reference batch entry that wrote them)."""
    _gpu_or_skip()
    fx = np.load(os.path.join(GOLDEN, f"full_{name}.npz"))
    pairs = make()
    assert np.array_equal(np.array([p.seed for p in pairs]), fx["seeds"])
    np.testing.assert_allclose(_checksum(pairs), fx["checksum"], rtol=1e-12)
    s, R, t = ia.tmscore_search_batch(pairs)
    _compare(name, fx, s, R, t)


def test_smoke_large_cases_vs_reference():
    """N = 6000 and N = 8192 (tests/test_cpp_smoke.py:87-105) through the
    single-pair entry with 8 OpenMP threads' semantics, as the reference ran
    them; checked per pair instead of the reference test's 0.3 threshold."""
    _gpu_or_skip()
    from test_gpu_parity import _smoke_inputs
    fx = np.load(os.path.join(GOLDEN, "smoke_large.npz"))
    cases = _smoke_inputs()[3:]
    pairs = [synth.Pair(p, q, np.ones(len(q), np.uint8), np.ones(len(q), np.uint8),
                        np.ones(len(q), np.uint8), float(len(q)), 42, 2.0) for p, q, _ in cases]
    np.testing.assert_allclose(_checksum(pairs), fx["checksum"], rtol=1e-12)
    old = os.environ.get("OMP_NUM_THREADS")
    os.environ["OMP_NUM_THREADS"] = "8"
    try:
        out = [ia.tmscore_search(p.pred, p.native) for p in pairs]
    finally:
        if old is None:
            os.environ.pop("OMP_NUM_THREADS", None)
        else:
            os.environ["OMP_NUM_THREADS"] = old
    s = np.array([o[0] for o in out])
    R = np.array([o[1] for o in out])
    t = np.array([o[2] for o in out])
    _compare("smoke N=6000, 8192", fx, s, R, t)
