"""The device Needleman-Wunsch of the chain kernel on supplied score matrices
(indialign_nw_matrix), bit for bit against the reference's nw_dp /
nw_traceback (indialign_c/nw.h:10-68): H, the trace codes and the aligned
pairs.

- the 36 matrices of tests/golden/functions.npz (written by the reference
  build: random, tie-heavy and zero-heavy matrices, gap opens -0.6, 0, -1);
- larger matrices against the oracle (itself pinned to those fixtures):
  shapes around the fill's 64-row rounds and 65-column period, wide and
  tall, a single row or column, tie-heavy, zero-heavy and SS-style {0, 1}
  matrices, every gap open the search uses;
- both trace placements (LDS and a global slot).
"""
import ctypes

import numpy as np
import pytest

import indialign_amd as ia
import libs

pytestmark = pytest.mark.gpu

D = ctypes.c_double
I = ctypes.c_int


def _gpu_or_skip():
    if not ia.gpu_available():
        pytest.fail("no GPU visible to the HIP runtime (GPU tests must run on the MI355X)")


def _lib():
    lib = ia.load_library()
    ip = ctypes.POINTER(ctypes.c_int)
    dp = ctypes.POINTER(ctypes.c_double)
    lib.indialign_nw_matrix.argtypes = [ctypes.c_int, dp, ip, ip, dp, ctypes.c_int, ctypes.c_int,
                                        dp, ctypes.POINTER(ctypes.c_int8), ip, ip, ip]
    lib.indialign_nw_matrix.restype = ctypes.c_int
    return lib


def device_nw(mats, gaps, fast=0, trace_global=0):
    """Run every (matrix, gap) on the device; per matrix (H, T, aligned p, aligned n)."""
    lib = _lib()
    pl = np.array([m.shape[0] for m in mats], np.int32)
    nl = np.array([m.shape[1] for m in mats], np.int32)
    S = np.ascontiguousarray(np.concatenate([m.reshape(-1) for m in mats]), np.float64)
    g = np.array(gaps, np.float64)
    hs = (pl.astype(np.int64) + 1) * (nl + 1)
    H = np.empty(hs.sum(), np.float64)
    T = np.empty(hs.sum(), np.int8)
    am = np.minimum(pl, nl).astype(np.int64)
    ap = np.empty(max(1, am.sum()), np.int32)
    an = np.empty(max(1, am.sum()), np.int32)
    na = np.empty(len(mats), np.int32)
    c = lambda a, t: a.ctypes.data_as(ctypes.POINTER(t))
    rc = lib.indialign_nw_matrix(len(mats), c(S, D), c(pl, I), c(nl, I), c(g, D), fast,
                                 trace_global, c(H, D), c(T, ctypes.c_int8), c(ap, I), c(an, I),
                                 c(na, I))
    assert rc == 0
    out, ho, ao = [], 0, 0
    for k in range(len(mats)):
        h = int(hs[k])
        out.append((H[ho:ho + h], T[ho:ho + h], ap[ao:ao + na[k]], an[ao:ao + na[k]]))
        ho += h
        ao += int(am[k])
    return out


def oracle_nw(S, gap):
    orc = libs.oracle()
    Np, Nn = S.shape
    S = np.ascontiguousarray(S, np.float64)
    H = np.zeros((Np + 1) * (Nn + 1))
    T = np.zeros((Np + 1) * (Nn + 1), np.int8)
    orc.oracle_nw_dp(libs.dptr(S), I(Np), I(Nn), D(gap), libs.dptr(H),
                     T.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)))
    ap = np.zeros(min(Np, Nn), np.int32)
    an = np.zeros(min(Np, Nn), np.int32)
    na = orc.oracle_nw_align(libs.dptr(S), I(Np), I(Nn), D(gap), libs.iptr(ap), libs.iptr(an))
    return H, T, ap[:na], an[:na]


@pytest.mark.parametrize("trace_global", [0, 1])
def test_nw_matches_reference_fixtures(golden_functions, trace_global):
    _gpu_or_skip()
    g = golden_functions
    mats, gaps = [], []
    for k in range(len(g["nw_shape"])):
        Np, Nn = (int(x) for x in g["nw_shape"][k])
        mats.append(np.ascontiguousarray(g["nw_S"][k][:Np * Nn]).reshape(Np, Nn))
        gaps.append(float(g["nw_gap"][k]))
    res = device_nw(mats, gaps, trace_global=trace_global)
    for k, (H, T, ap, an) in enumerate(res):
        Np, Nn = mats[k].shape
        h = (Np + 1) * (Nn + 1)
        assert np.array_equal(H, g["nw_H"][k][:h]), k
        assert np.array_equal(T, g["nw_T"][k][:h]), k
        na = int(g["nw_na"][k])
        assert len(ap) == na, k
        assert np.array_equal(ap, g["nw_ap"][k][:na]) and np.array_equal(an, g["nw_an"][k][:na]), k


def _matrices():
    rng = np.random.default_rng(2024)
    shapes = [(64, 65), (65, 64), (63, 66), (70, 130), (130, 70), (127, 129), (128, 128),
              (129, 65), (200, 200), (1, 7), (7, 1), (1, 1), (300, 190), (190, 300), (520, 90)]
    mats, gaps, kinds = [], [], []
    for k, (p, n) in enumerate(shapes):
        for kind in ("random", "ties", "zeros", "ss"):
            if kind == "random":
                S = rng.uniform(0, 1, (p, n))
            elif kind == "ties":
                S = rng.choice([0.0, 0.5, 1.0], (p, n))
            elif kind == "zeros":
                S = np.where(rng.uniform(size=(p, n)) < 0.85, 0.0, rng.uniform(0, 1, (p, n)))
            else:
                S = rng.choice([0.0, 1.0], (p, n))
            gap = -1.0 if kind == "ss" else (-0.6, 0.0, -1.0)[k % 3]
            mats.append(S)
            gaps.append(gap)
            kinds.append(f"{kind} {p}x{n} gap {gap}")
    return mats, gaps, kinds


@pytest.mark.parametrize("trace_global", [0, 1])
def test_nw_matches_oracle_on_larger_matrices(trace_global):
    _gpu_or_skip()
    mats, gaps, kinds = _matrices()
    res = device_nw(mats, gaps, trace_global=trace_global)
    for k, (H, T, ap, an) in enumerate(res):
        Ho, To, apo, ano = oracle_nw(mats[k], gaps[k])
        assert np.array_equal(H, Ho), kinds[k]
        assert np.array_equal(T, To), kinds[k]
        assert np.array_equal(ap, apo) and np.array_equal(an, ano), kinds[k]
    print(f"{len(mats)} matrices bit-identical (trace_global={trace_global})", flush=True)


def test_nw_fp32_fill_close_to_fp64():
    """The fast method's FP32 fill on the gap-0 matrices: there every H is the
    best path score whatever the tie decisions, so it must equal the FP64
    value within FP32 rounding of the path sums.  (With a gap open, H depends
    on how ties were broken, because the gap applies only after a diagonal
    step, and FP32 rounding may break near-ties the other way; the fast
    method rescores every alignment in FP64.)"""
    _gpu_or_skip()
    mats, gaps, kinds = _matrices()
    sel = [k for k in range(len(mats)) if gaps[k] == 0.0]
    assert len(sel) >= 10
    res = device_nw([mats[k] for k in sel], [0.0] * len(sel), fast=1)
    for (H, T, ap, an), k in zip(res, sel):
        Ho = oracle_nw(mats[k], 0.0)[0]
        steps = mats[k].shape[0] + mats[k].shape[1]
        assert np.abs(H - Ho).max() <= 2.0 ** -22 * steps * max(1.0, float(np.abs(Ho).max())), kinds[k]


def test_nw_rejects_bad_sizes():
    _gpu_or_skip()
    lib = _lib()
    pl = np.array([0], np.int32)
    nl = np.array([5], np.int32)
    z = np.zeros(8)
    zi = np.zeros(8, np.int32)
    c = lambda a, t: a.ctypes.data_as(ctypes.POINTER(t))
    assert lib.indialign_nw_matrix(1, c(z, D), c(pl, I), c(nl, I), c(z, D), 0, 0, c(z, D),
                                   c(np.zeros(8, np.int8), ctypes.c_int8), c(zi, I), c(zi, I),
                                   c(zi, I)) == -1
