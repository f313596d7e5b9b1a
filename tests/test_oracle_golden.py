"""The CPU oracle against the golden fixtures written from the reference build
(tests/golden/make_golden.py).  Integer / mask / trace outputs must match
exactly; FP64 outputs within a few ulps of accumulated rounding (the reference
is built with -ffast-math, the oracle with IEEE semantics)."""
import ctypes

import numpy as np
import pytest

import libs
from indialign_amd import synth

D, I = ctypes.c_double, ctypes.c_int


def dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def up(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def aos(x):
    return np.ascontiguousarray(np.asarray(x, np.float64).reshape(-1))


@pytest.fixture(scope="module")
def orc():
    lib = libs.oracle()
    lib.oracle_tm_score.restype = ctypes.c_double
    return lib


def test_d0(orc, golden_functions):
    g = golden_functions
    for L, ref in zip(g["d0_L"], g["d0_out"]):
        o = np.zeros(3)
        orc.oracle_d0(D(L), dp(o))
        np.testing.assert_allclose(o, ref, rtol=1e-15, atol=0)


def test_jacobi_and_rotation(orc, golden_functions):
    g = golden_functions
    for a0, A, V in zip(g["jac_in"], g["jac_A"], g["jac_V"]):
        a = np.ascontiguousarray(a0.copy())
        v = np.zeros(16)
        orc.oracle_jacobi4(dp(a), dp(v))
        np.testing.assert_allclose(a, A, atol=1e-10)
        np.testing.assert_allclose(v, V, atol=1e-12)
    for H, R in zip(g["rot_H"], g["rot_R"]):
        r = np.zeros(9)
        orc.oracle_cov_to_rotation(dp(np.ascontiguousarray(H)), dp(r))
        np.testing.assert_allclose(r, R, atol=1e-12)


def test_kabsch_tm_weighted(orc, golden_functions):
    g = golden_functions
    for k in range(len(g["kab_P"])):
        P, Q, m = aos(g["kab_P"][k]), aos(g["kab_Q"][k]), np.ascontiguousarray(g["kab_m"][k])
        n = len(m)
        R, t = np.zeros(9), np.zeros(3)
        orc.oracle_kabsch(dp(P), dp(Q), up(m), I(n), dp(R), dp(t))
        np.testing.assert_allclose(R, g["kab_R"][k], atol=1e-12)
        np.testing.assert_allclose(t, g["kab_t"][k], atol=1e-10)
        s8 = orc.oracle_tm_score(dp(P), dp(Q), up(m), dp(R), dp(t), D(3.0), D(n), D(4.0), I(n), I(1))
        s0 = orc.oracle_tm_score(dp(P), dp(Q), up(m), dp(R), dp(t), D(3.0), D(n), D(0.0), I(n), I(0))
        assert abs(s8 - g["tm_d8"][k]) < 1e-13 and abs(s0 - g["tm_no_d8"][k]) < 1e-13
        R2, t2, sc = R.copy(), t.copy(), np.zeros(1)
        orc.oracle_weighted_refine(dp(P), dp(Q), up(m), D(3.0), D(n), I(n), I(4), dp(R2), dp(t2), dp(sc))
        # With < 3 points the optimal rotation is not unique (rank-deficient
        # correlation, repeated eigenvalues), so only the score is pinned.
        if m.sum() >= 3:
            np.testing.assert_allclose(R2, g["wref_R"][k], atol=1e-10)
        assert abs(sc[0] - g["wref_s"][k]) < 1e-12


def test_nw_bit_exact(orc, golden_functions):
    g = golden_functions
    for k in range(len(g["nw_shape"])):
        Np, Nn = (int(x) for x in g["nw_shape"][k])
        S = np.ascontiguousarray(g["nw_S"][k][:Np * Nn])
        gap = float(g["nw_gap"][k])
        H = np.zeros((Np + 1) * (Nn + 1))
        T = np.zeros((Np + 1) * (Nn + 1), np.int8)
        orc.oracle_nw_dp(dp(S), I(Np), I(Nn), D(gap), dp(H), T.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)))
        assert np.array_equal(H, g["nw_H"][k][:H.size]), k
        assert np.array_equal(T, g["nw_T"][k][:T.size]), k
        ap = np.zeros(max(1, min(Np, Nn)), np.int32)
        an = np.zeros(max(1, min(Np, Nn)), np.int32)
        na = orc.oracle_nw_align(dp(S), I(Np), I(Nn), D(gap), ip(ap), ip(an))
        assert na == g["nw_na"][k]
        assert np.array_equal(ap[:na], g["nw_ap"][k][:na]) and np.array_equal(an[:na], g["nw_an"][k][:na])


def test_seed_masks_exact(orc, golden_functions):
    g = golden_functions
    pos = 0
    for k in range(len(g["seedmask_K"])):
        row = g["seedmask_in"][k]
        N, ms = int(row[0]), int(row[1])
        valid = np.ascontiguousarray(row[2:2 + N].astype(np.uint8))
        cap = 2000
        buf = np.zeros(cap * N, np.uint8)
        K = orc.oracle_seed_masks(I(N), up(valid), I(1), I(ms), up(buf), I(cap))
        assert K == g["seedmask_K"][k]
        L = int(g["seedmask_bits_len"][k])
        ref = np.unpackbits(g["seedmask_bits"][pos:pos + L])[:K * N]
        pos += L
        assert np.array_equal(buf[:K * N], ref)


def _pair(seed):
    if seed == 104:
        return synth.make_pair(104, n_range=(40, 70), sigma_range=(2, 4), domain_motion=True, drop_frac=0.1)
    if seed == 105:
        return synth.make_pair(105, n=48, sigma=0.7)
    return synth.make_pair(seed, n_range=(30, 70))


def _close(o, ref, score_tol=1e-10, rt_tol=1e-8):
    assert abs(o[12] - ref[12]) <= score_tol, (o[12], ref[12])
    np.testing.assert_allclose(o[:12], ref[:12], atol=rt_tol)


def test_pair_stages(orc, golden_functions):
    """Anchors, top-k, refine, seed bank, detailed search, dp_refine (lite and
    full), local-fragment rescue (normal and dense), threading DP, SS
    assignment and SS seeds, gapless threading: one pair set, every stage."""
    g = golden_functions
    apos = tpos = spos = 0
    for k, seed in enumerate(g["pair_seeds"]):
        p = _pair(int(seed))
        N = p.n
        P, Q, v, pv, nv = aos(p.pred), aos(p.native), p.valid, p.pvalid, p.nvalid
        d0 = np.zeros(3)
        orc.oracle_d0(D(p.lnorm), dp(d0))
        d0_, d0s, sd8 = d0
        buf = np.zeros(4000 * N, np.uint8)
        nfrag = ctypes.c_int(0)
        K = orc.oracle_all_seeds(dp(P), dp(Q), up(v), I(N), D(p.lnorm), up(buf), I(4000), ctypes.byref(nfrag))
        assert (K, nfrag.value) == tuple(g["anch_meta"][k])
        L = int(g["anch_bits_len"][k])
        assert np.array_equal(buf[:K * N], np.unpackbits(g["anch_bits"][apos:apos + L])[:K * N])
        apos += L
        d2 = np.ascontiguousarray(np.where(v > 0, np.sum((p.pred - p.native) ** 2, 1), 1e12))
        m = np.zeros(N, np.uint8)
        orc.oracle_topk_seed(dp(d2), up(v), I(N), D(0.5), up(m))
        assert np.array_equal(m, g["topk"][tpos:tpos + N])
        tpos += N
        R, t, npass = np.zeros(9), np.zeros(3), ctypes.c_int(0)
        orc.oracle_refine_seed(dp(P), dp(Q), up(v), up(v), D(d0s), I(N), I(20), dp(R), dp(t), ctypes.byref(npass))
        np.testing.assert_allclose(np.concatenate([R, t]), g["refine"][k], atol=1e-9)
        cut = np.ascontiguousarray(d0s * np.array([2.0, 1.5, 1.0, 0.75, 0.5, 0.25]))
        o = np.zeros(13)
        orc.oracle_seed_bank(dp(P), dp(Q), up(v), up(buf), I(min(K, 40)), dp(cut), I(6), D(d0_), D(sd8),
                             D(p.lnorm), I(N), I(20), I(1), D(0.99), dp(o))
        _close(o, g["bank"][k])
        idx = np.ascontiguousarray(np.nonzero(pv & nv)[0].astype(np.int32))
        o = np.zeros(13)
        orc.oracle_detailed_search(dp(P), dp(Q), ip(idx), ip(idx), I(len(idx)), D(d0_), D(d0s), D(sd8),
                                   D(p.lnorm), I(1), I(-1), I(50), dp(o))
        _close(o, g["det"][k])
        for key, s8, it, lite in (("dpl", 100.0, 2, 1), ("dpf", sd8, 1, 0)):
            o = np.zeros(13)
            orc.oracle_dp_refine(dp(P), dp(Q), up(v), up(pv), up(nv), dp(R), dp(t), D(d0_), D(d0s), D(s8),
                                 D(p.lnorm), I(N), I(it), I(lite), dp(o))
            _close(o, g[key][k])
        for key, dense in (("lfn", 0), ("lfd", 1)):
            o = np.zeros(13)
            orc.oracle_local_fragment_dp(dp(P), dp(Q), up(v), up(pv), up(nv), D(d0_), D(d0s), D(p.lnorm),
                                         I(N), I(dense), dp(o))
            _close(o, g[key][k])
        o = np.zeros(13)
        orc.oracle_threading_dp(dp(P), dp(Q), up(v), up(pv), up(nv), D(d0_), D(d0s), D(p.lnorm), I(N), dp(o))
        _close(o, g["thr"][k])
        sp = ctypes.create_string_buffer(N)
        sn = ctypes.create_string_buffer(N)
        orc.oracle_assign_ss(dp(P), up(pv), I(N), sp)
        orc.oracle_assign_ss(dp(Q), up(nv), I(N), sn)
        assert sp.raw + sn.raw == bytes(g["ss"][spos:spos + 2 * N])
        spos += 2 * N
        o = np.zeros(13)
        orc.oracle_ss_seeds(dp(P), dp(Q), up(v), up(pv), up(nv), dp(R), dp(t), D(d0_), D(d0s), D(sd8),
                            D(p.lnorm), I(N), I(1), dp(o))
        _close(o, g["sss"][k])
        o = np.zeros(13)
        orc.oracle_gapless(dp(P), dp(Q), up(pv), up(nv), up(v), D(d0_), D(d0s), D(p.lnorm), I(N), dp(o))
        _close(o, g["gap"][k])


def golden_pairs(gs):
    pairs, pos = [], 0
    for k, n in enumerate(gs["n"]):
        n = int(n)
        sl = slice(pos, pos + n)
        pairs.append(synth.Pair(gs["pred"][sl], gs["native"][sl], gs["valid"][sl], gs["pvalid"][sl],
                                gs["nvalid"][sl], float(gs["lnorm"][k]), int(gs["seeds"][k]), 0.0))
        pos += n
    return pairs


def test_full_search_matches_reference(golden_search):
    """Whole pipeline: oracle batch (serial bank) vs the reference batch
    entry, and oracle parallel-bank semantics vs the reference single-pair
    entry run with 8 OpenMP threads."""
    gs = golden_search
    pairs = golden_pairs(gs)
    serial = libs.oracle_search(pairs, serial_bank=1)
    par = libs.oracle_search(pairs, serial_bank=0)
    for k in range(len(pairs)):
        _close(serial[k], gs["batch"][k], 1e-9, 1e-7)
        _close(par[k], gs["single"][k], 1e-9, 1e-7)
