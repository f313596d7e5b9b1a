"""Pins for the CPU oracle that do not come from the reference build:
closed forms, textbook equivalents and brute force on tiny inputs, chosen so
that a dropped term, a wrong sign, a transposed operand or a wrong tie rule
fails one of them."""
import ctypes
import itertools
import math

import numpy as np

import libs
from indialign_amd import synth

D, I = ctypes.c_double, ctypes.c_int


def dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def up(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def orc():
    lib = libs.oracle()
    lib.oracle_tm_score.restype = ctypes.c_double
    return lib


def test_d0_closed_form():
    """RNA d0: piecewise constants below 30, 0.6*sqrt(L-0.5)-2.5 above; d0
    search clamp(1.24*cbrt(L-15)-1.8, 4.5, 8); d8 = 1.5 L^0.3 + 3.5."""
    lib = orc()
    for L, d0 in [(5, 0.3), (11, 0.3), (12, 0.4), (15, 0.4), (16, 0.5), (19, 0.5), (20, 0.6),
                  (23, 0.6), (24, 0.7), (29, 0.7)]:
        o = np.zeros(3)
        lib.oracle_d0(D(L), dp(o))
        assert o[0] == d0
    for L in [30, 64, 100, 300, 1500]:
        o = np.zeros(3)
        lib.oracle_d0(D(L), dp(o))
        assert math.isclose(o[0], 0.6 * math.sqrt(L - 0.5) - 2.5, rel_tol=1e-15)
        assert math.isclose(o[1], min(8.0, max(4.5, 1.24 * (L - 15) ** (1 / 3) - 1.8)), rel_tol=1e-14)
        assert math.isclose(o[2], 1.5 * L ** 0.3 + 3.5, rel_tol=1e-15)


def svd_kabsch(P, Q):
    cp, cq = P.mean(0), Q.mean(0)
    H = (P - cp).T @ (Q - cq)
    U, S, Vt = np.linalg.svd(H)
    d = np.sign(np.linalg.det(Vt.T @ U.T))
    Rm = Vt.T @ np.diag([1, 1, d]) @ U.T  # q = Rm p + t
    return Rm, cq - Rm @ cp


def test_kabsch_recovers_rigid_motion():
    """Noise-free rigid data: exact recovery, det +1, orthonormal, and the
    same optimum as an SVD Kabsch on noisy data."""
    lib = orc()
    rng = np.random.Generator(np.random.Philox(key=5))
    for k in range(20):
        n = 5 + k * 3
        P = rng.standard_normal((n, 3)) * 10
        Rt = synth.haar_rotation(rng)
        t = rng.uniform(-30, 30, 3)
        Q = P @ Rt.T + t
        R, tt = np.zeros(9), np.zeros(3)
        lib.oracle_kabsch(dp(np.ascontiguousarray(P.reshape(-1))), dp(np.ascontiguousarray(Q.reshape(-1))),
                          up(np.ones(n, np.uint8)), I(n), dp(R), dp(tt))
        Rm = R.reshape(3, 3)  # out_d = sum_k P_k R[3k+d]  ->  Q = P @ Rm + t
        np.testing.assert_allclose(P @ Rm + tt, Q, atol=1e-9)
        np.testing.assert_allclose(Rm.T @ Rm, np.eye(3), atol=1e-12)
        assert abs(np.linalg.det(Rm) - 1) < 1e-12
        Qn = Q + rng.standard_normal(Q.shape) * 2.0
        lib.oracle_kabsch(dp(np.ascontiguousarray(P.reshape(-1))), dp(np.ascontiguousarray(Qn.reshape(-1))),
                          up(np.ones(n, np.uint8)), I(n), dp(R), dp(tt))
        Rs, ts = svd_kabsch(P, Qn)
        np.testing.assert_allclose(R.reshape(3, 3), Rs.T, atol=1e-9)
        np.testing.assert_allclose(tt, ts, atol=1e-8)


def test_kabsch_mask_and_identity_fallback():
    lib = orc()
    rng = np.random.Generator(np.random.Philox(key=6))
    P = rng.standard_normal((10, 3))
    Q = rng.standard_normal((10, 3))
    m = np.zeros(10, np.uint8)
    m[[1, 4]] = 1  # < 3 selected -> identity, zero translation (core.h:104-107)
    R, t = np.zeros(9), np.zeros(3)
    lib.oracle_kabsch(dp(np.ascontiguousarray(P.reshape(-1))), dp(np.ascontiguousarray(Q.reshape(-1))),
                      up(m), I(10), dp(R), dp(t))
    assert np.array_equal(R, np.eye(3).reshape(-1)) and np.array_equal(t, np.zeros(3))
    m[[1, 4, 7, 8]] = 1  # masked rows only
    lib.oracle_kabsch(dp(np.ascontiguousarray(P.reshape(-1))), dp(np.ascontiguousarray(Q.reshape(-1))),
                      up(m), I(10), dp(R), dp(t))
    Rs, ts = svd_kabsch(P[m > 0], Q[m > 0])
    np.testing.assert_allclose(R.reshape(3, 3), Rs.T, atol=1e-10)


def test_tm_score_closed_form():
    """Sum of 1/(1+d^2/d0^2) over valid rows / max(Lnorm,1); the d8 cut
    drops rows with d^2 > d8^2."""
    lib = orc()
    rng = np.random.Generator(np.random.Philox(key=7))
    n = 50
    P = rng.standard_normal((n, 3)) * 10
    Q = rng.standard_normal((n, 3)) * 10
    v = (rng.uniform(size=n) > 0.2).astype(np.uint8)
    Rm = synth.haar_rotation(rng)
    t = rng.uniform(-3, 3, 3)
    R = np.ascontiguousarray(Rm.reshape(-1))
    d2 = np.sum((P @ Rm + t - Q) ** 2, 1)
    d0, L, d8 = 4.0, 60.0, 12.0
    exp0 = np.sum(v / (1 + d2 / d0 ** 2)) / L
    exp8 = np.sum(v * (d2 <= d8 ** 2) / (1 + d2 / d0 ** 2)) / L
    args = (dp(np.ascontiguousarray(P.reshape(-1))), dp(np.ascontiguousarray(Q.reshape(-1))), up(v),
            dp(R), dp(np.ascontiguousarray(t)), D(d0), D(L))
    assert math.isclose(lib.oracle_tm_score(*args, D(0), I(n), I(0)), exp0, rel_tol=1e-13)
    assert math.isclose(lib.oracle_tm_score(*args, D(d8), I(n), I(1)), exp8, rel_tol=1e-13)
    # identical structures, identity transform, Lnorm = n -> exactly 1
    ones = np.ones(n, np.uint8)
    Pf = np.ascontiguousarray(P.reshape(-1))
    s = lib.oracle_tm_score(dp(Pf), dp(Pf), up(ones), dp(np.eye(3).reshape(-1).copy()),
                            dp(np.zeros(3)), D(d0), D(n), D(0), I(n), I(0))
    assert s == 1.0


def _nw(lib, S, gap):
    Np, Nn = S.shape
    H = np.zeros((Np + 1) * (Nn + 1))
    T = np.zeros((Np + 1) * (Nn + 1), np.int8)
    lib.oracle_nw_dp(dp(np.ascontiguousarray(S.reshape(-1))), I(Np), I(Nn), D(gap), dp(H),
                     T.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)))
    ap = np.zeros(max(1, min(Np, Nn)), np.int32)
    an = np.zeros(max(1, min(Np, Nn)), np.int32)
    na = lib.oracle_nw_align(dp(np.ascontiguousarray(S.reshape(-1))), I(Np), I(Nn), D(gap), ip(ap), ip(an))
    return H.reshape(Np + 1, Nn + 1), T.reshape(Np + 1, Nn + 1), ap[:na], an[:na]


def _best_alignment_value(S):
    """Brute force: maximum of sum S[i,j] over all strictly increasing
    matchings (free gaps)."""
    Np, Nn = S.shape
    best = 0.0
    for k in range(1, min(Np, Nn) + 1):
        for I_ in itertools.combinations(range(Np), k):
            for J in itertools.combinations(range(Nn), k):
                best = max(best, sum(S[i, j] for i, j in zip(I_, J)))
    return best


def test_nw_zero_gap_is_optimal_alignment():
    lib = orc()
    rng = np.random.Generator(np.random.Philox(key=8))
    for k in range(60):
        Np, Nn = int(rng.integers(1, 6)), int(rng.integers(1, 6))
        S = rng.integers(0, 3, size=(Np, Nn)).astype(float) if k % 2 else rng.uniform(size=(Np, Nn))
        H, T, ap, an = _nw(lib, S, 0.0)
        assert math.isclose(H[Np, Nn], _best_alignment_value(S), rel_tol=0, abs_tol=1e-12)
        # the traced alignment attains the DP value
        assert math.isclose(sum(S[i, j] for i, j in zip(ap, an)), H[Np, Nn], abs_tol=1e-12)


def test_nw_traceback_consistent_with_gap_rule():
    """For any gap open the traced path re-adds to H[Np,Nn]: each diagonal
    step adds S, each gap step adds gap_open only when it leaves a cell that
    was entered diagonally; ties prefer diagonal, then left, then up."""
    lib = orc()
    rng = np.random.Generator(np.random.Philox(key=9))
    for k in range(80):
        Np, Nn = int(rng.integers(1, 9)), int(rng.integers(1, 9))
        S = rng.integers(0, 2, size=(Np, Nn)).astype(float) * 0.5
        gap = [-0.6, -1.0, 0.0][k % 3]
        H, T, ap, an = _nw(lib, S, gap)
        # recompute H with the stated recurrence in plain python
        Hr = np.zeros((Np + 1, Nn + 1))
        Tr = -np.ones((Np + 1, Nn + 1), int)
        for i in range(1, Np + 1):
            for j in range(1, Nn + 1):
                diag = Hr[i - 1, j - 1] + S[i - 1, j - 1]
                up_ = Hr[i - 1, j] + (gap if Tr[i - 1, j] == 0 else 0.0)
                left = Hr[i, j - 1] + (gap if Tr[i, j - 1] == 0 else 0.0)
                if diag >= up_ and diag >= left:
                    Hr[i, j], Tr[i, j] = diag, 0
                elif left >= up_:
                    Hr[i, j], Tr[i, j] = left, 2
                else:
                    Hr[i, j], Tr[i, j] = up_, 1
        assert np.array_equal(H, Hr) and np.array_equal(T, Tr)
        # walk: the aligned pairs are exactly the diagonal steps from (Np,Nn)
        i, j, pairs = Np, Nn, []
        while i > 0 and j > 0:
            if T[i, j] == 0:
                pairs.append((i - 1, j - 1)); i -= 1; j -= 1
            elif T[i, j] == 1:
                i -= 1
            else:
                j -= 1
        assert pairs[::-1] == list(zip(ap.tolist(), an.tolist()))


def test_seed_windows_closed_form():
    """Global seed + windows: lengths {V, V/2, V/4, V/8, 64, 32, 16, 8, 4}
    within [4, V], starts 0..n-len when that is <= 50 positions, else every
    (n-len+1)//50 and the last; windows with < 3 valid residues dropped."""
    lib = orc()
    for N in (3, 4, 10, 37, 64, 100, 150, 300):
        v = np.ones(N, np.uint8)
        buf = np.zeros(4000 * N, np.uint8)
        K = lib.oracle_seed_masks(I(N), up(v), I(1), I(50), up(buf), I(4000))
        masks = buf[:K * N].reshape(K, N)
        exp = [np.ones(N, np.uint8)]
        lens = sorted({x for x in (N, N // 2, N // 4, N // 8, 64, 32, 16, 8, 4) if 4 <= x <= N}, reverse=True)
        for L in lens:
            total = N - L + 1
            if total <= 50:
                starts = list(range(total))
            else:
                step = max(1, total // 50)
                starts = list(range(0, total, step))
                if starts[-1] != total - 1:
                    starts.append(total - 1)
            for s in starts:
                m = np.zeros(N, np.uint8)
                m[s:s + L] = 1
                exp.append(m)
        assert K == len(exp), N
        assert np.array_equal(masks, np.array(exp))


def test_search_invariants():
    """Whole search: an exact rigid copy scores 1 with the generating
    transform; Lnorm <= 2 and N <= 0 follow indialign.cpp:47-51, 288."""
    p = synth.make_pair(11, n=60, sigma=0.0)
    out = libs.oracle_search([p], serial_bank=1)[0]
    assert abs(out[12] - 1.0) < 1e-12
    R = out[:9].reshape(3, 3)
    np.testing.assert_allclose(p.pred @ R + out[9:12], p.native, atol=1e-8)
    q = synth.make_pair(12, n=40, sigma=1.0)
    q.lnorm = 2.0
    out = libs.oracle_search([q], serial_bank=1)[0]
    assert out[12] == 0.0 and np.array_equal(out[:9], np.eye(3).reshape(-1))


def test_fma_corrected_quotient_is_correctly_rounded():
    """The GPU kernels compute d2 / d0^2 as q = d2 * RN(1/d0^2) corrected by
    one FMA remainder step (ia_common.h div_by).  Markstein's theorem says this
    equals the IEEE quotient; checked on 2e7 random cases, including d0^2 of
    the TM-score formula and adversarial mantissas."""
    import ctypes
    lib = libs.oracle()
    f = lib.oracle_markstein_mismatches
    f.argtypes = [ctypes.c_longlong, ctypes.c_ulonglong]
    f.restype = ctypes.c_longlong
    assert f(20_000_000, 12345) == 0
