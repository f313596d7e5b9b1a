"""The library's rotation solvers (host build of the device source, through
indialign_rotation_from_covariance) against an SVD Kabsch in numpy:
- exact rigid data: both solvers recover the rotation;
- noisy data: both give the SVD optimum (same objective value, det +1);
- the closed form (fast method) agrees with the Jacobi solver (plain method)
  wherever it does not decline, and declines only near repeated top
  eigenvalues (planar / symmetric configurations);
- the rounding identity behind the device Jacobi's direct form of nearly
  converged rotations."""
import ctypes

import numpy as np
import pytest

import libs  # noqa: F401  (sys.path set up by conftest)
import indialign_amd as ia


def _lib():
    try:
        lib = ia.load_library()
    except ia.IndialignError as e:
        pytest.skip(str(e))
    f = lib.indialign_rotation_from_covariance
    f.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    f.restype = ctypes.c_int
    return f


def _solve(f, h, method):
    h = np.ascontiguousarray(h, dtype=np.float64).reshape(9)
    R = np.zeros(9)
    rc = f(h.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), method,
           R.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return rc, R.reshape(3, 3)


def _cov(P, Q):
    return (P - P.mean(0)).T @ (Q - Q.mean(0))  # h[k, d]


def _svd_rotation(H):
    """R with out = p @ R maximizing sum (p' R) . q' (det +1)."""
    U, S, Vt = np.linalg.svd(H)
    d = np.sign(np.linalg.det(U @ Vt))
    D = np.diag([1.0, 1.0, d])
    return U @ D @ Vt


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a*a+b*b-c*c-d*d, 2*(b*c-a*d), 2*(b*d+a*c)],
                     [2*(b*c+a*d), a*a-b*b+c*c-d*d, 2*(c*d-a*b)],
                     [2*(b*d-a*c), 2*(c*d+a*b), a*a-b*b-c*c+d*d]])


def test_exact_rigid_recovery():
    f = _lib()
    rng = np.random.default_rng(0)
    for _ in range(200):
        n = int(rng.integers(3, 60))
        P = rng.normal(size=(n, 3)) * 10
        Rt = _random_rotation(rng)
        Q = P @ Rt + rng.uniform(-50, 50, 3)
        H = _cov(P, Q)
        for method in (0, 1):
            rc, R = _solve(f, H, method)
            if method == 1 and rc == 1:
                continue
            assert rc == 0
            np.testing.assert_allclose(R, Rt, atol=1e-8)


def test_noisy_optimum_and_agreement():
    f = _lib()
    rng = np.random.default_rng(1)
    declined = 0
    for k in range(500):
        n = int(rng.integers(3, 200))
        P = rng.normal(size=(n, 3)) * rng.uniform(1, 30)
        Q = P @ _random_rotation(rng) + rng.normal(size=(n, 3)) * rng.uniform(0.1, 20)
        H = _cov(P, Q)
        Rs = _svd_rotation(H)
        best = np.sum(Rs * H)
        _, Rj = _solve(f, H, 0)
        rc, Rq = _solve(f, H, 1)
        for R in (Rj,) + ((Rq,) if rc == 0 else ()):
            assert abs(np.linalg.det(R) - 1) < 1e-9
            np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-9)
            assert np.sum(R * H) >= best - 1e-9 * max(1.0, abs(best))
        if rc == 0:
            np.testing.assert_allclose(Rq, Rj, atol=1e-7)
        else:
            declined += 1
    assert declined <= 5


def test_degenerate_cases():
    f = _lib()
    # collinear points (rank-1 covariance): any rotation about the line is
    # optimal; the closed form may decline, the Jacobi solver must return a
    # proper rotation with the optimal objective
    P = np.outer(np.arange(10.0), [1.0, 2.0, 3.0])
    Q = P @ _random_rotation(np.random.default_rng(2))
    H = _cov(P, Q)
    best = np.sum(_svd_rotation(H) * H)
    for method in (0, 1):
        rc, R = _solve(f, H, method)
        if rc == 1:
            continue
        assert abs(np.linalg.det(R) - 1) < 1e-9
        assert np.sum(R * H) >= best - 1e-9 * abs(best)
    # zero covariance: the closed form declines, Jacobi returns a rotation
    rc, R = _solve(f, np.zeros((3, 3)), 1)
    assert rc == 1
    rc, R = _solve(f, np.zeros((3, 3)), 0)
    assert rc == 0 and abs(np.linalg.det(R) - 1) < 1e-12


def test_converged_jacobi_rotation_identity():
    """The device Jacobi takes tn = 1/(2|theta|) and c = 1 directly when
    2^27 <= |theta| < 2^511.  That is exact only if, in IEEE binary64 with
    round to nearest, |theta| + sqrt(1 + theta^2) equals 2|theta| and
    1/sqrt(1 + tn^2) equals 1 there; numpy's sqrt and divide are correctly
    rounded, as the device sequences are.  Just below 2^27 the first identity
    fails for some theta, which is why the threshold sits there."""
    rng = np.random.default_rng(5)
    th = np.ldexp(rng.uniform(1.0, 2.0, 400_000), rng.integers(27, 511, 400_000))
    edge = np.array([2.0 ** 27, np.nextafter(2.0 ** 27, np.inf), np.nextafter(2.0 ** 28, 0.0),
                     np.nextafter(2.0 ** 511, 0.0)])
    near = np.nextafter(2.0 ** 27, np.inf) * (1.0 + rng.uniform(0.0, 1e-3, 200_000))
    th = np.concatenate([th, edge, near])
    assert th.min() >= 2.0 ** 27 and th.max() < 2.0 ** 511
    assert np.array_equal(th + np.sqrt(1.0 + th * th), th + th)
    tn = 1.0 / (th + th)
    assert np.all(1.0 / np.sqrt(1.0 + tn * tn) == 1.0)
    # below the threshold the shortcut would change some results
    lo = np.ldexp(rng.uniform(1.0, 2.0, 200_000), rng.integers(20, 27, 200_000))
    assert not np.array_equal(lo + np.sqrt(1.0 + lo * lo), lo + lo)
