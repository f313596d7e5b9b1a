"""Deterministic synthetic RNA C1' pairs for the TM-score search workloads.

Native chain: alternating A-form helix segments (radius 8.7 A, rise 2.81 A/nt,
twist 32.7 deg, segment length U{8,16}) joined by random-walk loops (5.9 A
steps, turn angle U(30,120) deg, uniform azimuth, loop length U{2,6}), then
centred.  Pred chain: the native under a Haar-random rotation, a U(-50,50) A
translation per axis and i.i.d. Gaussian noise.  Every pair draws from its own
counter-based generator (numpy Philox keyed by the pair seed), so pair p of a
workload is the same on every machine and independent of batch size.

Workloads (sizes as the bench and the parity tests use them):
  single   1 pair, seed 0, N=64, sigma=3 A, all masks valid, Lnorm=N
  short    pairs with seeds 1..n, N~U{40,300}, sigma~U(0.5,14) A
  mixed    pairs with seeds 50001.., N log-uniform over [27, 1500],
           sigma~U(0.5,14) A, all masks valid
  long     pairs with seeds 20001.., N~U{400,1500}, sigma~U(2,10) A, 20% of
           helix segments moved rigidly (20-60 deg about a random axis through
           the segment centroid, 5-15 A shift), 10% of pred-valid and of
           native-valid residues zeroed independently, valid = pv & nv,
           Lnorm = native-valid count
"""
from __future__ import annotations

import dataclasses
import numpy as np

HELIX_RADIUS = 8.7
HELIX_RISE = 2.81
HELIX_TWIST = np.deg2rad(32.7)
LOOP_STEP = 5.9


@dataclasses.dataclass
class Pair:
    pred: np.ndarray      # (N, 3) float64
    native: np.ndarray    # (N, 3) float64
    valid: np.ndarray     # (N,) uint8
    pvalid: np.ndarray    # (N,) uint8
    nvalid: np.ndarray    # (N,) uint8
    lnorm: float
    seed: int
    sigma: float

    @property
    def n(self) -> int:
        return int(self.native.shape[0])


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(key=int(seed)))


def _unit(v):
    return v / np.linalg.norm(v)


def _frame_from_z(z, rng):
    """Orthonormal frame whose third column is z, random spin about z."""
    z = _unit(z)
    a = rng.standard_normal(3)
    x = _unit(a - np.dot(a, z) * z)
    y = np.cross(z, x)
    return np.stack([x, y, z], axis=1)


def haar_rotation(rng) -> np.ndarray:
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def _axis_angle(axis, ang):
    k = _unit(axis)
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * (K @ K)


def _turn(d, rng):
    """New unit direction at angle U(30,120) deg from d, uniform azimuth."""
    ang = np.deg2rad(rng.uniform(30.0, 120.0))
    f = _frame_from_z(d, rng)
    return _unit(f @ np.array([np.sin(ang), 0.0, np.cos(ang)]))


def make_native(n: int, rng, segments_out: list | None = None) -> np.ndarray:
    out = np.zeros((n, 3))
    k = 0
    pos = np.zeros(3)
    d = _unit(rng.standard_normal(3))
    first = True
    while k < n:
        L = int(rng.integers(8, 17))
        frame = _frame_from_z(d, rng)
        t = np.arange(L)
        local = np.stack([HELIX_RADIUS * np.cos(t * HELIX_TWIST),
                          HELIX_RADIUS * np.sin(t * HELIX_TWIST),
                          HELIX_RISE * t], axis=1)
        pts = local @ frame.T
        start = pos if first else pos + LOOP_STEP * d
        pts = pts - pts[0] + start
        m = min(L, n - k)
        out[k:k + m] = pts[:m]
        if segments_out is not None and m > 0:
            segments_out.append((k, k + m))
        k += m
        first = False
        if k >= n:
            break
        pos = pts[m - 1]
        d = _unit(pts[m - 1] - pts[m - 2]) if m >= 2 else d
        for _ in range(int(rng.integers(2, 7))):
            if k >= n:
                break
            d = _turn(d, rng)
            pos = pos + LOOP_STEP * d
            out[k] = pos
            k += 1
    return out - out.mean(axis=0)


def make_pair(seed: int, n: int | None = None, sigma: float | None = None,
              n_range=(40, 300), sigma_range=(0.5, 14.0), domain_motion=False,
              drop_frac=0.0) -> Pair:
    rng = _rng(seed)
    if n is None:
        n = int(rng.integers(n_range[0], n_range[1] + 1))
    if sigma is None:
        sigma = float(rng.uniform(sigma_range[0], sigma_range[1]))
    segs: list = []
    native = make_native(n, rng, segs)
    moved = native.copy()
    if domain_motion:
        for (a, b) in segs:
            if rng.uniform() < 0.2:
                c = moved[a:b].mean(axis=0)
                Rm = _axis_angle(rng.standard_normal(3), np.deg2rad(rng.uniform(20, 60)))
                shift = _unit(rng.standard_normal(3)) * rng.uniform(5, 15)
                moved[a:b] = (moved[a:b] - c) @ Rm.T + c + shift
    R = haar_rotation(rng)
    t = rng.uniform(-50.0, 50.0, size=3)
    pred = moved @ R.T + t + rng.standard_normal((n, 3)) * sigma
    if drop_frac > 0:
        pv = (rng.uniform(size=n) >= drop_frac).astype(np.uint8)
        nv = (rng.uniform(size=n) >= drop_frac).astype(np.uint8)
    else:
        pv = np.ones(n, np.uint8)
        nv = np.ones(n, np.uint8)
    valid = (pv & nv).astype(np.uint8)
    lnorm = float(nv.sum()) if drop_frac > 0 else float(n)
    return Pair(np.ascontiguousarray(pred), np.ascontiguousarray(native),
                valid, pv, nv, lnorm, int(seed), float(sigma))


def indel_pair(seed: int, n_range=(40, 160), sigma_range=(0.5, 6.0)) -> Pair:
    """A pair whose residue correspondence is not the identity: pred is the
    native with k = U{3,10} residues deleted at a and k random-walk residues
    (5.9 A steps) inserted at b > a, so residues between the two are shifted
    by k; then the Haar rotation, translation and noise of make_pair.  Same
    length N on both sides, all masks valid, Lnorm = N."""
    rng = _rng(seed + 9_000_000)
    n = int(rng.integers(n_range[0], n_range[1] + 1))
    sigma = float(rng.uniform(sigma_range[0], sigma_range[1]))
    native = make_native(n, rng)
    k = int(rng.integers(3, 11))
    a = int(rng.integers(n // 8, n // 2))
    b = int(rng.integers(a + k + 4, n - 2))
    ins = [native[b - 1]]
    d = _unit(rng.standard_normal(3))
    for _ in range(k):
        d = _turn(d, rng)
        ins.append(ins[-1] + LOOP_STEP * d)
    moved = np.concatenate([native[:a], native[a + k:b], np.array(ins[1:]), native[b:]])
    R = haar_rotation(rng)
    t = rng.uniform(-50.0, 50.0, size=3)
    pred = moved @ R.T + t + rng.standard_normal((n, 3)) * sigma
    m = np.ones(n, np.uint8)
    return Pair(np.ascontiguousarray(pred), np.ascontiguousarray(native), m, m.copy(), m.copy(),
                float(n), int(seed), sigma)


def single_pair() -> Pair:
    return make_pair(0, n=64, sigma=3.0)


def short_pairs(count: int, first_seed: int = 1) -> list[Pair]:
    return [make_pair(s) for s in range(first_seed, first_seed + count)]


def mixed_pairs(count: int, first_seed: int = 50001) -> list[Pair]:
    """Pairs of uneven length: N log-uniform over [27, 1500] (as many pairs
    of 27-100 nt as of 400-1500 nt), sigma~U(0.5,14) A, all masks valid."""
    out = []
    for s in range(first_seed, first_seed + count):
        u = float(_rng(s + 7_000_000).uniform())
        n = int(round(np.exp(np.log(27.0) + u * (np.log(1500.0) - np.log(27.0)))))
        out.append(make_pair(s, n=n))
    return out


def long_pairs(count: int, first_seed: int = 20001) -> list[Pair]:
    return [make_pair(s, n_range=(400, 1500), sigma_range=(2.0, 10.0),
                      domain_motion=True, drop_frac=0.10)
            for s in range(first_seed, first_seed + count)]
