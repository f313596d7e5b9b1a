"""Write the golden fixtures under tests/golden/ from the reference build.

Runs only where the reference sources were compiled into oracle/_ref (see
oracle/Makefile).  The reference ships no golden vectors for its C path
(its tests/test_cpp_smoke.py only checks thresholds), so these fixtures are
the reference's outputs on inputs drawn here:

  functions.npz  single functions (core.h, seeds.h, search.h, nw.h, align.h,
                 rescue.h, ss.h, threading.h) through oracle/_ref/libref_harness.so,
                 OpenMP pinned to 1 thread (serial semantics everywhere).
  search.npz     full indialign_tmscore_search results through
                 oracle/_ref/libindialign_ref.so: the batch entry (serial seed
                 bank) and the single-pair entry with 8 threads (parallel seed
                 bank), inputs stored alongside.

Usage: python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import libs  # noqa: E402
from indialign_amd import synth  # noqa: E402

D = ctypes.c_double
I = ctypes.c_int


def dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def up(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def aos(x):
    return np.ascontiguousarray(x.reshape(-1), dtype=np.float64)


def rigid_pair(rng, n, noise):
    P = rng.standard_normal((n, 3)) * 8.0
    R = synth.haar_rotation(rng)
    t = rng.uniform(-20, 20, 3)
    Q = P @ R.T + t + rng.standard_normal((n, 3)) * noise
    return P, Q


def functions(h):
    out = {}
    rng = np.random.Generator(np.random.Philox(key=777))
    h.ref_set_threads(1)
    # d0 over lengths incl. the piecewise breakpoints
    Ls = np.array([0.5, 1, 2, 3, 11, 11.5, 12, 15, 16, 19, 20, 23, 24, 29, 29.9, 30, 31, 40, 64,
                   100, 150, 300, 1000, 1500, 8192], float)
    d0 = np.zeros((len(Ls), 3))
    for k, L in enumerate(Ls):
        o = np.zeros(3)
        h.ref_d0(D(L), dp(o))
        d0[k] = o
    out["d0_L"], out["d0_out"] = Ls, d0
    # jacobi + rotation on random symmetric / correlation matrices
    As, Aout, Vout, Hs, Rs = [], [], [], [], []
    for k in range(40):
        A = rng.standard_normal((4, 4)) * 10
        A = (A + A.T) / 2
        a = np.ascontiguousarray(A.reshape(-1))
        v = np.zeros(16)
        As.append(a.copy())
        h.ref_jacobi4(dp(a), dp(v))
        Aout.append(a)
        Vout.append(v)
        H = np.ascontiguousarray(rng.standard_normal(9) * 50)
        R = np.zeros(9)
        h.ref_cov_to_rotation(dp(H), dp(R))
        Hs.append(H)
        Rs.append(R)
    out["jac_in"], out["jac_A"], out["jac_V"] = np.array(As), np.array(Aout), np.array(Vout)
    out["rot_H"], out["rot_R"] = np.array(Hs), np.array(Rs)
    # kabsch / tm-score / weighted refine on noisy rigid sets with masks
    kin_P, kin_Q, kin_m, kR, kt, tm8, tm0, wR, wt, ws = [], [], [], [], [], [], [], [], [], []
    for k in range(30):
        n = 40
        P, Q = rigid_pair(rng, n, 0.5 + k * 0.2)
        m = (rng.uniform(size=n) > (0.2 if k % 3 else 0.97)).astype(np.uint8)
        R, t = np.zeros(9), np.zeros(3)
        h.ref_kabsch(dp(aos(P)), dp(aos(Q)), up(m), I(n), dp(R), dp(t))
        kin_P.append(P); kin_Q.append(Q); kin_m.append(m); kR.append(R); kt.append(t)
        tm8.append(h.ref_tm_score(dp(aos(P)), dp(aos(Q)), up(m), dp(R), dp(t), D(3.0), D(n), D(4.0), I(n)))
        tm0.append(h.ref_tm_score_no_d8(dp(aos(P)), dp(aos(Q)), up(m), dp(R), dp(t), D(3.0), D(n), I(n)))
        R2, t2, sc = R.copy(), t.copy(), np.zeros(1)
        h.ref_weighted_tm_refine(dp(aos(P)), dp(aos(Q)), up(m), D(3.0), D(n), I(n), I(4), dp(R2), dp(t2), dp(sc))
        wR.append(R2); wt.append(t2); ws.append(sc[0])
    out.update(kab_P=np.array(kin_P), kab_Q=np.array(kin_Q), kab_m=np.array(kin_m),
               kab_R=np.array(kR), kab_t=np.array(kt), tm_d8=np.array(tm8), tm_no_d8=np.array(tm0),
               wref_R=np.array(wR), wref_t=np.array(wt), wref_s=np.array(ws))
    # NW: random, tie-heavy and zero-heavy score matrices, three gap opens
    nw_in, nw_shape, nw_gap, nw_H, nw_T, nw_ap, nw_an, nw_na = [], [], [], [], [], [], [], []
    for k in range(36):
        Np, Nn = int(rng.integers(1, 25)), int(rng.integers(1, 25))
        kind = k % 3
        if kind == 0:
            S = rng.uniform(size=(Np, Nn))
        elif kind == 1:
            S = rng.integers(0, 2, size=(Np, Nn)).astype(float)
        else:
            S = np.where(rng.uniform(size=(Np, Nn)) < 0.7, 0.0, rng.choice([0.25, 0.5, 1.0], size=(Np, Nn)))
        gap = [-0.6, 0.0, -1.0][(k // 3) % 3]
        S = np.ascontiguousarray(S)
        H = np.zeros((Np + 1) * (Nn + 1))
        T = np.zeros((Np + 1) * (Nn + 1), np.int8)
        h.ref_nw_dp(dp(S), I(Np), I(Nn), D(gap), dp(H), T.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)))
        ap = np.zeros(min(Np, Nn), np.int32)
        an = np.zeros(min(Np, Nn), np.int32)
        na = h.ref_nw_align(dp(S), I(Np), I(Nn), D(gap), ip(ap), ip(an))
        pad = np.zeros(25 * 25)
        pad[:S.size] = S.reshape(-1)
        nw_in.append(pad)
        nw_shape.append((Np, Nn))
        nw_gap.append(gap)
        Hp = np.zeros(26 * 26); Hp[:H.size] = H
        Tp = np.zeros(26 * 26, np.int8); Tp[:T.size] = T
        nw_H.append(Hp); nw_T.append(Tp)
        a1 = np.full(25, -1, np.int32); a1[:na] = ap[:na]
        a2 = np.full(25, -1, np.int32); a2[:na] = an[:na]
        nw_ap.append(a1); nw_an.append(a2); nw_na.append(na)
    out.update(nw_S=np.array(nw_in), nw_shape=np.array(nw_shape), nw_gap=np.array(nw_gap),
               nw_H=np.array(nw_H), nw_T=np.array(nw_T), nw_ap=np.array(nw_ap),
               nw_an=np.array(nw_an), nw_na=np.array(nw_na))
    # seed masks (frag windows with holes) and the full seed set with anchors
    sm_valid, sm_masks, sm_K = [], [], []
    for k in range(12):
        N = int(rng.integers(3, 140))
        valid = (rng.uniform(size=N) > (0.1 if k % 2 else 0.0)).astype(np.uint8)
        cap = 2000
        buf = np.zeros(cap * N, np.uint8)
        ms = [50, 5][k % 2]
        K = h.ref_build_seed_masks(I(N), up(valid), I(1), I(ms), up(buf), I(cap))
        v = np.zeros(140, np.uint8); v[:N] = valid
        sm_valid.append(np.concatenate([[N, ms], v]).astype(np.int32))
        sm_K.append(K)
        sm_masks.append(np.packbits(buf[:K * N]))
    out["seedmask_in"] = np.array(sm_valid)
    out["seedmask_K"] = np.array(sm_K)
    out["seedmask_bits"] = np.array(np.concatenate(sm_masks))
    out["seedmask_bits_len"] = np.array([len(x) for x in sm_masks])
    return out


def pair_fixtures(h):
    """Per-stage functions on a few synthetic pairs (small N, fast on CPU)."""
    out = {}
    h.ref_set_threads(1)
    pairs = [synth.make_pair(s, n_range=(30, 70)) for s in (101, 102, 103)]
    pairs.append(synth.make_pair(104, n_range=(40, 70), sigma_range=(2, 4), domain_motion=True,
                                 drop_frac=0.1))
    pairs.append(synth.make_pair(105, n=48, sigma=0.7))
    keys = ["anch", "topk", "refine", "bank", "det", "dpl", "dpf", "lfn", "lfd", "thr", "ss",
            "sss", "gap"]
    res = {k: [] for k in keys}
    for p in pairs:
        N = p.n
        P, Q = aos(p.pred), aos(p.native)
        v, pv, nv = p.valid, p.pvalid, p.nvalid
        d0 = np.zeros(3)
        h.ref_d0(D(p.lnorm), dp(d0))
        d0_, d0s, sd8 = d0
        cap = 4000
        buf = np.zeros(cap * N, np.uint8)
        nfrag = ctypes.c_int(0)
        K = h.ref_all_seeds(dp(P), dp(Q), up(v), I(N), D(p.lnorm), up(buf), I(cap), ctypes.byref(nfrag))
        res["anch"].append(np.packbits(buf[:K * N]).astype(np.uint8))
        # top-k mask from the identity transform distances
        d2 = np.where(v > 0, np.sum((p.pred - p.native) ** 2, 1), 1e12)
        m = np.zeros(N, np.uint8)
        h.ref_topk_seed(dp(np.ascontiguousarray(d2)), up(v), I(N), D(0.5), up(m))
        res["topk"].append(m)
        # iterative refine of the global seed at d0s
        R, t = np.zeros(9), np.zeros(3)
        h.ref_iterative_seed_refine(dp(P), dp(Q), up(v), up(v), D(d0s), I(N), I(20), dp(R), dp(t))
        res["refine"].append(np.concatenate([R, t]))
        # seed bank over the first 40 seeds x 6 cutoffs, with d8 and early term
        cut = np.ascontiguousarray(d0s * np.array([2.0, 1.5, 1.0, 0.75, 0.5, 0.25]))
        o = np.zeros(13)
        h.ref_evaluate_seed_bank(dp(P), dp(Q), up(v), up(buf), I(min(K, 40)), dp(cut), I(6), D(d0_),
                                 D(sd8), D(p.lnorm), I(N), I(20), I(1), D(0.99), dp(o))
        res["bank"].append(o)
        # detailed search on the identity alignment of valid residues
        idx = np.nonzero(pv & nv)[0].astype(np.int32)
        o = np.zeros(13)
        h.ref_alignment_detailed_search(dp(P), dp(Q), ip(idx), ip(idx), I(len(idx)), D(d0_), D(d0s),
                                        D(sd8), D(p.lnorm), I(N), I(1), I(-1), I(50), dp(o))
        res["det"].append(o)
        # dp_refine lite (2 iters, sd8=100) and full (1 iter) from the refined seed
        o = np.zeros(13)
        h.ref_dp_refine(dp(P), dp(Q), up(v), up(pv), up(nv), dp(R), dp(t), D(d0_), D(d0s), D(100.0),
                        D(p.lnorm), I(N), I(2), I(1), dp(o))
        res["dpl"].append(o)
        o = np.zeros(13)
        h.ref_dp_refine(dp(P), dp(Q), up(v), up(pv), up(nv), dp(R), dp(t), D(d0_), D(d0s), D(sd8),
                        D(p.lnorm), I(N), I(1), I(0), dp(o))
        res["dpf"].append(o)
        for key, dense in (("lfn", 0), ("lfd", 1)):
            o = np.zeros(13)
            h.ref_local_fragment_dp(dp(P), dp(Q), up(v), up(pv), up(nv), D(d0_), D(d0s), D(p.lnorm),
                                    I(N), I(dense), dp(o))
            res[key].append(o)
        o = np.zeros(13)
        h.ref_threading_dp(dp(P), dp(Q), up(v), up(pv), up(nv), D(d0_), D(d0s), D(p.lnorm), I(N), dp(o))
        res["thr"].append(o)
        ssp = ctypes.create_string_buffer(N)
        h.ref_assign_rna_ss(dp(P), up(pv), I(N), ssp)
        ssn = ctypes.create_string_buffer(N)
        h.ref_assign_rna_ss(dp(Q), up(nv), I(N), ssn)
        res["ss"].append(np.frombuffer(ssp.raw + ssn.raw, np.uint8).copy())
        o = np.zeros(13)
        h.ref_evaluate_ss_seeds(dp(P), dp(Q), up(v), up(pv), up(nv), dp(R), dp(t), D(d0_), D(d0s), D(sd8),
                                D(p.lnorm), I(N), I(1), dp(o))
        res["sss"].append(o)
        o = np.zeros(13)
        h.ref_gapless_threading(dp(P), dp(Q), up(pv), up(nv), up(v), D(d0_), D(d0s), D(p.lnorm), I(N), dp(o))
        res["gap"].append(o)
    out["pair_seeds"] = np.array([p.seed for p in pairs])
    for k in keys:
        if k == "anch":
            out["anch_bits"] = np.concatenate(res[k])
            out["anch_bits_len"] = np.array([len(x) for x in res[k]])
        elif k in ("topk", "ss"):
            out[k] = np.concatenate(res[k])
            out[k + "_len"] = np.array([len(x) for x in res[k]])
        else:
            out[k] = np.array(res[k])
    # store K / n_frag plainly
    meta = []
    for p in pairs:
        N = p.n
        buf = np.zeros(4000 * N, np.uint8)
        nfrag = ctypes.c_int(0)
        K = h.ref_all_seeds(dp(aos(p.pred)), dp(aos(p.native)), up(p.valid), I(N), D(p.lnorm),
                            up(buf), I(4000), ctypes.byref(nfrag))
        meta.append((K, nfrag.value))
    out["anch_meta"] = np.array(meta, np.int32)
    return out


def search_fixtures():
    pairs = [synth.single_pair()] + synth.short_pairs(24)
    pairs += [synth.make_pair(s, n_range=(60, 200), sigma_range=(1.0, 6.0), domain_motion=True,
                              drop_frac=0.1) for s in (30001, 30002, 30003, 30004)]
    pairs += [synth.make_pair(40001, n=64, sigma=0.3), synth.make_pair(40002, n=40, sigma=0.0)]
    os.environ["OMP_NUM_THREADS"] = "8"
    batch = libs.ref_batch(pairs)
    lib = libs.ref()
    cfg = libs.Cfg()
    lib.indialign_default_config(ctypes.byref(cfg))
    single = []
    pk = libs.PackedPairs(pairs)
    for k in range(len(pairs)):
        r = libs.Res()
        lib.indialign_tmscore_search(ctypes.byref(pk.inps[k]), ctypes.byref(cfg), ctypes.byref(r))
        single.append(libs.results_to_numpy([r], 1)[0])
    out = {"batch": batch, "single": np.array(single), "seeds": np.array([p.seed for p in pairs])}
    out["n"] = np.array([p.n for p in pairs])
    out["lnorm"] = np.array([p.lnorm for p in pairs])
    out["pred"] = np.concatenate([p.pred for p in pairs])
    out["native"] = np.concatenate([p.native for p in pairs])
    out["valid"] = np.concatenate([p.valid for p in pairs])
    out["pvalid"] = np.concatenate([p.pvalid for p in pairs])
    out["nvalid"] = np.concatenate([p.nvalid for p in pairs])
    return out


def main():
    if not libs.have_ref():
        sys.exit("oracle/_ref not built (needs the reference sources): run make -C oracle")
    h = libs.harness()
    np.savez_compressed(os.path.join(HERE, "functions.npz"), **functions(h), **pair_fixtures(h))
    np.savez_compressed(os.path.join(HERE, "search.npz"), **search_fixtures())
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
