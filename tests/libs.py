"""ctypes bindings used by the tests: the product library (through the public
C ABI), the CPU oracle (test infrastructure) and, where it was built, the
reference library compiled from source (oracle/_ref)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libindialign_ref.so")
REF_HARNESS_SO = os.path.join(ROOT, "oracle", "_ref", "libref_harness.so")

c_double_p = ctypes.POINTER(ctypes.c_double)
c_u8_p = ctypes.POINTER(ctypes.c_uint8)
c_int_p = ctypes.POINTER(ctypes.c_int)


class Cfg(ctypes.Structure):
    """NativeConfig, api.h:10-20."""
    _fields_ = [("backend_mode", ctypes.c_int), ("use_fragment_search", ctypes.c_int),
                ("max_iter", ctypes.c_int), ("dp_iter", ctypes.c_int),
                ("max_mem_gb", ctypes.c_double), ("request_cuda", ctypes.c_int),
                ("rescue_score_1", ctypes.c_double), ("rescue_score_2", ctypes.c_double),
                ("early_term_score", ctypes.c_double)]


class Inp(ctypes.Structure):
    """NativePairInput, api.h:22-30."""
    _fields_ = [("length", ctypes.c_int), ("pred_xyz", c_double_p), ("native_xyz", c_double_p),
                ("valid_mask", c_u8_p), ("pred_valid_mask", c_u8_p),
                ("native_valid_mask", c_u8_p), ("lnorm", ctypes.c_double)]


class Res(ctypes.Structure):
    """NativePairResult, api.h:32-36."""
    _fields_ = [("score", ctypes.c_double), ("rotation", ctypes.c_double * 9),
                ("translation", ctypes.c_double * 3)]


def dptr(a):
    return a.ctypes.data_as(c_double_p)


def u8ptr(a):
    return a.ctypes.data_as(c_u8_p)


def iptr(a):
    return a.ctypes.data_as(c_int_p)


class PackedPairs:
    """Keeps numpy buffers alive behind an array of Inp structs."""

    def __init__(self, pairs):
        self.keep = []
        self.inps = (Inp * max(1, len(pairs)))()
        for k, p in enumerate(pairs):
            pf = np.ascontiguousarray(p.pred.reshape(-1), dtype=np.float64)
            nf = np.ascontiguousarray(p.native.reshape(-1), dtype=np.float64)
            v = np.ascontiguousarray(p.valid, dtype=np.uint8)
            pv = np.ascontiguousarray(p.pvalid, dtype=np.uint8)
            nv = np.ascontiguousarray(p.nvalid, dtype=np.uint8)
            self.keep += [pf, nf, v, pv, nv]
            self.inps[k] = Inp(p.n, dptr(pf), dptr(nf), u8ptr(v), u8ptr(pv), u8ptr(nv), p.lnorm)
        self.n = len(pairs)


def results_to_numpy(res, n):
    out = np.zeros((n, 13))
    for k in range(n):
        out[k, :9] = res[k].rotation[:]
        out[k, 9:12] = res[k].translation[:]
        out[k, 12] = res[k].score
    return out


_oracle = None


def oracle():
    global _oracle
    if _oracle is None:
        lib = ctypes.CDLL(ORACLE_SO)
        lib.oracle_default_config.argtypes = [ctypes.POINTER(Cfg)]
        lib.oracle_search_traced.argtypes = [ctypes.POINTER(Inp), ctypes.POINTER(Cfg),
                                             ctypes.POINTER(Res), ctypes.c_int, c_double_p]
        lib.oracle_search_batch.argtypes = [ctypes.c_int, ctypes.POINTER(Inp),
                                            ctypes.POINTER(Cfg), ctypes.POINTER(Res),
                                            ctypes.c_int]
        lib.oracle_tm_score.restype = ctypes.c_double
        _oracle = lib
    return _oracle


def oracle_config(**kw):
    c = Cfg()
    oracle().oracle_default_config(ctypes.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def oracle_search(pairs, cfg=None, serial_bank=1, trace=False):
    lib = oracle()
    cfg = cfg or oracle_config()
    pk = PackedPairs(pairs)
    res = (Res * pk.n)()
    traces = np.full((pk.n, 10, 13), np.nan)
    if trace:
        for k in range(pk.n):
            tr = np.zeros(130)
            lib.oracle_search_traced(ctypes.byref(pk.inps[k]), ctypes.byref(cfg),
                                     ctypes.byref(res[k]), serial_bank, dptr(tr))
            traces[k] = tr.reshape(10, 13)
    else:
        lib.oracle_search_batch(pk.n, pk.inps, ctypes.byref(cfg), res, serial_bank)
    out = results_to_numpy(res, pk.n)
    return (out, traces) if trace else out


def have_ref():
    return os.path.exists(REF_SO) and os.path.exists(REF_HARNESS_SO)


_ref = None


def ref():
    """The reference library compiled from source (oracle/_ref)."""
    global _ref
    if _ref is None:
        lib = ctypes.CDLL(REF_SO)
        lib.indialign_default_config.argtypes = [ctypes.POINTER(Cfg)]
        lib.indialign_tmscore_search.argtypes = [ctypes.POINTER(Inp), ctypes.POINTER(Cfg),
                                                 ctypes.POINTER(Res)]
        lib.indialign_tmscore_search_batch.argtypes = [ctypes.c_int, ctypes.POINTER(Inp),
                                                       ctypes.POINTER(Cfg),
                                                       ctypes.POINTER(Res)]
        _ref = lib
    return _ref


_harness = None


def harness():
    global _harness
    if _harness is None:
        lib = ctypes.CDLL(REF_HARNESS_SO)
        lib.ref_tm_score.restype = ctypes.c_double
        lib.ref_tm_score_no_d8.restype = ctypes.c_double
        _harness = lib
    return _harness


def ref_batch(pairs, cfg=None):
    """Reference batch entry (OpenMP over pairs; serial seed-bank semantics)."""
    lib = ref()
    if cfg is None:
        cfg = Cfg()
        lib.indialign_default_config(ctypes.byref(cfg))
    pk = PackedPairs(pairs)
    res = (Res * pk.n)()
    lib.indialign_tmscore_search_batch(pk.n, pk.inps, ctypes.byref(cfg), res)
    return results_to_numpy(res, pk.n)
