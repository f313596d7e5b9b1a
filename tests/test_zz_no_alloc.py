"""The no-allocation gradient entry (indialign_mlp_grad_noalloc) on the GPU,
through the C ABI with device buffers, against the allocating entry
indialign_mlp_grad on the same inputs: two calls in a row, the second with
fewer rows, and the row limit."""
import ctypes

import numpy as np
import pytest

import indialign_amd as ia

pytestmark = pytest.mark.gpu


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime the library uses."""

    def __init__(self):
        ia.load_library()
        self.h = ctypes.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.ptrs = []

    def to_dev(self, a):
        a = np.ascontiguousarray(a, np.float32)
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), max(4, a.nbytes)) == 0
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        self.ptrs.append(p)
        return p

    def empty(self, n):
        return self.to_dev(np.full(n, np.nan, np.float32))

    def to_host(self, p, n):
        out = np.empty(n, np.float32)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.ptrs:
            self.h.hipFree(p)
        self.ptrs = []


def _problem(n, seed):
    rng = np.random.default_rng(seed)
    p = ia.init_params(seed) * 0.5
    p[90881 - 1] = 0.1
    X = rng.uniform(0, 1, (n, ia.NFEAT)).astype(np.float32)
    t = rng.uniform(0, 1, n).astype(np.float32)
    return p.astype(np.float32), X, t


def test_noalloc_grad_matches_allocating_entry():
    if not ia.gpu_available():
        pytest.fail("no GPU visible to the HIP runtime (GPU tests must run on the MI355X)")
    lib = ia.load_library()
    lib.indialign_mlp_grad_noalloc.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int] + \
        [ctypes.c_void_p] * 3
    lib.indialign_mlp_grad_noalloc.restype = ctypes.c_int
    hip = _Hip()
    try:
        for n, seed in ((3000, 11), (1000, 12)):
            p, X, t = _problem(n, seed)
            g_ref, loss_ref = ia.mlp_grad(p, X, t)
            dp, dX, dt = hip.to_dev(p), hip.to_dev(X), hip.to_dev(t)
            dg, dl = hip.empty(ia.NPARAM), hip.empty(1)
            assert lib.indialign_mlp_grad_noalloc(dp, dX, dt, n, dg, dl, None) == 0
            g = hip.to_host(dg, ia.NPARAM)
            loss = float(hip.to_host(dl, 1)[0])
            rel = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
            print(f"n={n}: loss {loss:.8e} vs {loss_ref:.8e}; grad max |d| "
                  f"{np.abs(g - g_ref).max():.2e}, relative {rel:.2e}", flush=True)
            # the same kernels in the same order on the same inputs
            np.testing.assert_array_equal(g, g_ref)
            assert loss == loss_ref
        # beyond the preallocated rows: refused, outputs untouched
        p, X, t = _problem(8, 13)
        dp, dX, dt = hip.to_dev(p), hip.to_dev(X), hip.to_dev(t)
        dg = hip.to_dev(np.zeros(ia.NPARAM, np.float32))
        dl = hip.to_dev(np.zeros(1, np.float32))
        assert lib.indialign_mlp_grad_noalloc(dp, dX, dt, 4097, dg, dl, None) == -1
        assert lib.indialign_mlp_grad_noalloc(dp, dX, dt, 0, dg, dl, None) == -1
        assert not hip.to_host(dg, ia.NPARAM).any()
    finally:
        hip.free()
