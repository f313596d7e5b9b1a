"""The C-ABI library: it loads without a GPU, exports every function that
include/indialign.h declares, and its structs have the reference layout
(api.h:10-36; x86-64 SysV offsets)."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

import indialign_amd as ia

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "indialign.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(indialign_\w+)\s*\(", text)))


def test_library_exports_header():
    assert os.path.exists(ia.LIB_PATH), "build first: python -m indialign_amd.build"
    names = declared_functions()
    assert {"indialign_default_config", "indialign_tmscore_search", "indialign_tmscore_search_batch",
            "indialign_cuda_available"} <= set(names)
    out = subprocess.run(["nm", "-D", "--defined-only", ia.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    lib = ia.load_library()
    for n in names:
        assert hasattr(lib, n)


def test_default_config_and_no_compute_without_gpu():
    cfg = ia.default_config()
    assert (cfg.backend_mode, cfg.use_fragment_search, cfg.max_iter, cfg.dp_iter) == (0, 1, 20, 1)
    assert (cfg.max_mem_gb, cfg.request_cuda) == (20.0, 0)
    assert (cfg.rescue_score_1, cfg.rescue_score_2, cfg.early_term_score) == (0.5, 0.3, 0.99)
    lib = ia.load_library()
    inp = ia.NativePairInput()
    inp.length = 0
    res = ia.NativePairResult()
    # N <= 0 returns -1 before touching the device (indialign.cpp:288)
    assert lib.indialign_tmscore_search(ctypes.byref(inp), ctypes.byref(cfg), ctypes.byref(res)) == -1
    # no resident batch: batch_run is refused without touching the device
    assert lib.indialign_batch_run(1, ctypes.byref(cfg), ctypes.byref(res), 1, 0) == -5
    assert lib.indialign_set_learned_options(2, 512, 64) == -1
    assert lib.indialign_set_learned_options(0, -1, 64) == -1


LAYOUT_C = r"""
#include <stddef.h>
#include <stdio.h>
#include "indialign.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(NativeConfig),
    offsetof(NativeConfig, backend_mode), offsetof(NativeConfig, use_fragment_search),
    offsetof(NativeConfig, max_iter), offsetof(NativeConfig, dp_iter),
    offsetof(NativeConfig, max_mem_gb), offsetof(NativeConfig, request_cuda),
    offsetof(NativeConfig, rescue_score_1), offsetof(NativeConfig, rescue_score_2),
    offsetof(NativeConfig, early_term_score));
  printf("%zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(NativePairInput),
    offsetof(NativePairInput, length), offsetof(NativePairInput, pred_xyz),
    offsetof(NativePairInput, native_xyz), offsetof(NativePairInput, valid_mask),
    offsetof(NativePairInput, pred_valid_mask), offsetof(NativePairInput, native_valid_mask),
    offsetof(NativePairInput, lnorm));
  printf("%zu %zu %zu %zu\n", sizeof(NativePairResult), offsetof(NativePairResult, score),
    offsetof(NativePairResult, rotation), offsetof(NativePairResult, translation));
  return 0;
}
"""


def test_struct_layout_matches_reference():
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "l.c"), os.path.join(d, "l")
        open(src, "w").write(LAYOUT_C)
        r = subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), src, "-o", exe],
                           capture_output=True, text=True)
        if r.returncode:
            pytest.fail(r.stderr)
        rows = [list(map(int, l.split())) for l in subprocess.run([exe], capture_output=True, text=True).stdout.splitlines()]
    # reference layout (api.h:10-36 on x86-64)
    assert rows[0] == [56, 0, 4, 8, 12, 16, 24, 32, 40, 48]
    assert rows[1] == [56, 0, 8, 16, 24, 32, 40, 48]
    assert rows[2] == [104, 0, 8, 80]
    # the Python mirror agrees
    assert ctypes.sizeof(ia.NativeConfig) == 56 and ia.NativeConfig.early_term_score.offset == 48
    assert ctypes.sizeof(ia.NativePairInput) == 56 and ia.NativePairInput.lnorm.offset == 48
    assert ctypes.sizeof(ia.NativePairResult) == 104 and ia.NativePairResult.translation.offset == 80


def test_device_error_marks_outputs():
    """Without a usable GPU the search returns -2 and every pair with
    length > 0 reads score NaN and the identity (callers that ignore the
    return code, as the reference's batch callers do, see the failure);
    pairs with length <= 0 are left untouched.  Runs where no GPU is visible."""
    if ia.gpu_available():
        pytest.skip("a GPU is visible: the search would run")
    import numpy as np
    lib = ia.load_library()
    cfg = ia.default_config()
    p = np.zeros(30)
    m = np.ones(10, np.uint8)
    inps = (ia.NativePairInput * 2)()
    for k, n in enumerate((10, 0)):
        inps[k] = ia.NativePairInput(n, p.ctypes.data_as(ia.c_double_p), p.ctypes.data_as(ia.c_double_p),
                                     m.ctypes.data_as(ia.c_u8_p), m.ctypes.data_as(ia.c_u8_p),
                                     m.ctypes.data_as(ia.c_u8_p), 10.0)
    res = (ia.NativePairResult * 2)()
    res[0].score = res[1].score = 42.0
    assert lib.indialign_tmscore_search_batch(2, inps, ctypes.byref(cfg), res) == -2
    assert np.isnan(res[0].score) and list(res[0].rotation) == [1, 0, 0, 0, 1, 0, 0, 0, 1]
    assert res[1].score == 42.0
    r = ia.NativePairResult()
    assert lib.indialign_tmscore_search(ctypes.byref(inps[0]), ctypes.byref(cfg), ctypes.byref(r)) == -2
    assert np.isnan(r.score)
