"""Launch tuning of the chain and seed-bank kernels (indialign_set_launch_tuning and the
tuning file the autotuner writes): range checks and file parsing run on the
CPU (the library reads the file before touching a device); on the GPU every
shape must give the same results, bit for bit."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import indialign_amd as ia
from indialign_amd import autotune, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROBE = r"""
import sys
sys.path.insert(0, %r)
import indialign_amd as ia
print(list(ia.get_launch_tuning()))
"""


def read_tuning_in_subprocess(tuning_path):
    env = dict(os.environ, INDIALIGN_TUNING=tuning_path)
    r = subprocess.run([sys.executable, "-c", PROBE % ROOT], capture_output=True, text=True, env=env)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1]), r.stderr


def test_setter_range_checks():
    before = ia.get_launch_tuning()
    for bad in [(1, 0, 0, 0), (4, 0, 0, 0), (3, -1, 0, 0), (3, 65, 0, 0), (3, 0, 3, 0),
                (2, 0, -1, 0), (3, 0, 0, -1), (3, 0, 0, 65)]:
        with pytest.raises(ia.IndialignError):
            ia.set_launch_tuning(*bad)
        assert ia.get_launch_tuning() == before
    ia.set_launch_tuning(2, 8, 2, 5)
    assert ia.get_launch_tuning() == (2, 8, 2, 5)
    ia.set_launch_tuning(*before)


def test_tuning_file_read_by_library(tmp_path):
    p = tmp_path / "t.txt"
    autotune.write_tuning(str(p), (2, 4, 1, 3), "test")
    got, _ = read_tuning_in_subprocess(str(p))
    assert got == [2, 4, 1, 3]
    # a file may set some keys only
    p.write_text("# partial\nbank_pass_cap 4\n")
    got, _ = read_tuning_in_subprocess(str(p))
    assert got == [3, 0, 0, 4]  # (default cap 5 replaced)
    # malformed or out-of-range files are reported and ignored
    p.write_text("chain_waves 7\n")
    got, err = read_tuning_in_subprocess(str(p))
    assert got == [3, 0, 0, 5] and "malformed" in err
    p.write_text("chain_waves\n")
    got, err = read_tuning_in_subprocess(str(p))
    assert got == [3, 0, 0, 5] and "malformed" in err
    # a missing file keeps the defaults
    got, _ = read_tuning_in_subprocess(str(tmp_path / "absent.txt"))
    assert got == [3, 0, 0, 5]


def test_candidate_shapes():
    c = autotune.candidates()
    assert c[0] == (3, 0, 0) and len(set(c)) == len(c)
    for w, s, t in c:
        assert w in (2, 3) and 0 <= s <= 64 and t in (0, 1, 2)
    assert autotune.BANK_CAPS[0] == 0 and all(0 <= x <= 64 for x in autotune.BANK_CAPS)


@pytest.mark.gpu
def test_launch_shapes_give_identical_results():
    """Short pairs, one long masked pair and mixed-length pairs, plain and
    fast methods: every launch shape (chain kernel variant, slots, trace
    placement, seed-bank pass cap with its compacted continuation launch,
    chain length classes) returns the same scores, rotations and
    translations, and the same device-counted algorithmic work (the work
    depends on the search alone, so a counter lost or counted twice under
    some shape shows here)."""
    if not ia.gpu_available():
        pytest.fail("no GPU visible to the HIP runtime")
    before = ia.get_launch_tuning()
    # lengths in every chain length class (<= 160, <= 400, <= 900, longer)
    pairs = synth.short_pairs(24) + synth.long_pairs(1) + synth.mixed_pairs(6, first_seed=50101)
    try:
        for method in (0, 1):
            ref = ref_work = None
            for shape in [(3, 0, 0, 0), (2, 0, 0, 0), (3, 4, 2, 0), (2, 8, 1, 2), (3, 0, 1, 3),
                          (3, 0, 0, 1), (3, 0, 0, 6), (3, 0, 0, 5, 1), (3, 0, 0, 5, 2),
                          (3, 0, 2, 5, 4)]:
                ia.set_launch_tuning(*shape[:4])
                ia.set_chain_buckets(shape[4] if len(shape) > 4 else 1)
                out = ia.tmscore_search_batch(pairs, method=method)
                st = ia.last_stats()
                work = (st["bank_ops"], st["refine_passes"], st["chain_ops"], st["nw_cells"])
                print(f"method {method} shape {shape}: mean score {out[0].mean():.12f}, "
                      f"work {work}", flush=True)
                # the chain kernel counts its work in the FP64 methods only
                assert st["bank_ops"] > 0 and (st["chain_ops"] > 0) == (method == 0)
                if ref is None:
                    ref, ref_work = out, work
                    continue
                for a, b in zip(out, ref):
                    assert np.array_equal(a, b), (method, shape)
                assert work == ref_work, (method, shape)
    finally:
        ia.set_launch_tuning(*before)
        ia.set_chain_buckets(1)
