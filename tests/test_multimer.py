"""Multimer chain-permutation scoring (indialign_amd/multimer.py): grouping,
candidate permutations (degenerate copies, Hungarian, ICP), and the best
candidate's score, with the CPU oracle as the search on the CPU and the
library's batched search on the GPU."""
import numpy as np
import pytest

import libs
from indialign_amd import multimer as mm
from indialign_amd import synth


def test_group_indices_first_seen_order():
    chains = ["A", "A", "B", "A", "A", "B"]
    copies = [1, 1, 1, 2, 2, 1]
    blocks, labels = mm.group_indices(chains, copies)
    assert [b.tolist() for b in blocks] == [[0, 1], [2, 5], [3, 4]]
    assert labels == ["A", "B", "A"]


def test_group_targets_by_length():
    lens = {"a": 10, "b": 19, "c": 21, "d": 50, "e": 100, "f": 101}
    assert mm.group_targets_by_length(list(lens), lens) == [["a", "b"], ["c"], ["d", "e"], ["f"]]


def _dimer(seed, n=40, sigma=0.3, same_shape=True):
    """Native: two copies of chain A (copy 2 is copy 1 moved rigidly, or a
    different chain when same_shape is False); pred: the native under a
    rigid motion with noise, its two copies listed in swapped order."""
    rng = np.random.default_rng(seed)
    c1 = synth.make_native(n, rng)
    if same_shape:
        Rm = synth.haar_rotation(rng)
        c2 = c1 @ Rm.T + np.array([35.0, 0.0, 0.0])
    else:
        c2 = synth.make_native(n, rng) + np.array([35.0, 0.0, 0.0])
    native = np.concatenate([c1, c2])
    R = synth.haar_rotation(rng)
    pred_true = native @ R.T + rng.uniform(-20, 20, 3) + rng.normal(0, sigma, native.shape)
    pred = np.concatenate([pred_true[n:], pred_true[:n]])  # copies swapped
    blocks, labels = mm.group_indices(["A"] * (2 * n), [1] * n + [2] * n)
    return pred, pred_true, native, blocks, labels


def test_candidates_degenerate_and_hungarian():
    pred, _, native, blocks, labels = _dimer(1, same_shape=True)
    v = np.ones(len(native), np.uint8)
    c = mm.frame_candidates(pred, native, v, v, blocks, labels)
    assert len(c) == 2  # identical copies: both orders
    pred2, _, native2, blocks2, labels2 = _dimer(2, same_shape=False)
    c2 = mm.frame_candidates(pred2, native2, v, v, blocks2, labels2)
    assert len(c2) == 1  # different shapes: the Hungarian assignment, which un-swaps
    n = len(native2) // 2
    assert c2[0].tolist() == list(range(n, 2 * n)) + list(range(n))


def test_icp_candidates_include_the_true_order():
    """Six interchangeable copies placed on a ring: the prediction lists them
    in a shuffled order; ICP on the centroids recovers the order that
    undoes the shuffle."""
    rng = np.random.default_rng(3)
    unit = synth.make_native(12, rng)
    ring = [np.array([40 * np.cos(a), 40 * np.sin(a), 7.0 * k]) for k, a in
            enumerate(np.linspace(0, 2 * np.pi, 6, endpoint=False))]
    native = np.concatenate([unit + c for c in ring])
    shuffle = np.array([3, 0, 5, 1, 4, 2])
    R = synth.haar_rotation(rng)
    pred = np.concatenate([native[12 * s:12 * s + 12] for s in shuffle]) @ R.T + 5.0
    blocks, labels = mm.group_indices(["A"] * 72, np.repeat(np.arange(6), 12))
    v = np.ones(72, np.uint8)
    cands = mm.icp_candidates(pred, native, v, v, blocks, list(range(6)))
    assert tuple(int(x) for x in np.argsort(shuffle)) in cands


def test_best_candidate_score_with_oracle_search():
    pred, pred_true, native, blocks, labels = _dimer(5)
    v = np.ones(len(native), np.uint8)

    def oracle_search(tuples):
        pairs = [synth.Pair(p, q, a, b, c, L, 0, 0.0) for p, q, a, b, c, L in tuples]
        out = libs.oracle_search(pairs, serial_bank=1)
        return out[:, 12], out[:, :9], out[:, 9:12]

    s = mm.score_multimer(pred[None], v[None], native[None], v[None], blocks, labels,
                          search=oracle_search)
    direct = libs.oracle_search([synth.Pair(pred_true, native, v, v, v, float(len(v)), 0, 0.0)],
                                serial_bank=1)[0, 12]
    swapped = libs.oracle_search([synth.Pair(pred, native, v, v, v, float(len(v)), 0, 0.0)],
                                 serial_bank=1)[0, 12]
    assert s == direct and s > swapped + 0.2


@pytest.mark.gpu
def test_multimer_on_gpu_matches_oracle():
    import indialign_amd as ia
    if not ia.gpu_available():
        pytest.fail("no GPU visible to the HIP runtime (GPU tests must run on the MI355X)")
    pred, pred_true, native, blocks, labels = _dimer(7)
    v = np.ones(len(native), np.uint8)
    # two pred frames x one native frame
    P = np.stack([pred, pred_true])
    s = mm.score_multimer(P, np.stack([v, v]), native[None], v[None], blocks, labels)
    ref = libs.oracle_search([synth.Pair(pred_true, native, v, v, v, float(len(v)), 0, 0.0)],
                             serial_bank=1)[0, 12]
    print(f"multimer best {s:.15f} oracle (true order) {ref:.15f}", flush=True)
    assert abs(s - ref) <= 1e-6
