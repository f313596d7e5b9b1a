"""The CPU oracle of the work-item ranker (oracle/mlp_oracle.py) pinned by
mathematics that does not restate it: finite differences of its own forward
pass in float64, torch's float32 -> bfloat16 conversion, and the textbook
Adam update."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as mo  # noqa: E402


def test_layout_matches_header():
    hdr = open(os.path.join(ROOT, "include", "indialign.h")).read()
    assert f"#define INDIALIGN_MLP_NPARAM {mo.NPARAM}" in hdr
    assert f"#define INDIALIGN_MLP_NFEAT {mo.NFEAT}" in hdr
    kh = open(os.path.join(ROOT, "indialign_amd", "csrc", "ia_kernels.h")).read()
    assert f"constexpr int kSkipCol = {mo.SKIP_COL};" in kh


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = rng.normal(0, 0.1, mo.NPARAM)
    X = rng.uniform(0, 1, (7, mo.NFEAT))
    t = rng.uniform(0, 1, 7)
    g, loss = mo.backward(p, X, t, rounding=False)

    def L(q):
        y, _ = mo.forward(q, X, rounding=False)
        return float(np.mean((y - t) ** 2))

    assert abs(loss - L(p)) < 1e-12
    # every block, including the scalar output bias
    idx = [mo.OFF_B3, mo.OFF_W3 + 5, mo.OFF_B2 + 17, mo.OFF_B1 + 9, mo.OFF_W1 + 11 * 96 + 40,
           mo.OFF_W2 + 200 * 256 + 3]
    idx += list(rng.integers(0, mo.NPARAM, 40))
    for i in idx:
        e = 1e-6
        qp, qm = p.copy(), p.copy()
        qp[i] += e
        qm[i] -= e
        fd = (L(qp) - L(qm)) / (2 * e)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])


def test_bf16_matches_torch():
    import torch
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-3, 2000), rng.normal(0, 1e4, 2000),
                        [0.0, -0.0, 1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9]]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(mo.bf16(x), ref)


def test_adam_textbook():
    p, g, m, v = 0.5, 0.2, 0.0, 0.0
    lr, b1, b2, eps = 1e-2, 0.9, 0.999, 1e-8
    q, m1, v1 = mo.adam(np.array([p]), np.array([g]), np.array([m]), np.array([v]), 1, lr)
    # first step: m_hat = g, v_hat = g^2 -> step = lr * g / (|g| + eps)
    assert abs(q[0] - (p - lr * g / (abs(g) + eps))) < 1e-15
    q2, _, _ = mo.adam(q, np.array([-0.1]), m1, v1, 2, lr)
    m2 = b1 * m1[0] + (1 - b1) * -0.1
    v2 = b2 * v1[0] + (1 - b2) * 0.01
    step = lr * (m2 / (1 - b1 ** 2)) / (np.sqrt(v2 / (1 - b2 ** 2)) + eps)
    assert abs(q2[0] - (q[0] - step)) < 1e-15


def test_forward_matches_torch_linear():
    """The oracle's forward pass against torch.nn.functional.linear / relu on
    the same BF16-rounded operands (torch's own conversion), accumulated in
    float64: a transposed weight or a misplaced bias or skip term fails."""
    import torch
    import torch.nn.functional as F
    rng = np.random.default_rng(21)
    p = (rng.normal(0, 0.2, mo.NPARAM)).astype(np.float32)
    X = rng.uniform(0, 1, (33, mo.NFEAT)).astype(np.float32)
    W1, b1, W2, b2, w3, b3 = mo.unpack(p)

    def r(a):  # BF16 operand, as float64
        return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(torch.bfloat16).to(torch.float64)

    b = lambda a: torch.from_numpy(np.asarray(a, np.float64))
    h1 = F.relu(F.linear(r(X), r(W1), b(b1))).to(torch.float32)
    h2 = F.relu(F.linear(r(h1.numpy()), r(W2), b(b2))).to(torch.float32)
    y = F.linear(r(h2.numpy()), r(w3.reshape(1, -1)), b(np.array([b3]))).reshape(-1)
    y = (y + b(X[:, mo.SKIP_COL])).to(torch.float32).numpy()
    yo, (h1o, h2o) = mo.forward(p, X)
    np.testing.assert_allclose(h1o, h1.numpy(), rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(h2o, h2.numpy(), rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(yo, y, rtol=1e-6, atol=1e-6)


def test_forward_hand_computed_two_features():
    """A sparse net on two input features, values exact in BF16:
    h1 = relu([x0, x1, x0 - x1 + 0.5]), h2 = relu([h1_0 + h1_1, h1_2]),
    y = 2 h2_0 - h2_1 + 0.25 + x31."""
    p = np.zeros(mo.NPARAM, np.float32)
    W1 = p[mo.OFF_W1:mo.OFF_B1].reshape(mo.NHID, mo.NFEAT)
    W1[0, 0] = 1.0
    W1[1, 1] = 1.0
    W1[2, 0], W1[2, 1] = 1.0, -1.0
    p[mo.OFF_B1 + 2] = 0.5
    W2 = p[mo.OFF_W2:mo.OFF_B2].reshape(mo.NHID, mo.NHID)
    W2[0, 0] = W2[0, 1] = 1.0
    W2[1, 2] = 1.0
    p[mo.OFF_W3] = 2.0
    p[mo.OFF_W3 + 1] = -1.0
    p[mo.OFF_B3] = 0.25
    X = np.zeros((2, mo.NFEAT), np.float32)
    X[0, :2] = 1.5, 0.25
    X[1, :2] = 0.25, 1.5
    X[:, mo.SKIP_COL] = 0.125
    y, (h1, h2) = mo.forward(p, X)
    # row 0: h1 = [1.5, 0.25, 1.75], h2 = [1.75, 1.75], y = 3.5 - 1.75 + 0.25 + 0.125
    # row 1: h1 = [0.25, 1.5, 0 (clipped)], h2 = [1.75, 0], y = 3.5 + 0.25 + 0.125
    np.testing.assert_array_equal(h1[:, :3], [[1.5, 0.25, 1.75], [0.25, 1.5, 0.0]])
    np.testing.assert_array_equal(h2[:, :2], [[1.75, 1.75], [1.75, 0.0]])
    np.testing.assert_array_equal(y, [2.125, 3.875])
