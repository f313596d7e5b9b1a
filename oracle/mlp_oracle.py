"""CPU restatement of the work-item ranker MLP (forward, MSE backward, Adam).

TEST INFRASTRUCTURE ONLY: tests/ use it as the checker of the GPU kernels
(indialign_amd/csrc/k_learn.hip, mlp.hip); the product never imports it.
The ranker has no counterpart in the reference; its definition is the
library's own (include/indialign.h): 96 -> 256 -> 256 -> 1 with ReLU, plus
descriptor column 31 added to the output, MSE loss, every matrix-product
operand rounded to BF16 (round to nearest even) and products accumulated in
wide precision.

Pins (tests/test_mlp_oracle.py):
  - backward() with rounding off against central finite differences of
    forward() in float64 (a dropped term, wrong transpose or missing ReLU
    mask fails);
  - bf16() against torch's float32 -> bfloat16 conversion;
  - adam() against the textbook update written out for one scalar.
"""
from __future__ import annotations

import numpy as np

NFEAT, NHID = 96, 256
OFF_W1 = 0
OFF_B1 = OFF_W1 + NHID * NFEAT
OFF_W2 = OFF_B1 + NHID
OFF_B2 = OFF_W2 + NHID * NHID
OFF_W3 = OFF_B2 + NHID
OFF_B3 = OFF_W3 + NHID
NPARAM = OFF_B3 + 1
SKIP_COL = 31  # the output adds this descriptor column (FP32, not rounded)


def bf16(a):
    """Round float32 values to the nearest BF16 (ties to even), as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).reshape(np.shape(a))


def unpack(p):
    p = np.asarray(p)
    W1 = p[OFF_W1:OFF_B1].reshape(NHID, NFEAT)
    b1 = p[OFF_B1:OFF_W2]
    W2 = p[OFF_W2:OFF_B2].reshape(NHID, NHID)
    b2 = p[OFF_B2:OFF_W3]
    w3 = p[OFF_W3:OFF_B3]
    b3 = p[OFF_B3]
    return W1, b1, W2, b2, w3, b3


def _mm(A, B, rounding):
    if rounding:
        A = bf16(A.astype(np.float32))
        B = bf16(B.astype(np.float32))
    return A.astype(np.float64) @ B.astype(np.float64)


def forward(p, X, rounding=True):
    """y (n,) and the activations (h1, h2)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    dt = np.float32 if rounding else np.float64
    h1 = np.maximum(_mm(X, W1.T, rounding) + b1, 0).astype(dt)
    h2 = np.maximum(_mm(h1, W2.T, rounding) + b2, 0).astype(dt)
    y = (_mm(h2, w3.reshape(-1, 1), rounding)[:, 0] + b3 + X[:, SKIP_COL]).astype(dt)
    return y, (h1, h2)


def backward(p, X, t, rounding=True):
    """(grad flat in the parameter layout, loss) of mean (y - t)^2."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    dt = np.float32 if rounding else np.float64
    n = X.shape[0]
    y, (h1, h2) = forward(p, X, rounding)
    d = (y.astype(np.float64) - t)
    loss = float(np.mean(d * d))
    dy = (2.0 * d / n).astype(dt)
    g = np.zeros(NPARAM, np.float64)
    g[OFF_W3:OFF_B3] = _mm(h2.T, dy.reshape(-1, 1), rounding)[:, 0]
    g[OFF_B3] = _mm(dy.reshape(1, -1), np.ones((n, 1), dt), rounding)[0, 0]
    dh2 = np.where(h2 > 0, dy.reshape(-1, 1) * w3.reshape(1, -1), 0).astype(dt)
    g[OFF_W2:OFF_B2] = _mm(dh2.T, h1, rounding).reshape(-1)
    g[OFF_B2:OFF_W3] = _mm(dh2.T, np.ones((n, 1), dt), rounding)[:, 0]
    dh1 = np.where(h1 > 0, _mm(dh2, W2, rounding), 0).astype(dt)
    g[OFF_W1:OFF_B1] = _mm(dh1.T, X, rounding).reshape(-1)
    g[OFF_B1:OFF_W2] = _mm(dh1.T, np.ones((n, 1), dt), rounding)[:, 0]
    return g, loss


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (bias-corrected); returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
