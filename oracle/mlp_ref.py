"""ORACLE -- test infrastructure only.

Float64 numpy restatement of the learned policy's MLP (12 -> 32 -> 32 -> 3, ReLU, softmax
cross-entropy) with the blob layout of include/msc.h (W1[32][12] b1[32] W2[32][32] b2[32]
W3[3][32] b3[3] mean[12] std[12]).  The device forward / backward kernels
(connectivity_amd/csrc/msc_mlp.hip) are checked against it, and it is checked against
central finite differences in tests/test_policy.py.
"""
import numpy as np

IN, H, OUT = 12, 32, 3
PARAMS = IN * H + H + H * H + H + H * OUT + OUT
BLOB = PARAMS + 2 * IN


def unpack(blob):
    b = np.asarray(blob, dtype=np.float64)
    o = 0
    W1 = b[o:o + H * IN].reshape(H, IN); o += H * IN
    b1 = b[o:o + H]; o += H
    W2 = b[o:o + H * H].reshape(H, H); o += H * H
    b2 = b[o:o + H]; o += H
    W3 = b[o:o + OUT * H].reshape(OUT, H); o += OUT * H
    b3 = b[o:o + OUT]; o += OUT
    mean = b[o:o + IN]; o += IN
    std = b[o:o + IN]
    return W1, b1, W2, b2, W3, b3, mean, std


def forward(blob, x):
    W1, b1, W2, b2, W3, b3, mean, std = unpack(blob)
    xs = (np.asarray(x, dtype=np.float64) - mean) / std
    a1 = xs @ W1.T + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0.0)
    z = h2 @ W3.T + b3
    return xs, a1, h1, a2, h2, z


def loss_and_grad(blob, x, y, w=None):
    """sum_b w_b * CE(softmax(z_b), y_b) and its gradient w.r.t. the PARAMS weights."""
    xs, a1, h1, a2, h2, z = forward(blob, x)
    y = np.asarray(y)
    n = z.shape[0]
    w = np.ones(n) if w is None else np.asarray(w, dtype=np.float64)
    zmax = z.max(axis=1, keepdims=True)
    e = np.exp(z - zmax)
    p = e / e.sum(axis=1, keepdims=True)
    loss = float(np.sum(w * (np.log(e.sum(axis=1)) - (z[np.arange(n), y] - zmax[:, 0]))))
    dz = p.copy()
    dz[np.arange(n), y] -= 1.0
    dz *= w[:, None]
    W1, b1, W2, b2, W3, b3, _, _ = unpack(blob)
    gW3 = dz.T @ h2
    gb3 = dz.sum(axis=0)
    d2 = (dz @ W3) * (a2 > 0)
    gW2 = d2.T @ h1
    gb2 = d2.sum(axis=0)
    d1 = (d2 @ W2) * (a1 > 0)
    gW1 = d1.T @ xs
    gb1 = d1.sum(axis=0)
    grad = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, grad
