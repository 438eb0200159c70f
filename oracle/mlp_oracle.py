"""FP64 numpy restatement of the direction-policy MLP (TEST INFRASTRUCTURE
ONLY): the checker of liblgmi's training kernels (lightgraphs.jl_amd/csrc/
mlp.hip).  No counterpart exists in the reference, whose BFS is top-down
only (src/traversals/bfs.jl:108-122).

Parameter layout (1250 floats): W1[64][16], b1[64], W2[2][64], b2[2],
feature mean mu[16], feature scale sd[16].  Forward: xs = (x - mu) / sd,
h = relu(W1 xs + b1), z = W2 h + b2.  Loss: sum_i w_i CE(softmax(z_i), y_i)
/ sum_i w_i.

Parity pinning: loss_grad() is checked against central finite differences
of loss() in tests/test_policy.py, so the analytic gradient here is not
trusted on its own.
"""
from __future__ import annotations

import numpy as np

F, H = 16, 64
NW = H * F + H + 2 * H + 2 + F + F
NTRAIN = H * F + H + 2 * H + 2


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + H * F].reshape(H, F); o += H * F
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + 2 * H].reshape(2, H); o += 2 * H
    b2 = p[o:o + 2]; o += 2
    mu = p[o:o + F]; o += F
    sd = p[o:o + F]
    return W1, b1, W2, b2, mu, sd


def forward(x, p):
    W1, b1, W2, b2, mu, sd = unpack(p)
    xs = (np.asarray(x, np.float64) - mu) / sd
    a = xs @ W1.T + b1
    h = np.maximum(a, 0.0)
    z = h @ W2.T + b2
    return xs, a, h, z


def loss(x, y, w, p):
    _, _, _, z = forward(x, p)
    y = np.asarray(y)
    w = np.ones(len(y)) if w is None else np.asarray(w, np.float64)
    m = z.max(axis=1)
    lse = np.log(np.exp(z - m[:, None]).sum(axis=1)) + m
    ce = lse - z[np.arange(len(y)), y]
    return float((w * ce).sum() / w.sum())


def loss_grad(x, y, w, p):
    """Loss and gradient (NW entries; mu/sd entries are 0)."""
    W1, b1, W2, b2, mu, sd = unpack(p)
    xs, a, h, z = forward(x, p)
    y = np.asarray(y)
    w = np.ones(len(y)) if w is None else np.asarray(w, np.float64)
    W = w / w.sum()
    zm = z - z.max(axis=1, keepdims=True)
    e = np.exp(zm)
    pr = e / e.sum(axis=1, keepdims=True)
    onehot = np.zeros_like(pr)
    onehot[np.arange(len(y)), y] = 1.0
    dz = W[:, None] * (pr - onehot)
    gW2 = dz.T @ h
    gb2 = dz.sum(axis=0)
    dh = (dz @ W2) * (a > 0)
    gW1 = dh.T @ xs
    gb1 = dh.sum(axis=0)
    g = np.zeros(NW)
    g[:H * F] = gW1.ravel()
    g[H * F:H * F + H] = gb1
    g[H * F + H:H * F + 3 * H] = gW2.ravel()
    g[H * F + 3 * H:NTRAIN] = gb2
    return loss(x, y, w, p), g
