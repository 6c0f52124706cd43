"""FP64 numpy restatement of the learned PPE warm start (TEST INFRASTRUCTURE ONLY).

Same network as fluidized_bed_master_amd/csrc/learned.hip: edge MLP
6 -> 64 -> 64 (ReLU), sum aggregation scaled by 1/64, node MLP 64 -> 64
(ReLU) -> 1.  With bf16=True the forward rounds to BF16 at the points where
the GPU feeds MFMA or reads BF16 weights (edge inputs, W1, h1, W2, W3, W4);
the backward is the exact FP64 gradient of that function with the roundings
treated as identity.  With bf16=False everything is FP64 (used for the
finite-difference pin of the backward).
"""
from __future__ import annotations

import numpy as np
from scipy.spatial import cKDTree

H, EIN = 64, 6
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * EIN
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H * H
OFF_W4 = OFF_B3 + H
OFF_B4 = OFF_W4 + H
NPARAM = OFF_B4 + 1


def bf16(x):
    """Round to bfloat16 (round to nearest even) and back to float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r.astype(np.uint32)).view(np.float32).astype(np.float64)


def unpack(w):
    w = np.asarray(w, np.float64)
    return (w[OFF_W1:OFF_B1].reshape(H, EIN), w[OFF_B1:OFF_W2], w[OFF_W2:OFF_B2].reshape(H, H), w[OFF_B2:OFF_W3],
            w[OFF_W3:OFF_B3].reshape(H, H), w[OFF_B3:OFF_W4], w[OFF_W4:OFF_B4], w[OFF_B4])


def edges(X, h, kappa):
    """All (i, j) with |x_i - x_j| + 1e-20 < kappa h_i (self included)."""
    tree = cKDTree(X)
    src, dst = [], []
    for i, cand in enumerate(tree.query_ball_point(X, kappa * h.max() * 1.0001)):
        c = np.asarray(cand)
        d = X[i] - X[c]
        r = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]) + 1e-20
        c = c[r < kappa * h[i]]
        src.append(np.full(len(c), i))
        dst.append(c)
    return np.concatenate(src), np.concatenate(dst)


def edge_inputs(X, h, nf, src, dst):
    d = (X[src] - X[dst]) / h[src][:, None]
    r = np.sqrt((d * d).sum(1))
    return np.column_stack([d, r, nf[dst, 0], nf[dst, 1]]).astype(np.float32).astype(np.float64)


def forward(w, X, h, nf, rows, src, dst, bf=True):
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(w)
    R = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    e = R(edge_inputs(X, h, nf, src, dst))
    z1 = e @ R(W1).T + b1
    h1 = R(np.maximum(z1, 0.0))
    z2 = h1 @ R(W2).T + b2
    h2 = np.maximum(z2, 0.0)
    n = len(X)
    M = np.zeros((n, H))
    np.add.at(M, src, h2)
    M /= 64.0
    z3 = M @ R(W3).T + b3
    g = np.maximum(z3, 0.0)
    y = g @ R(W4) + b4
    cache = dict(e=e, z1=z1, h1=h1, z2=z2, M=M, z3=z3, g=g, R=R)
    return y, cache


def loss_grad(w, X, h, nf, rows, src, dst, t, bf=True):
    """Mean-square loss of y against t over rows and its gradient."""
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(w)
    y, c = forward(w, X, h, nf, rows, src, dst, bf)
    R = c["R"]
    m = rows.astype(bool)
    nr = m.sum()
    err = np.where(m, y - t, 0.0)
    loss = float((err * err).sum() / nr)
    dy = 2 * err / nr
    g = np.zeros(NPARAM)
    g[OFF_B4] = dy.sum()
    g[OFF_W4:OFF_B4] = dy @ c["g"]
    dz3 = (dy[:, None] * R(W4)[None, :]) * (c["z3"] > 0)
    g[OFF_W3:OFF_B3] = (dz3.T @ c["M"]).ravel()
    g[OFF_B3:OFF_W4] = dz3.sum(0)
    dM = dz3 @ R(W3)
    dh2 = dM[src] / 64.0
    dz2 = dh2 * (c["z2"] > 0)
    g[OFF_W2:OFF_B2] = (dz2.T @ c["h1"]).ravel()
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    dh1 = dz2 @ R(W2)
    dz1 = dh1 * (c["z1"] > 0)
    g[OFF_W1:OFF_B1] = (dz1.T @ c["e"]).ravel()
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    return loss, g, y
