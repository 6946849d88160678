"""CPU restatement of the schedule cost model's forward and backward passes.

TEST INFRASTRUCTURE ONLY: loaded by tests/ as the checker of the GPU MLP
kernels (bitstreamzoo_amd/csrc/mlp_kernels.hip, mlp_host.cpp), never by the
shipped library.  The cost model has no counterpart in the reference; this
file restates the textbook two-hidden-layer ReLU network with mean squared
error, in numpy:

    H1 = relu(X W1 + b1);  H2 = relu(H1 W2 + b2);  out = H2 w3 + b3
    L  = mean((out - y)^2)

`bf16=True` rounds every matrix-product operand to bfloat16 (round to
nearest even) exactly where the GPU kernels do (inputs, weights, activations
and back-propagated gradients feeding an MFMA product) and accumulates in
float32, so the GPU result should match it to float32 summation-order
differences; `bf16=False` in float64 is the plain reference that the
finite-difference test pins.
"""

from __future__ import annotations

import numpy as np


def bf16_round(a: np.ndarray) -> np.ndarray:
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unpack(p: np.ndarray, n_in: int, h1: int, h2: int):
    o = 0
    W1 = p[o:o + n_in * h1].reshape(n_in, h1); o += n_in * h1
    b1 = p[o:o + h1]; o += h1
    W2 = p[o:o + h1 * h2].reshape(h1, h2); o += h1 * h2
    b2 = p[o:o + h2]; o += h2
    w3 = p[o:o + h2]; o += h2
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def n_params(n_in: int, h1: int, h2: int) -> int:
    return n_in * h1 + h1 + h1 * h2 + h2 + h2 + 1


def forward(p, X, n_in, h1, h2, bf16=True):
    dt = np.float32 if bf16 else np.float64
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, w3, b3 = unpack(np.asarray(p, dtype=dt), n_in, h1, h2)
    Xb = r(X)
    H1 = np.maximum(r(Xb) @ r(W1) + b1, 0).astype(dt)
    H2 = np.maximum(r(H1) @ r(W2) + b2, 0).astype(dt)
    out = (r(H2) @ r(w3) + b3).astype(dt)
    return Xb, H1, H2, out


def loss_and_grad(p, X, y, n_in, h1, h2, bf16=True):
    dt = np.float32 if bf16 else np.float64
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, w3, b3 = unpack(np.asarray(p, dtype=dt), n_in, h1, h2)
    Xb, H1, H2, out = forward(p, X, n_in, h1, h2, bf16)
    n = len(y)
    e = out - np.asarray(y, dtype=dt)
    loss = float(np.mean(e.astype(np.float64) ** 2))
    d = (2.0 * e / n).astype(dt)
    dw3 = r(H2).T @ r(d)
    db3 = d.sum()
    dH2 = (r(d)[:, None] @ r(w3)[None, :]) * (H2 > 0)
    dW2 = r(H1).T @ r(dH2)
    db2 = dH2.sum(axis=0)
    dH1 = (r(dH2) @ r(W2).T) * (H1 > 0)
    dW1 = r(Xb).T @ r(dH1)
    db1 = dH1.sum(axis=0)
    g = np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dw3, [db3]]).astype(dt)
    return loss, g
