"""CPU oracle of the warm-start MLP -- TEST INFRASTRUCTURE ONLY.

Forward and backward of x -> GELU(x W1 + b1) -> GELU(. W2 + b2) -> . W3 + b3
with an MSE loss, in numpy.  With ``bf16=True`` every GEMM operand is rounded
to bfloat16 (round to nearest even) exactly where the GPU kernels round them,
and products accumulate in float64; with ``bf16=False`` everything is float64
(used for the finite-difference pins of the backward formulas).  GELU is the
tanh form 0.5 x (1 + tanh(0.7978845608 (x + 0.044715 x^3))).
The warm start has no counterpart in the reference.
"""
from __future__ import annotations

import numpy as np

C_GELU = 0.7978845608
A_GELU = 0.044715


def bf16_round(x: np.ndarray) -> np.ndarray:
    """float32 -> nearest-even bfloat16, returned as float64 values."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(C_GELU * (z + A_GELU * z ** 3)))


def gelu_grad(z):
    t = np.tanh(C_GELU * (z + A_GELU * z ** 3))
    dt = (1.0 - t * t) * C_GELU * (1.0 + 3.0 * A_GELU * z * z)
    return 0.5 * (1.0 + t) + 0.5 * z * dt


def forward(params, X, bf16=True):
    W1, b1, W2, b2, W3, b3 = params
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    f = (lambda a: np.asarray(a, np.float32).astype(np.float64)) if bf16 else (lambda a: a)
    Xb = r(X)
    Z1 = f(Xb @ r(W1) + b1)
    G1 = f(gelu(Z1))
    Z2 = f(r(G1) @ r(W2) + b2)
    G2 = f(gelu(Z2))
    Y = f(r(G2) @ r(W3) + b3)
    return Y, (Xb, Z1, G1, Z2, G2)


def loss_and_grads(params, X, T, bf16=True):
    W1, b1, W2, b2, W3, b3 = params
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    f = (lambda a: np.asarray(a, np.float32).astype(np.float64)) if bf16 else (lambda a: a)
    Y, (Xb, Z1, G1, Z2, G2) = forward(params, X, bf16)
    M, N = Y.shape
    E = Y - T
    loss = float(np.sum(E * E) / (M * N))
    dY = f(2.0 * E / (M * N))
    dW3 = r(G2).T @ r(dY)
    db3 = dY.sum(0)
    dG2 = f(r(dY) @ r(W3).T)
    dZ2 = f(dG2 * gelu_grad(Z2))
    dW2 = r(G1).T @ r(dZ2)
    db2 = dZ2.sum(0)
    dG1 = f(r(dZ2) @ r(W2).T)
    dZ1 = f(dG1 * gelu_grad(Z1))
    dW1 = Xb.T @ r(dZ1)
    db1 = dZ1.sum(0)
    return loss, [dW1, db1, dW2, db2, dW3, db3]


def adam_step(w, g, m1, m2, step, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    """One Adam step (bias-corrected moments, L2 term wd * w added to g)."""
    g = g + wd * w
    m1 = b1 * m1 + (1 - b1) * g
    m2 = b2 * m2 + (1 - b2) * g * g
    w = w - lr * (m1 / (1 - b1 ** step)) / (np.sqrt(m2 / (1 - b2 ** step)) + eps)
    return w, m1, m2


def init_params(sizes, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for k_in, k_out in zip(sizes[:-1], sizes[1:]):
        out.append((rng.standard_normal((k_in, k_out)) * np.sqrt(2.0 / k_in)).astype(np.float32))
        out.append(np.zeros(k_out, np.float32))
    return out
