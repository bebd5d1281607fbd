"""CPU FP64 restatement of the fold-in encoder's MLP (test infrastructure only).

The encoder has no counterpart in the reference; this module is the checker
for the GPU kernels in netflix_prize_amd/csrc/enc_kernels.hip:

    Z = X W1 + b1,  H = gelu(Z),  Y = H W2 + b2
    loss = sum_{r,c} colw[c] (Y - T)^2 / M

with gelu(z) = z Phi(z) in its exact erf form.  forward/backward are written
out from the chain rule; tests/test_encoder_cpu.py pins backward() against
central finite differences of loss(), and gelu against its closed-form
values, so a dropped term or a transposed operand fails there.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf


def gelu(z):
    z = np.asarray(z, np.float64)
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    z = np.asarray(z, np.float64)
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(X, W1, b1, W2, b2):
    """Returns Z, H, Y in FP64."""
    X, W1, b1, W2, b2 = (np.asarray(a, np.float64) for a in (X, W1, b1, W2, b2))
    Z = X @ W1 + b1
    H = gelu(Z)
    Y = H @ W2 + b2
    return Z, H, Y


def loss_and_grad_y(Y, T, colw):
    Y, T, colw = (np.asarray(a, np.float64) for a in (Y, T, colw))
    d = Y - T
    M = Y.shape[0]
    return float((colw * d * d).sum() / M), 2.0 * colw * d / M


def backward(X, W2, Z, H, dY):
    """Gradients of the loss given dY = dloss/dY: dW1, db1, dW2, db2."""
    X, W2, Z, H, dY = (np.asarray(a, np.float64) for a in (X, W2, Z, H, dY))
    dW2 = H.T @ dY
    db2 = dY.sum(axis=0)
    dH = dY @ W2.T
    dZ = dH * gelu_grad(Z)
    dW1 = X.T @ dZ
    db1 = dZ.sum(axis=0)
    return dW1, db1, dW2, db2


def loss(params, X, T, colw):
    W1, b1, W2, b2 = params
    _, _, Y = forward(X, W1, b1, W2, b2)
    return loss_and_grad_y(Y, T, colw)[0]


def grads(params, X, T, colw):
    W1, b1, W2, b2 = params
    Z, H, Y = forward(X, W1, b1, W2, b2)
    val, dY = loss_and_grad_y(Y, T, colw)
    return val, backward(X, W2, Z, H, dY)
