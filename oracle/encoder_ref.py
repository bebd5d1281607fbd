"""CPU FP64 restatement of the fold-in encoder's MLP (test infrastructure only).

The encoder has no counterpart in the reference; this module is the checker
for the GPU kernels in netflix_prize_amd/csrc/enc_kernels.hip:

    Z = X W1 + b1,  H = gelu(Z),  Y = H W2 + b2
    loss = sum_{r,c} colw[c] (Y - T)^2 / M

with gelu(z) = z Phi(z) in its exact erf form.  forward/backward are written
out from the chain rule; tests/test_encoder_cpu.py pins backward() against
central finite differences of loss(), and gelu against its closed-form
values, so a dropped term or a transposed operand fails there.

item_day_stats and user_features restate the encoder's input: per-rating
tokens (item factor row, rating one-hot, per-item normalised day, residual
against the biases) mean-pooled per user, plus per-user scalars, in FP64
with plain per-user loops.  tests/test_encoder_cpu.py pins them on hand-made
users whose pooled blocks and scalars are written out by hand (a shifted
one-hot class, a wrong sample position, a day normalised by the user's
statistics instead of the item's, or a residual without the item bias fails
there).
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


# --------------------------------------------------------------------------
# features
# --------------------------------------------------------------------------

DAYS = 2243.0
SCALAR_COLS = 64


def item_day_stats(items, days, n_items):
    """Per-item mean day and 1 / max(std, 1) (population std); 0, 0 for an
    item without ratings."""
    items = np.asarray(items, np.int64)
    days = np.asarray(days, np.float64)
    cnt = np.bincount(items, minlength=n_items).astype(np.float64)
    s1 = np.bincount(items, weights=days, minlength=n_items)
    s2 = np.bincount(items, weights=days * days, minlength=n_items)
    mean = np.zeros(n_items)
    inv = np.zeros(n_items)
    has = cnt > 0
    mean[has] = s1[has] / cnt[has]
    var = np.zeros(n_items)
    var[has] = np.maximum(s2[has] / cnt[has] - mean[has] ** 2, 0.0)
    inv[has] = 1.0 / np.maximum(np.sqrt(var[has]), 1.0)
    return mean, inv


def user_features(q, ibias, mu, day_mean, day_inv, off, items, ratings, targets, days, max_items):
    """Row u: mean over the sampled ratings of one-hot(class) x q_j (5 blocks),
    d q_j and w q_j, then 64 scalars over all the user's ratings (the column
    contract of enc_user_features in include/tsvd_enc.h)."""
    q = np.asarray(q, np.float64)
    f = q.shape[1]
    U = len(off) - 1
    out = np.zeros((U, 7 * f + SCALAR_COLS))
    for u in range(U):
        s0, cnt = int(off[u]), int(off[u + 1] - off[u])
        j = np.asarray(items[s0:s0 + cnt], np.int64)
        r = np.asarray(ratings[s0:s0 + cnt], np.int64)
        w = np.asarray(targets[s0:s0 + cnt], np.float64) - mu - np.asarray(ibias, np.float64)[j]
        day = np.asarray(days[s0:s0 + cnt], np.float64)
        d = (day - np.asarray(day_mean, np.float64)[j]) * np.asarray(day_inv, np.float64)[j]
        ns = min(cnt, max_items)
        for t in range(ns):
            k = (t * cnt) // ns          # position within the user's ratings
            row = q[j[k]]
            if 1 <= r[k] <= 5:
                out[u, (r[k] - 1) * f:r[k] * f] += row
            out[u, 5 * f:6 * f] += d[k] * row
            out[u, 6 * f:7 * f] += w[k] * row
        if ns:
            out[u, :7 * f] /= ns
        sc = out[u, 7 * f:]
        sc[0] = np.log1p(cnt) / 8.0
        sc[14] = 1.0
        if cnt:
            for c in range(1, 6):
                sc[c] = np.mean(r == c)
            sc[6], sc[7] = w.mean(), w.std()
            sc[8], sc[9] = d.mean(), d.std()
            sc[10] = day.mean() / DAYS
            sc[11] = (day.max() - day.min()) / DAYS
            sc[12] = day.min() / DAYS
            sc[13] = day.max() / DAYS
    return out
