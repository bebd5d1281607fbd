"""CPU restatement of KnnfModel (src/knnf.rs) -- test infrastructure only, the
checker of the GPU knnf kernels; never imported by the product.

Follows src/knnf.rs line by line:
  * factor_norms[i] = sqrt(row_i . row_i)                  (:64-68)
  * per-user rated lists (item, residual, day as f32)      (:70-88)
  * sim(i, j) = max(0, row_i . row_j / (n_i n_j)), 0 if a norm is 0   (:46-52)
  * predict: skip j == i; w = sim^scaling / (1 + tau |day - d_j|); skip
    w == 0; no candidates -> 0; keep the k largest w; sum(w r) / sum(w)
    (:101-141).
Arithmetic is FP64 here (the reference is FP32): predictions agree to ~1e-6
except where the k-th and (k+1)-th weights nearly tie, which the reference's
unstable selection leaves open too; `margins` reports that gap per query.
"""
from __future__ import annotations

import numpy as np


def norms(F: np.ndarray) -> np.ndarray:
    F = np.asarray(F, np.float64)
    return np.sqrt(np.einsum("ij,ij->i", F, F))


def sim_matrix(F: np.ndarray) -> np.ndarray:
    """max(0, cos) between all rows, 0 where a row norm is 0."""
    F = np.asarray(F, np.float64)
    n = norms(F)
    G = F @ F.T
    den = np.outer(n, n)
    S = np.zeros_like(G)
    ok = den != 0.0
    S[ok] = G[ok] / den[ok]
    return np.maximum(S, 0.0)


class KnnfRef:
    def __init__(self, tr, F: np.ndarray, k: int = 25, scaling: float = 1.0, tau: float = 0.025):
        self.S = sim_matrix(F)
        self.k, self.scaling, self.tau = int(k), float(scaling), float(tau)
        u = np.asarray(tr.user_idxs)
        self.off = np.zeros(tr.n_users + 1, np.int64)
        np.add.at(self.off, u + 1, 1)
        self.off = np.cumsum(self.off)
        self.items = np.asarray(tr.item_idxs, np.int64)
        self.res = np.asarray(tr.residuals, np.float64)
        self.days = np.asarray(tr.dates, np.float64)

    def _weights(self, u: int, i: int, day: float):
        a, b = self.off[u], self.off[u + 1]
        j, r, d = self.items[a:b], self.res[a:b], self.days[a:b]
        s = self.S[i, j]
        if self.scaling != 1.0:
            s = s ** self.scaling
        w = s / (1.0 + self.tau * np.abs(day - d))
        keep = (j != i) & (w != 0.0)
        return w[keep], r[keep]

    def predict_one(self, u: int, i: int, day: float) -> float:
        w, r = self._weights(u, i, day)
        if len(w) == 0:
            return 0.0
        if len(w) > self.k:
            top = np.argsort(-w, kind="stable")[: self.k]
            w, r = w[top], r[top]
        return float(np.sum(w * r) / np.sum(w))

    def predict(self, users, items, days) -> np.ndarray:
        return np.array([self.predict_one(int(u), int(i), float(d)) for u, i, d in zip(users, items, days)])

    def margins(self, users, items, days) -> np.ndarray:
        """Relative gap between the k-th and (k+1)-th weights (inf when at most k)."""
        out = np.full(len(users), np.inf)
        for q, (u, i, d) in enumerate(zip(users, items, days)):
            w, _ = self._weights(int(u), int(i), float(d))
            if len(w) > self.k:
                s = np.sort(w)[::-1]
                out[q] = (s[self.k - 1] - s[self.k]) / s[self.k - 1]
        return out
