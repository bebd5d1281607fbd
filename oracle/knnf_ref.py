"""CPU restatement of KnnfModel (src/knnf.rs), KnnsModel (src/knns.rs) and the
KnnModel form of the `*__knn-d` job (src/knn.rs, Support + use_stat1) --
test infrastructure only, the checker of the GPU kNN kernels; never imported
by the product.

Follows src/knnf.rs line by line:
  * factor_norms[i] = sqrt(row_i . row_i)                  (:64-68)
  * per-user rated lists (item, residual, day as f32)      (:70-88)
  * sim(i, j) = max(0, row_i . row_j / (n_i n_j)), 0 if a norm is 0   (:46-52)
  * predict: skip j == i; w = sim^scaling / (1 + tau |day - d_j|); skip
    w == 0; no candidates -> 0; keep the k largest w; sum(w r) / sum(w)
    (:101-141).  Equal weights at the k-th place as the reference's
    select_nth_unstable_by (core::slice::select): with k + 1 candidates the
    first smallest is dropped (index == len - 1, max_index keeps the first
    of equals), otherwise equal weights are kept in rating order (its
    insertion sort up to 16 elements; beyond that its quickselect
    partitioning decides, which this restatement does not follow).
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
        if len(w) == self.k + 1:
            keep = np.ones(len(w), bool)
            keep[np.argmin(w)] = False  # the first smallest
            w, r = w[keep], r[keep]
        elif len(w) > self.k:
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


N_USERS = 480_189  # src/lib.rs:15


def knns_sim_matrix(tr, pr, n_items: int, shrinkage: float) -> np.ndarray:
    """KnnsModel::sim (src/knns.rs:64-70) over build_bin_nbstats(tr, pr,
    "supp") (src/nbstats.rs:73-117): supp[i, j] = number of users with both i
    and j among their training and held-out items (diagonal: users of i)."""
    n_users = int(max(tr.n_users, pr.n_users))
    has_train = np.bincount(np.asarray(tr.user_idxs), minlength=n_users) > 0  # others skipped (:79-80)
    pu = np.asarray(pr.user_idxs)
    users = np.concatenate([np.asarray(tr.user_idxs), pu[has_train[pu]]]).astype(np.int64)
    items = np.concatenate([np.asarray(tr.item_idxs), np.asarray(pr.item_idxs)[has_train[pu]]]).astype(np.int64)
    B = np.zeros((n_users, n_items), np.float64)
    np.add.at(B, (users, items), 1.0)
    supp = B.T @ B
    d = np.diag(supp).copy()
    den = np.outer(d, d)
    phi = np.zeros_like(supp)
    ok = den > 0
    phi[ok] = supp[ok] * N_USERS / den[ok]
    return phi * supp / (supp + shrinkage)


class KnnsRef(KnnfRef):
    def __init__(self, tr, pr, n_items: int, k: int = 25, shrinkage: float = 100.0, scaling: float = 1.0,
                 tau: float = 0.025):
        super().__init__(tr, np.ones((n_items, 1), np.float32), k=k, scaling=scaling, tau=tau)
        self.S = knns_sim_matrix(tr, pr, n_items, shrinkage)


def stat_sim_matrix(stat0: np.ndarray, shrinkage: float) -> np.ndarray:
    """KnnModel::support_sim (src/knn.rs:105-113) over a loaded support stat:
    phi = n * N_USERS / (s_ii s_jj) (0 if that product is 0), phi n / (n + shrinkage)."""
    S0 = np.asarray(stat0, np.float64)
    d = np.diag(S0).copy()
    den = np.outer(d, d)
    phi = np.zeros_like(S0)
    ok = den > 0
    phi[ok] = S0[ok] * N_USERS / den[ok]
    return phi * S0 / (S0 + shrinkage)


class KnnStatRef(KnnfRef):
    """KnnModel (src/knn.rs) with SimType::Support and use_stat1, the form of the
    `*__knn-d` job: weights as KnnfModel's over stat_sim_matrix, and the
    averaged value stat1[i, j] / (stat0[i, j] + shrinkage) in place of the
    residual (:321-334)."""

    def __init__(self, tr, stat0, stat1, k: int = 15, shrinkage: float = 100.0, scaling: float = 3.5,
                 tau: float = 0.0):
        n = np.asarray(stat0).shape[0]
        super().__init__(tr, np.ones((n, 1), np.float32), k=k, scaling=scaling, tau=tau)
        self.S = stat_sim_matrix(stat0, shrinkage)
        self.Dm = (np.asarray(stat1, np.float32) / (np.asarray(stat0, np.float32) + np.float32(shrinkage))).astype(
            np.float64)

    def _weights(self, u: int, i: int, day: float):
        a, b = self.off[u], self.off[u + 1]
        j, d = self.items[a:b], self.days[a:b]
        s = self.S[i, j]
        if self.scaling != 1.0:
            s = s ** self.scaling
        w = s / (1.0 + self.tau * np.abs(day - d))
        keep = (j != i) & (w != 0.0)
        return w[keep], self.Dm[i, j][keep]
