"""CPU restatement of Knn3Model (src/knn3.rs) -- test infrastructure only, the
checker of the GPU knn3 kernels; never imported by the product.

Follows src/knn3.rs:
  * compute_baseline (:123-165): mu = f32(FP64 mean of the non-test raw
    ratings); bm[m] = f32(sum (raw - mu) / (cnt + reg_m)), the differences
    in FP32 and the sums in FP64 in rating order; bu[u] likewise over
    raw - mu - bm[m]                                           (FP32 as written)
  * r_prime = raw - mu - bm[m] - bu[u] (FP32, left to right)   (:295-300)
  * co-rating counts over non-test ratings; for item m every j != m with
    count > 0 and n_j >= 1 is a candidate with
    prox = (n_mj N_USERS / (n_m n_j)) n_mj / (n_mj + shrinkage), FP32 (:195-210)
  * candidates by prox, largest first; K = clamp(#{prox >= threshold},
    k_min, k_max), at most the candidate count                 (:211-216)
  * C[a, b] = sum over users u of m of x_a(u) x_b(u) and d[a] = sum x_a(u)
    x_m(u), FP64 of FP32 products, for the residual (err) and r_prime (bias)
    signals; users that rated m only                           (:222-270)
  * M = x C_err + (1 - x) C_bias + reg I, rhs = x d_err + (1 - x) d_bias
                                                               (:272-286)
  * solve_nnls (:79-120): Cholesky solve on the active set, drop the most
    negative component (the first of equals), repeat; a failed Cholesky
    gives all zeros
  * weights w > 0 kept as FP32 in candidate order; predict sums w * err over
    the neighbours the user rated, FP32, in that order      (:379-389)
Two orders the reference leaves to its unstable sort, equal prox values
within the candidate list, are fixed here as the smaller item index first;
the GPU follows the same rule.
"""
from __future__ import annotations

import numpy as np

N_USERS = 480_189  # src/lib.rs:15

F32 = np.float32


class Knn3Config:
    """Knn3Config (src/knn3.rs:7-31): defaults of Knn3Config::default."""

    def __init__(self, threshold=0.25, k_min=10, k_max=60, shrinkage=20000.0, reg=0.01, x=0.8, bl_reg_m=25.0,
                 bl_reg_u=10.0):
        self.threshold, self.k_min, self.k_max = float(threshold), int(k_min), int(k_max)
        self.shrinkage, self.reg, self.x = float(shrinkage), float(reg), float(x)
        self.bl_reg_m, self.bl_reg_u = float(bl_reg_m), float(bl_reg_u)


def baseline(tr, reg_m: float, reg_u: float):
    """compute_baseline (src/knn3.rs:123-165) -> (mu, bm[I], bu[U]) as FP32."""
    keep = np.asarray(tr.is_test) == 0
    u = np.asarray(tr.user_idxs)[keep].astype(np.int64)
    m = np.asarray(tr.item_idxs)[keep].astype(np.int64)
    raw = np.asarray(tr.raw_ratings)[keep].astype(F32)
    s = np.cumsum(raw.astype(np.float64))  # sequential FP64 sum
    mu = F32(s[-1] / len(raw)) if len(raw) else F32(np.nan)
    bm_sum = np.zeros(tr.n_items, np.float64)
    np.add.at(bm_sum, m, (raw - mu).astype(np.float64))  # in rating order per item
    bm_cnt = np.bincount(m, minlength=tr.n_items).astype(np.float64)
    bm = (bm_sum / (bm_cnt + np.float64(F32(reg_m)))).astype(F32)
    bu_sum = np.zeros(tr.n_users, np.float64)
    np.add.at(bu_sum, u, ((raw - mu) - bm[m]).astype(np.float64))
    bu_cnt = np.bincount(u, minlength=tr.n_users).astype(np.float64)
    bu = np.zeros(tr.n_users, F32)
    ok = bu_cnt > 0
    bu[ok] = (bu_sum[ok] / (bu_cnt[ok] + np.float64(F32(reg_u)))).astype(F32)
    return mu, bm, bu


def r_prime(tr, mu, bm, bu) -> np.ndarray:
    raw = np.asarray(tr.raw_ratings).astype(F32)
    return ((raw - mu) - bm[np.asarray(tr.item_idxs)]) - bu[np.asarray(tr.user_idxs)]


def prox_row(cnt_row: np.ndarray, n: np.ndarray, m: int, shrinkage: float):
    """(candidate items, prox) of item m from its co-rating count row."""
    n_m = F32(n[m])
    j = np.nonzero((cnt_row > 0) & (n >= 1))[0]
    j = j[j != m]
    n_mj = cnt_row[j].astype(F32)
    phi = (n_mj * F32(N_USERS)) / (n_m * n[j].astype(F32))
    prox = (phi * n_mj) / (n_mj + F32(shrinkage))
    order = np.lexsort((j, -prox.astype(np.float64)))  # largest first, then smaller index
    return j[order], prox[order]


def solve_nnls(mat: np.ndarray, rhs: np.ndarray) -> np.ndarray:
    """solve_nnls (src/knn3.rs:79-120)."""
    n = len(rhs)
    active = np.ones(n, bool)
    for _ in range(n):
        idx = np.nonzero(active)[0]
        if len(idx) == 0:
            return np.zeros(n)
        a = mat[np.ix_(idx, idx)]
        try:
            L = np.linalg.cholesky(a)
        except np.linalg.LinAlgError:
            return np.zeros(n)
        y = _forward(L, rhs[idx])
        sol = _backward(L, y)
        worst, wv = 0, 0.0
        for ri in range(len(idx)):
            if sol[ri] < wv:
                worst, wv = ri, sol[ri]
        if wv >= 0.0:
            out = np.zeros(n)
            out[idx] = sol
            return out
        active[idx[worst]] = False
    return np.zeros(n)


def _forward(L, b):
    y = np.array(b, np.float64)
    for i in range(len(y)):
        y[i] = y[i] / L[i, i]
        y[i + 1:] -= y[i] * L[i + 1:, i]
    return y


def _backward(L, y):
    x = np.array(y, np.float64)
    for i in range(len(x) - 1, -1, -1):
        x[i] = (x[i] - float(np.dot(L[i + 1:, i], x[i + 1:]))) / L[i, i]
    return x


class Knn3Ref:
    def __init__(self, tr, cfg: Knn3Config | None = None):
        cfg = cfg or Knn3Config()
        self.cfg = cfg
        I, U = tr.n_items, tr.n_users
        keep = np.asarray(tr.is_test) == 0
        mu, bm, bu = baseline(tr, cfg.bl_reg_m, cfg.bl_reg_u)
        rp = r_prime(tr, mu, bm, bu)
        u = np.asarray(tr.user_idxs)[keep].astype(np.int64)
        m = np.asarray(tr.item_idxs)[keep].astype(np.int64)
        err = np.asarray(tr.residuals, F32)[keep]
        rp = rp[keep]
        self.mu, self.bm, self.bu = mu, bm, bu
        # dense per-user values (one rating per user and item)
        E = np.zeros((U, I), np.float64)
        R = np.zeros((U, I), np.float64)
        B = np.zeros((U, I), np.float64)
        E[u, m] = err
        R[u, m] = rp
        B[u, m] = 1.0
        cnt = B.T @ B
        n = np.diag(cnt).copy()
        self.neighbors: list[list[tuple[int, float]]] = []
        self.candidates: list[np.ndarray] = []
        x, reg = np.float64(F32(cfg.x)), np.float64(F32(cfg.reg))
        for mi in range(I):
            if n[mi] < 1:
                self.neighbors.append([])
                self.candidates.append(np.zeros(0, np.int64))
                continue
            j, prox = prox_row(cnt[mi], n, mi, cfg.shrinkage)
            above = int(np.sum(prox >= F32(cfg.threshold)))
            k = min(max(above, cfg.k_min), cfg.k_max, len(j))
            j = j[:k]
            self.candidates.append(j)
            if k == 0:
                self.neighbors.append([])
                continue
            users = np.nonzero(B[:, mi])[0]
            Xe, Xr = E[np.ix_(users, j)], R[np.ix_(users, j)]
            ce, cb = Xe.T @ Xe, Xr.T @ Xr
            de, db = Xe.T @ E[users, mi], Xr.T @ R[users, mi]
            sysm = x * ce + (1.0 - x) * cb
            sysm[np.diag_indices(k)] += reg
            w = solve_nnls(sysm, x * de + (1.0 - x) * db)
            self.neighbors.append([(int(j[a]), float(F32(w[a]))) for a in range(k) if w[a] > 0.0])
        # per-user lookup of the non-test residuals (UserErrors, src/knn3.rs:37-77)
        self._err = {(int(a), int(b)): F32(e) for a, b, e in zip(u, m, err)}

    def predict_one(self, u: int, i: int) -> float:
        pred = F32(0.0)
        for j, w in self.neighbors[i]:
            e = self._err.get((u, j))
            if e is not None:
                pred = F32(pred + F32(F32(w) * e))
        return float(pred)

    def predict(self, users, items) -> np.ndarray:
        return np.array([self.predict_one(int(a), int(b)) for a, b in zip(users, items)], np.float32)
