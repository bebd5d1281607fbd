"""The knn3 oracle (oracle/knn3_ref.py) pinned on the CPU: the reference's KAT
and cases whose answer the mathematics fixes (the regularised means, the
proximity formula and its clamp, NNLS on systems with closed-form solutions,
a planted regression the weights must recover, the prediction sum)."""
import numpy as np
import pytest

from oracle.knn3_ref import Knn3Config, Knn3Ref, baseline, prox_row, solve_nnls
from tests.fixtures import dataset_from_rows, tiny

# tests/regressors_test.rs:177-196 (knn3_probe_rmse_regression)
KNN3_KAT_CFG = dict(threshold=0.25, k_min=2, k_max=5, shrinkage=5.0, reg=0.01, x=0.8, bl_reg_m=2.0, bl_reg_u=1.0)
KNN3_KAT_RMSE = 3.0427174790608045


def rmse(pred, ds):
    e = np.asarray(pred, np.float64) - np.asarray(ds.residuals, np.float64)
    return float(np.sqrt(np.mean(e[np.asarray(ds.is_test) == 0] ** 2)))


def test_oracle_reproduces_reference_kat():
    _, tr, pr = tiny()
    ref = Knn3Ref(tr, Knn3Config(**KNN3_KAT_CFG))
    r = rmse(ref.predict(pr.user_idxs, pr.item_idxs), pr)
    assert r == pytest.approx(KNN3_KAT_RMSE, abs=1e-6), r


def test_baseline_is_the_regularised_mean_without_test_rows():
    rows = [(0, 0, 5, 1), (0, 1, 3, 2), (1, 0, 4, 3), (1, 2, 1, 4), (2, 1, 2, 5), (2, 2, 5, 6)]
    tr = dataset_from_rows(rows, 3, 3)
    tr.is_test = np.array([0, 0, 0, 0, 0, 1], np.int8)  # (2, 2, 5) must not count
    mu, bm, bu = baseline(tr, 2.0, 1.0)
    assert mu == np.float32(15 / 5)
    # bm[m] = sum(r - mu) / (n_m + 2): item 0 -> (2 + 1) / 4, item 1 -> (0 - 1) / 4, item 2 -> (-2) / 3
    np.testing.assert_allclose(bm, [3 / 4, -1 / 4, -2 / 3], rtol=1e-6)
    # bu[u] = sum(r - mu - bm) / (n_u + 1)
    want_bu = [((5 - 3 - 0.75) + (3 - 3 + 0.25)) / 3, ((4 - 3 - 0.75) + (1 - 3 + 2 / 3)) / 3, (2 - 3 + 0.25) / 2]
    np.testing.assert_allclose(bu, want_bu, rtol=1e-6)


def test_prox_formula_order_and_clamp():
    # item 0 co-rated with item 1 by 2 users, with item 2 by 3 users; n = (3, 4, 6, 0)
    cnt = np.array([3, 2, 3, 0], np.float64)
    n = np.array([3, 4, 6, 0], np.float64)
    j, p = prox_row(cnt, n, 0, 5.0)
    p1 = 2 * 480189 / (3 * 4) * 2 / (2 + 5)
    p2 = 3 * 480189 / (3 * 6) * 3 / (3 + 5)
    assert list(j) == [2, 1]  # p2 > p1
    np.testing.assert_allclose(p, [p2, p1], rtol=1e-6)
    # equal prox: the smaller item first
    j, _ = prox_row(np.array([5, 2, 2], np.float64), np.array([5, 4, 4], np.float64), 0, 5.0)
    assert list(j) == [1, 2]


def test_nnls_closed_forms():
    # diagonal: each component independently max(0, b / d)
    d = np.array([2.0, 4.0, 1.0, 5.0])
    b = np.array([1.0, -2.0, 3.0, -0.5])
    np.testing.assert_allclose(solve_nnls(np.diag(d), b), np.maximum(b / d, 0.0), rtol=1e-15)
    # positive unconstrained solution: the plain solve
    rng = np.random.default_rng(1)
    A = rng.normal(size=(8, 5))
    M = A.T @ A + np.eye(5)
    w = np.abs(rng.normal(size=5)) + 0.1
    np.testing.assert_allclose(solve_nnls(M, M @ w), w, rtol=1e-10)
    # 2 x 2 with a negative component: it is dropped and the other solved alone
    M = np.array([[2.0, 1.0], [1.0, 2.0]])
    b = np.array([3.0, -3.0])  # unconstrained (3, -3)
    np.testing.assert_allclose(solve_nnls(M, b), [1.5, 0.0], rtol=1e-15)
    # not positive definite: all zeros
    assert np.all(solve_nnls(np.array([[1.0, 2.0], [2.0, 1.0]]), np.array([1.0, 1.0])) == 0.0)


def _planted(n_users=400, seed=3):
    """Every user rates items 0..5; the residual of item 0 is 0.5 e1 + 0.3 e2
    + 0.2 e4 of the user's residuals on items 1, 2 and 4."""
    rng = np.random.default_rng(seed)
    E = rng.normal(size=(n_users, 6)).astype(np.float32)
    E[:, 0] = (0.5 * E[:, 1].astype(np.float64) + 0.3 * E[:, 2] + 0.2 * E[:, 4]).astype(np.float32)
    rows = [(u, m, 3, 100) for u in range(n_users) for m in range(6)]
    tr = dataset_from_rows(rows, n_users, 6)
    tr.residuals = E.reshape(-1).copy()
    return tr, E


def test_planted_weights_are_recovered():
    tr, E = _planted()
    cfg = Knn3Config(threshold=0.25, k_min=10, k_max=60, shrinkage=20000.0, reg=1e-6, x=1.0)
    ref = Knn3Ref(tr, cfg)
    got = dict(ref.neighbors[0])
    assert set(ref.candidates[0]) == {1, 2, 3, 4, 5}  # every other item, k_min above the count
    for j, w in ((1, 0.5), (2, 0.3), (4, 0.2)):
        assert got[j] == pytest.approx(w, abs=1e-4)
    assert got.get(3, 0.0) < 1e-4 and got.get(5, 0.0) < 1e-4
    # prediction: the weighted sum of the user's residuals on the neighbours
    u = 7
    assert ref.predict_one(u, 0) == pytest.approx(float(E[u, 0]), abs=1e-4)


def test_prediction_skips_unrated_neighbours():
    tr, E = _planted(50)
    ref = Knn3Ref(tr, Knn3Config(reg=1e-6, x=1.0))
    ref._err.pop((4, 2))  # user 4 without item 2
    want = sum(np.float32(w) * E[4, j] for j, w in ref.neighbors[0] if j != 2)
    assert ref.predict_one(4, 0) == pytest.approx(float(want), abs=1e-6)
    assert Knn3Ref(tr, Knn3Config(k_min=0, threshold=1e30)).predict_one(0, 0) == 0.0  # no neighbours
