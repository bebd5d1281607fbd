"""The knnf oracle (oracle/knnf_ref.py) pinned on the CPU: the reference's KAT
and closed-form cases that each fail for a plausible mistake (decay sign or
form, the j == i skip, the cosine clamp, top-k direction, the weighted mean)."""
import numpy as np
import pytest

from oracle.knnf_ref import KnnfRef, sim_matrix
from tests.fixtures import dataset_from_rows, tiny

# tests/regressors_test.rs:403-422 (knnf_probe_rmse_regression)
KNNF_KAT_RMSE = 1.9509877627957246


def kat_factors():
    i, k = np.meshgrid(np.arange(10), np.arange(4), indexing="ij")
    return (((i * 13 + k * 7 + 3) % 17).astype(np.float32) / np.float32(17.0) - np.float32(0.5)).astype(np.float32)


def rmse(pred, ds):
    e = np.asarray(pred, np.float64) - np.asarray(ds.residuals, np.float64)
    return float(np.sqrt(np.mean(e[np.asarray(ds.is_test) == 0] ** 2)))


def test_oracle_reproduces_reference_kat():
    _, tr, pr = tiny()
    ref = KnnfRef(tr, kat_factors(), k=3, scaling=1.0, tau=0.025)
    r = rmse(ref.predict(pr.user_idxs, pr.item_idxs, pr.dates), pr)
    assert r == pytest.approx(KNNF_KAT_RMSE, abs=1e-6), r


def _one_user(rows, n_items):
    return dataset_from_rows([(0, i, r, d) for i, r, d in rows], 1, n_items)


def test_weighted_mean_with_temporal_decay():
    # item 0 at 60 degrees from item 1 (cos 0.5) and parallel to item 2 (cos 1)
    F = np.array([[1.0, 0.0], [0.5, np.sqrt(3) / 2], [2.0, 0.0]], np.float32)
    tr = _one_user([(1, 4, 100), (2, 1, 140)], 3)
    ref = KnnfRef(tr, F, k=25, scaling=1.0, tau=0.025)
    # day 100: w1 = 0.5 / 1, w2 = 1 / (1 + 0.025 * 40) = 0.5
    assert ref.predict_one(0, 0, 100.0) == pytest.approx((0.5 * 4 + 0.5 * 1) / 1.0, abs=1e-6)
    # day 140: w1 = 0.5 / 2, w2 = 1
    assert ref.predict_one(0, 0, 140.0) == pytest.approx((0.25 * 4 + 1.0 * 1) / 1.25, abs=1e-6)
    # k = 1 keeps the larger weight only
    assert KnnfRef(tr, F, k=1).predict_one(0, 0, 140.0) == pytest.approx(1.0)
    assert KnnfRef(tr, F, k=1).predict_one(0, 0, 100.0 - 1000.0) == pytest.approx(1.0)
    # scaling 2 squares the cosine: w1 = 0.25, w2 = 0.5 at day 100
    assert KnnfRef(tr, F, k=25, scaling=2.0).predict_one(0, 0, 100.0) == pytest.approx((0.25 * 4 + 0.5) / 0.75)


def test_self_item_skipped_and_negative_cosine_clamped():
    F = np.array([[1.0, 0.0], [-1.0, 0.1], [0.0, 1.0]], np.float32)
    tr = _one_user([(0, 5, 10), (1, 3, 10), (2, 2, 10)], 3)
    ref = KnnfRef(tr, F)
    # querying item 0: item 0 itself is skipped, item 1 has cos < 0, item 2 cos 0 -> no candidate
    assert ref.predict_one(0, 0, 10.0) == 0.0
    S = sim_matrix(F)
    assert S[0, 1] == 0.0 and S[0, 0] == pytest.approx(1.0)


def test_constant_residuals_give_the_constant():
    rng = np.random.default_rng(0)
    F = np.abs(rng.normal(size=(30, 8))).astype(np.float32)
    tr = _one_user([(j, 3, int(d)) for j, d in zip(range(1, 30), rng.integers(0, 2000, 29))], 30)
    ref = KnnfRef(tr, F, k=7)
    assert ref.predict_one(0, 0, 1000.0) == pytest.approx(3.0, abs=1e-12)


# tests/regressors_test.rs:218-233 (knns_probe_rmse_regression)
KNNS_KAT_RMSE = 1.2349219913552398


def test_knns_oracle_reproduces_reference_kat():
    from oracle.knnf_ref import KnnsRef
    fx, tr, pr = tiny()
    ref = KnnsRef(tr, pr, fx["n_items"], k=3, shrinkage=5.0, scaling=1.0, tau=0.025)
    r = rmse(ref.predict(pr.user_idxs, pr.item_idxs, pr.dates), pr)
    assert r == pytest.approx(KNNS_KAT_RMSE, abs=1e-6), r


def test_knns_support_counts_training_and_heldout_items():
    from oracle.knnf_ref import N_USERS, knns_sim_matrix
    # user 0 rates items 0, 1 (train) and 2 (held out); user 1 rates 0 and 2
    tr = dataset_from_rows([(0, 0, 4, 1), (0, 1, 3, 2), (1, 0, 5, 3)], 2, 3)
    pr = dataset_from_rows([(0, 2, 1, 4), (1, 2, 2, 5)], 2, 3)
    S = knns_sim_matrix(tr, pr, 3, shrinkage=1.0)
    # supp: n(0)=2, n(1)=1, n(2)=2, n(0,2)=2, n(0,1)=1
    assert S[0, 2] == pytest.approx(2 * N_USERS / (2 * 2) * 2 / (2 + 1.0))
    assert S[0, 1] == pytest.approx(1 * N_USERS / (2 * 1) * 1 / (1 + 1.0))
    assert S[1, 2] == pytest.approx(1 * N_USERS / (1 * 2) * 1 / (1 + 1.0))


def test_knn_stat_oracle_closed_form():
    """KnnModel's `*__knn-d` form (src/knn.rs:105-113, 263-336): one user
    with items 1 and 2 rated; the prediction for item 0 is the weighted mean
    of stat1 / (stat0 + shrinkage) with weights (support lift)^scaling."""
    from oracle.knnf_ref import KnnStatRef
    s0 = np.array([[3, 2, 1], [2, 4, 1], [1, 1, 5]], np.float32)
    s1 = np.array([[0, 6, -4], [-6, 0, 2], [4, -2, 0]], np.float32)
    tr = _one_user([(1, 4, 100), (2, 1, 140)], 3)
    shr, sc = 10.0, 2.0
    ref = KnnStatRef(tr, s0, s1, k=15, shrinkage=shr, scaling=sc, tau=0.0)
    sim01 = (2 * 480189 / (3 * 4)) * 2 / (2 + shr)
    sim02 = (1 * 480189 / (3 * 5)) * 1 / (1 + shr)
    w1, w2 = sim01 ** sc, sim02 ** sc
    want = (w1 * 6 / (2 + shr) + w2 * (-4) / (1 + shr)) / (w1 + w2)
    assert ref.predict_one(0, 0, 120.0) == pytest.approx(want, rel=1e-6)
    # k = 1 keeps the larger weight (item 1: sim01 > sim02)
    assert KnnStatRef(tr, s0, s1, k=1, shrinkage=shr, scaling=sc).predict_one(0, 0, 0.0) == pytest.approx(
        6 / (2 + shr), rel=1e-6)
