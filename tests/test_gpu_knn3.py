"""Knn3Model on the GPU (knn3_create / knn3_predict) against the oracle's
restatement of src/knn3.rs (needs an MI355X).

The baseline is computed in the reference's order on the host and compared
bit for bit; candidate lists are integer work and compared exactly; the
normal equations are exact FP64 sums in the reference's order, so the
weights differ from the oracle's only by the Cholesky's FP64 rounding
(LAPACK's blocked order there): within 1e-5, relative."""
import numpy as np
import pytest

from netflix_prize_amd import tx
from netflix_prize_amd.knn import Knn3Config, Knn3Model
from oracle.knn3_ref import Knn3Config as RefConfig, Knn3Ref
from tests.fixtures import dataset_from_rows, synthetic, tiny
from tests.test_knn3_cpu import KNN3_KAT_CFG, KNN3_KAT_RMSE, rmse

pytestmark = pytest.mark.gpu


def _compare(tr, cfg: dict, pred_atol=1e-5):
    ref = Knn3Ref(tr, RefConfig(**cfg))
    with Knn3Model(tr, Knn3Config(**cfg)) as m:
        mu, bm, bu = m.baseline()
        assert mu == ref.mu
        np.testing.assert_array_equal(bm, ref.bm)
        np.testing.assert_array_equal(bu, ref.bu)
        cand = m.candidates()
        for i in range(tr.n_items):
            np.testing.assert_array_equal(cand[i], ref.candidates[i], err_msg=f"candidates of item {i}")
        nb = m.neighbors()
        for i in range(tr.n_items):
            a, b = nb[i], ref.neighbors[i]
            assert [j for j, _ in a] == [j for j, _ in b], f"neighbours of item {i}"
            np.testing.assert_allclose([w for _, w in a], [w for _, w in b], rtol=1e-5, atol=1e-7,
                                       err_msg=f"weights of item {i}")
        rng = np.random.default_rng(0)
        q = rng.integers(0, len(tr.user_idxs), 2000)
        u = np.concatenate([tr.user_idxs[q], rng.integers(0, tr.n_users, 500)]).astype(np.int32)
        i = np.concatenate([rng.integers(0, tr.n_items, 2000), rng.integers(0, tr.n_items, 500)]).astype(np.int32)
        np.testing.assert_allclose(m.predict(u, i), ref.predict(u, i), rtol=0, atol=pred_atol)
    return ref


def test_knn3_reference_kat(gpu):
    """tests/regressors_test.rs:177-196 on the GPU: the probe RMSE of the KAT
    configuration, and every table equal to the oracle's."""
    _, tr, pr = tiny()
    _compare(tr, KNN3_KAT_CFG)
    with Knn3Model(tr, Knn3Config(**KNN3_KAT_CFG)) as m:
        r = rmse(m.predict(pr.user_idxs, pr.item_idxs), pr)
    assert r == pytest.approx(KNN3_KAT_RMSE, abs=1e-6), r


@pytest.mark.parametrize("cfg", [
    {},                                                        # Knn3Config::default
    dict(threshold=0.5, k_min=3, k_max=12, shrinkage=50.0),   # threshold decides K
    dict(k_min=64, k_max=64, reg=0.5, x=0.3),                  # a full wave of candidates
])
def test_knn3_synthetic_matches_oracle(gpu, cfg):
    """600 x 150 x 30k set: baseline, candidates, weights and predictions."""
    tr, _ = synthetic("tiny_synth")
    _compare(tr, cfg)


def test_knn3_skips_test_rows(gpu):
    """Rows with is_test != 0 are left out of the baseline, the counts, the
    regressions and the lookups (src/knn3.rs:55, 134, 185, 244, 305)."""
    tr, _ = synthetic("tiny_synth")
    rng = np.random.default_rng(5)
    tr2 = dataset_from_rows(np.stack([tr.user_idxs, tr.item_idxs, tr.raw_ratings, tr.dates], 1), tr.n_users,
                            tr.n_items)
    tr2.residuals = (tr.raw_ratings.astype(np.float32) - np.float32(3.6)).astype(np.float32)
    tr2.is_test = (rng.random(len(tr2.user_idxs)) < 0.2).astype(np.int8)
    _compare(tr2, dict(k_min=5, k_max=20))


def test_knn3_benchmark_shape_matches_oracle(gpu):
    """The 10k x 2k x 1M set with Knn3Config::default (long user lists take
    the binary-search path of the normal-equation kernel)."""
    tr, _ = synthetic("small")
    assert np.max(np.bincount(tr.user_idxs)) > 512
    _compare(tr, {}, pred_atol=1e-4)


def test_knn3_rejects_unsupported_input(gpu):
    tr, _ = synthetic("tiny_synth")
    with pytest.raises(RuntimeError, match="k_max"):
        Knn3Model(tr, Knn3Config(k_max=65))
    rows = [(0, 1, 4, 1), (0, 1, 5, 2), (1, 0, 3, 3)]
    with pytest.raises(RuntimeError, match="twice"):
        Knn3Model(dataset_from_rows(rows, 2, 2), Knn3Config())


def test_knn3_edge_cases(gpu):
    """Items nobody rated (no candidates, predictions 0), a user whose every
    rating is a test row (bu 0, no lookups), and k_min 0 with a threshold no
    pair reaches (every item without neighbours)."""
    tr, _ = synthetic("tiny_synth")
    rows = np.stack([tr.user_idxs, tr.item_idxs, tr.raw_ratings, tr.dates], 1)
    tr2 = dataset_from_rows(rows, tr.n_users, tr.n_items + 3)
    tr2.residuals = (tr.raw_ratings.astype(np.float32) - np.float32(3.5)).astype(np.float32)
    tr2.is_test = (tr.user_idxs == 0).astype(np.int8)
    ref = _compare(tr2, dict(k_min=4, k_max=16))
    for m in range(tr.n_items, tr.n_items + 3):
        assert ref.neighbors[m] == []
    with Knn3Model(tr2, Knn3Config(k_min=4, k_max=16)) as m:
        p = m.predict(np.array([1, 2, 0], np.int32), np.array([tr.n_items, tr.n_items + 2, 5], np.int32))
    assert p[0] == 0.0 and p[1] == 0.0 and p[2] == 0.0
    with Knn3Model(tr2, Knn3Config(k_min=0, threshold=1e30)) as m:
        assert all(len(x) == 0 for x in m.neighbors())
        assert np.all(m.predict(tr2.user_idxs[:100], tr2.item_idxs[:100]) == 0.0)
