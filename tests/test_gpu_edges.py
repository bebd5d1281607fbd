"""Edge cases of the trainer and predictor on the GPU: empty and invalid
queries, inputs the host rejects before any device work (the reference's
calc_user_offsets panics on unsorted input, tests/lib_test.rs:46-51), and users
with one to three training ratings, held-out-only days and held-out-only items
against the serial oracle."""
import numpy as np
import pytest

from netflix_prize_amd import tx
from oracle import oracle as O
from tests.fixtures import lm_config, tiny

pytestmark = pytest.mark.gpu


def test_predict_empty_and_out_of_range_queries(gpu):
    fx, tr, pr = tiny()
    cfg = tx.TxConfig.from_dict(fx["kat"]["tsvdx4"]["config"])
    with tx.TxModel(tr, pr, cfg) as m:
        m.fit_epoch(tr, pr, 1)
        ref = m.predict(pr.user_idxs, pr.item_idxs, pr.dates)
        empty = m.predict(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.int16))
        assert empty.shape == (0,)
        for u, i in ((tr.n_users, 0), (-1, 0), (0, tr.n_items), (0, -5)):
            with pytest.raises(tx.TxError, match="out of range"):
                m.predict(np.array([0, u], np.int32), np.array([0, i], np.int32), np.array([100, 100], np.int16))
        # a rejected call leaves no error behind: the same queries again
        np.testing.assert_array_equal(m.predict(pr.user_idxs, pr.item_idxs, pr.dates), ref)
        assert np.isfinite(m.calc_rmse(pr))


def test_create_rejects_bad_rating_sets(gpu):
    fx, tr, pr = tiny()
    cfg = tx.TxConfig.from_dict(fx["kat"]["tsvdx4"]["config"])
    bad_item = tx.Dataset.from_arrays(tr.user_idxs, np.where(np.arange(tr.n_ratings) == 3, tr.n_items, tr.item_idxs),
                                      tr.raw_ratings, tr.dates, tr.n_users, tr.n_items)
    with pytest.raises(tx.TxError, match="item index"):
        tx.TxModel(bad_item, pr, cfg)
    order = np.argsort(tr.item_idxs, kind="stable")
    unsorted = tx.Dataset.from_arrays(tr.user_idxs[order], tr.item_idxs[order], tr.raw_ratings[order],
                                      tr.dates[order], tr.n_users, tr.n_items)
    with pytest.raises(tx.TxError, match="not user-sorted"):
        tx.TxModel(unsorted, pr, cfg)
    # the library still works after the rejected calls
    with tx.TxModel(tr, pr, cfg) as m:
        m.fit_epoch(tr, pr, 1)
        assert np.isfinite(m.calc_rmse(pr))


def _short_users(n_users=64, seed=11):
    """Users that share no item, with 1, 2, 3 or 17 training ratings, one or
    two held-out ratings (some on a day without training ratings), and one
    item per user that only a held-out rating names."""
    rng = np.random.default_rng(seed)
    per = 24
    rows = {"tr": [], "pr": []}
    for u in range(n_users):
        n_tr = (1, 2, 3, 17)[u % 4]
        items = u * per + rng.permutation(per)
        days = np.sort(rng.choice(np.arange(200, 2100), n_tr, replace=True))
        tr_rows = sorted(zip(days, items[:n_tr]))
        for d, i in tr_rows:
            rows["tr"].append((u, i, int(rng.integers(1, 6)), d))
        n_h = 1 + u % 2
        for k in range(n_h):
            d = int(days[-1]) if k == 0 else 2200          # a training day, then a held-out-only day
            rows["pr"].append((u, items[n_tr + k], int(rng.integers(1, 6)), d))
    out = []
    for part in ("tr", "pr"):
        a = np.array(rows[part], np.int64)
        out.append(tx.Dataset.from_arrays(a[:, 0], a[:, 1], a[:, 2], a[:, 3], n_users, n_users * per))
    return out


@pytest.mark.parametrize("low_memory", [True, False])
def test_short_users_parallel_matches_serial(gpu, low_memory):
    """The default parallel schedule on users of 1 to 17 training ratings: every
    table within 1e-3 of its movement from the serial oracle over two epochs
    (users share no item, so only float association separates them)."""
    tr, pr = _short_users()
    cfg = lm_config(n_feat=24, n_epochs=2, n_time_bins=30, low_memory=low_memory, lr_ud=0.01, lr_yd=0.01,
                    sigma_yd=0.05, reset_u_epoch=100)
    m_o = O.OracleModel(cfg, tr, pr)
    with tx.TxModel(tr, pr, cfg) as m_g:   # the same initial draws as the oracle's
        names = [n for n in tx.PARAMS if n not in tx.NBR_PARAMS and n != "thresholds"]
        init = {n: m_o.get(n).copy() for n in names}
        for e in (1, 2):
            m_o.fit_epoch(e)
            m_g.fit_epoch(epoch=e)
        for n in names:
            o, g = m_o.get(n), m_g.get(n)
            move = float(np.max(np.abs(o.astype(np.float64) - init[n]))) if o.size else 0.0
            err = float(np.max(np.abs(g.astype(np.float64) - o))) if o.size else 0.0
            assert err <= 1e-3 * move + 1e-6, (n, err, move)
        g = m_g.predict(pr.user_idxs, pr.item_idxs, pr.dates)
        o = m_o.predict(pr.user_idxs, pr.item_idxs, pr.dates)
        np.testing.assert_allclose(g, o, atol=1e-5)
