"""tx_predict_noalloc, the entry point that allocates nothing (needs an
MI355X): its predictions equal tx_predict's bit for bit, on two calls in a
row, the second with a smaller input (the scratch array keeps the first
call's contents), and a batch above its maximum is refused."""
import numpy as np
import pytest

from netflix_prize_amd import tx
from tests.fixtures import lm_config, synthetic

pytestmark = pytest.mark.gpu


def test_noalloc_predict_equals_predict(gpu):
    tr, pr = synthetic("small")
    cfg = lm_config(n_feat=64, n_epochs=2, n_time_bins=30)
    with tx.TxModel(tr, pr, cfg) as m:
        for e in (1, 2):
            m.fit_epoch(epoch=e)
        n1 = tx.NOALLOC_MAX_ROWS          # a full batch first (training rows repeated)
        u, i, d = (np.resize(a, n1 + 1) for a in (tr.user_idxs, tr.item_idxs, tr.dates))
        a = m.predict_noalloc(u[:n1], i[:n1], d[:n1])
        np.testing.assert_array_equal(a, m.predict(u[:n1], i[:n1], d[:n1]))
        lo, hi = 12_345, 12_345 + 777     # then a smaller one at another offset
        b = m.predict_noalloc(u[lo:hi], i[lo:hi], d[lo:hi])
        np.testing.assert_array_equal(b, m.predict(u[lo:hi], i[lo:hi], d[lo:hi]))
        np.testing.assert_array_equal(b, a[lo:hi])
        # BF16 tables: refreshed from the masters without allocating either
        m.set_table_precision("bf16")
        c = m.predict_noalloc(pr.user_idxs, pr.item_idxs, pr.dates)
        np.testing.assert_array_equal(c, m.predict(pr.user_idxs, pr.item_idxs, pr.dates))
        m.set_table_precision("f32")
        with pytest.raises(tx.TxError, match="exceed the maximum"):
            m.predict_noalloc(u[:n1 + 1], i[:n1 + 1], d[:n1 + 1])
        assert len(m.predict_noalloc(u[:0], i[:0], d[:0])) == 0
