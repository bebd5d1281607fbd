"""Concurrent reads of one model from several host threads (needs an MI355X).

The reference calls `predict(&self)` from rayon workers (gen_preds_parallel,
src/lib.rs:710-722), so a model behind the `Regressor` trait must be `Sync`.
tx_predict, tx_predict_subscores and tx_calc_rmse each stage through device
memory of their own; here 8 threads call them on one ctx at once (ctypes
drops the GIL for the call) and every result must equal the single-threaded
one bit for bit.
"""
import threading

import numpy as np
import pytest

from netflix_prize_amd import tx
from tests.fixtures import lm_config, synthetic

pytestmark = pytest.mark.gpu


def test_eight_threads_predict_on_one_ctx(gpu):
    tr, pr = synthetic("small")
    cfg = lm_config(n_feat=32, n_epochs=2, n_time_bins=30)
    with tx.TxModel(tr, pr, cfg) as m:
        for e in (1, 2):
            m.fit_epoch(epoch=e)
        u, i, d = tr.user_idxs, tr.item_idxs, tr.dates
        ref = m.predict(u, i, d)
        ref_sub = m.predict_subscores(pr.user_idxs, pr.item_idxs, pr.dates)
        ref_rmse = m.calc_rmse(pr)
        n = len(u)
        errors = []

        def worker(k):
            try:
                rng = np.random.default_rng(k)
                for rep in range(6):
                    lo = int(rng.integers(0, n - 1))
                    hi = min(n, lo + int(rng.integers(1, 400_000)))
                    got = m.predict(u[lo:hi], i[lo:hi], d[lo:hi])
                    if not np.array_equal(got, ref[lo:hi]):
                        errors.append(f"thread {k} rep {rep}: predict rows {lo}:{hi} differ")
                    if rep % 3 == 0:
                        sub = m.predict_subscores(pr.user_idxs, pr.item_idxs, pr.dates)
                        if not all(np.array_equal(a, b) for a, b in zip(sub, ref_sub)):
                            errors.append(f"thread {k} rep {rep}: subscores differ")
                    if rep % 3 == 1:
                        r = m.calc_rmse(pr)
                        if abs(r - ref_rmse) > 1e-12:   # FP64 sums: block order varies
                            errors.append(f"thread {k} rep {rep}: rmse {r!r} vs {ref_rmse!r}")
            except Exception as e:  # reported below
                errors.append(f"thread {k}: {e!r}")

        threads = [threading.Thread(target=worker, args=(k,)) for k in range(8)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        assert not errors, errors[:5]
        # training after the concurrent reads still sees a consistent model
        m.fit_epoch(epoch=3)
        assert np.isfinite(m.calc_rmse(pr))
