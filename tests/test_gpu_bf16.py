"""BF16 factor tables for predictions (FP32 masters) against the FP64 predict
oracle on identical parameters (needs an MI355X)."""
import numpy as np
import pytest

from netflix_prize_amd import tx
from oracle import oracle as O
from tests.fixtures import lm_config, synthetic

pytestmark = pytest.mark.gpu

ALL_PARAMS = [p for p in tx.PARAMS if p != "thresholds"]


@pytest.mark.parametrize("f", [256, 512])
def test_bf16_tables_prediction_error(gpu, f):
    """Per-prediction max and RMS error vs the FP64 oracle with FP32 and with
    BF16 tables, and the holdout RMSE each gives (10k x 2k x 1M, 3 epochs)."""
    tr, pr = synthetic("small")
    cfg = lm_config(n_feat=f, n_epochs=3, n_time_bins=30)
    m = tx.TxModel(tr, pr, cfg)
    for e in range(1, 4):
        m.fit_epoch(epoch=e)
    o = O.OracleModel(cfg, tr, pr)
    for n in ALL_PARAMS:
        o.set(n, m.get(n))
    u, i, d = pr.user_idxs, pr.item_idxs, pr.dates
    ref = o.predict_f64(u, i, d)
    g32 = m.predict(u, i, d).astype(np.float64)
    r32 = m.calc_rmse(pr)
    m.set_table_precision("bf16")
    g16 = m.predict(u, i, d).astype(np.float64)
    r16 = m.calc_rmse(pr)
    m.set_table_precision("f32")
    assert m.calc_rmse(pr) == pytest.approx(r32, abs=1e-12)   # FP64 atomic sums: order varies
    e32, e16 = g32 - ref, g16 - ref
    print(f"f={f}: FP32 max {np.abs(e32).max():.2e} rms {np.sqrt(np.mean(e32**2)):.2e} rmse {r32:.6f}; "
          f"BF16 max {np.abs(e16).max():.2e} rms {np.sqrt(np.mean(e16**2)):.2e} rmse {r16:.6f}")
    assert np.abs(e32).max() <= 1e-5 and np.sqrt(np.mean(e32 ** 2)) <= 1e-6
    assert np.abs(e16).max() <= 5e-2 and np.sqrt(np.mean(e16 ** 2)) <= 5e-3
    assert abs(r16 - r32) <= 1e-3
