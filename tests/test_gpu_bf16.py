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


@pytest.mark.parametrize("f", [256, 512])
def test_bf16_frozen_sides_full_shape(gpu, f):
    """Mixed precision on the tsvdx4p schedule at the full Netflix shape: the
    user passes read q (and q_f) and the item pass reads p, p2 and ycache from
    BF16 copies of the FP32 masters (tx_set_frozen_precision).  After 4
    epochs the holdout RMSE stays within 1e-3 of the all-FP32 run; the
    predictions of the BF16-trained model (FP32 predict) against the FP64
    predict oracle on its own parameters, over a 1M-rating subset, are within
    the FP32 bounds; and its BF16 prediction copies stay within the BF16 bounds."""
    from tests.fixtures import full_data
    tr, pr = full_data(1)
    cfg = lm_config(n_feat=f, n_epochs=4, n_time_bins=30, nsvd_norm_exp=1.0, lr_yb=2e-4, reset_u_epoch=100)
    rmse, secs = {}, {}
    for prec in ("f32", "bf16"):
        m = tx.TxModel(tr, pr, cfg)
        m.set_schedule(mode=tx.SCHED_TSVDX4P)
        m.set_frozen_precision(prec)
        ms = []
        for e in range(1, cfg.n_epochs + 1):
            m.fit_epoch(epoch=e)
            ms.append(m.last_epoch_stats()["ms_total"])
        rmse[prec], secs[prec] = m.calc_rmse(pr), ms
        if prec == "f32":
            m.close()
    print(f"f={f} tsvdx4p epochs ms FP32 {[round(x, 1) for x in secs['f32']]} BF16 {[round(x, 1) for x in secs['bf16']]}; "
          f"holdout RMSE FP32 {rmse['f32']:.6f} BF16 {rmse['bf16']:.6f}")
    assert abs(rmse["bf16"] - rmse["f32"]) <= 1e-3
    # FP64 predict oracle on the BF16-trained model's parameters: the first 1M
    # held-out ratings
    n = 1_000_000
    assert pr.n_ratings >= n
    u, i, d = pr.user_idxs[:n], pr.item_idxs[:n], pr.dates[:n]
    o = O.OracleModel(cfg, tr, pr)
    for n in ALL_PARAMS:
        o.set(n, m.get(n))
    ref = o.predict_f64(u, i, d)
    e32 = m.predict(u, i, d).astype(np.float64) - ref
    m.set_table_precision("bf16")
    e16 = m.predict(u, i, d).astype(np.float64) - ref
    m.close()
    print(f"f={f} vs FP64 oracle on {len(u)} ratings: FP32 predict max {np.abs(e32).max():.2e} "
          f"rms {np.sqrt(np.mean(e32 ** 2)):.2e}; BF16 predict copies max {np.abs(e16).max():.2e} "
          f"rms {np.sqrt(np.mean(e16 ** 2)):.2e}")
    assert np.abs(e32).max() <= 1e-5 and np.sqrt(np.mean(e32 ** 2)) <= 1e-6
    assert np.abs(e16).max() <= 5e-2 and np.sqrt(np.mean(e16 ** 2)) <= 5e-3
