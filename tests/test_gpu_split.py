"""Split teams (tx_schedule.split_waves): a rating wave and a streaming wave
per user stream.  Where users share no item the schedule of users cannot
matter, so every table matches the serial oracle up to float association;
at the full shape the holdout curve ends within 1e-3 of the serial trainer's."""
import json
import os

import numpy as np
import pytest

from netflix_prize_amd import tx
from oracle import oracle as O
from tests.fixtures import lm_config
from tests.test_gpu_edges import _short_users

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("f,teams", [(24, 8), (200, 8), (256, 1)])
def test_split_teams_match_serial_on_disjoint_users(gpu, f, teams):
    tr, pr = _short_users()
    cfg = lm_config(n_feat=f, n_epochs=2, n_time_bins=30, reset_u_epoch=100)
    m_o = O.OracleModel(cfg, tr, pr)
    with tx.TxModel(tr, pr, cfg) as m_g:
        m_g.set_schedule(mode=tx.SCHED_STRATA, split_waves=True, blocks=teams)
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
        np.testing.assert_allclose(m_g.predict(pr.user_idxs, pr.item_idxs, pr.dates),
                                   m_o.predict(pr.user_idxs, pr.item_idxs, pr.dates), atol=1e-5)


def test_split_teams_schedule_rules(gpu):
    tr, pr = _short_users()
    with tx.TxModel(tr, pr, lm_config(n_feat=24, n_epochs=1, n_time_bins=30)) as m:
        for bad in (dict(n_strata=8), dict(user_order=tx.USER_ORDER_BUCKETS), dict(y_update=0)):
            with pytest.raises(tx.TxError, match="split_waves"):
                m.set_schedule(mode=tx.SCHED_STRATA, split_waves=True, **bad)
        m.set_schedule(mode=tx.SCHED_STRATA, split_waves=True, use_graph=True)
        m.fit_epoch(epoch=1)
        assert np.isfinite(m.calc_rmse(pr))
    with tx.TxModel(tr, pr, lm_config(n_feat=24, n_epochs=1, n_time_bins=30, low_memory=False)) as m:
        with pytest.raises(tx.TxError, match="split_waves"):
            m.set_schedule(mode=tx.SCHED_STRATA, split_waves=True)


def test_full_shape_split_teams_match_serial_reference(gpu):
    """Plain form, f=256, 20 epochs at the full shape with split teams (the
    default teams in flight), against the serial oracle's holdout RMSE."""
    from tests.fixtures import full_data
    golden = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "oracle_curve_full_f256.json")))
    tr, pr = full_data(golden["seed"])
    cfg = tx.plain_form(lm_config(n_feat=256, n_epochs=20, n_time_bins=30, reset_u_epoch=10_000))
    with tx.TxModel(tr, pr, cfg) as m:
        m.set_schedule(mode=tx.SCHED_STRATA, split_waves=True, use_graph=True)
        curve = []
        for e in range(1, 21):
            m.fit_epoch(epoch=e)
            curve.append(m.calc_rmse(pr))
    ref = [x["holdout_rmse"] for x in golden["epochs"]]
    print("gpu", [round(x, 5) for x in curve])
    print("ref", [round(x, 5) for x in ref])
    assert abs(curve[-1] - ref[-1]) <= 1e-3
