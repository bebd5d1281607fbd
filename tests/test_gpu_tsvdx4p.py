"""Tsvdx4pModel epochs and user-only (fold-in) epochs on the GPU against the
oracle's restatement of src/tsvdx4p.rs (needs an MI355X).

The GPU sums the NSVD1 accumulators with atomics and reduces dot products
across lanes, so it matches the oracle to rounding, not bit for bit; the
tolerances below are written for FP32 work over a few epochs."""
import numpy as np
import pytest

from netflix_prize_amd import tx
from oracle import oracle as O
from tests.fixtures import lm_config, synthetic, tiny

pytestmark = pytest.mark.gpu

ALL_PARAMS = [p for p in tx.PARAMS if p != "thresholds"]


def _p_config(base: dict, qf: bool = False, **kw):
    """A Tsvdx4p config in the low-memory form the GPU schedule runs, every
    epoch parallel (seq_epochs empty); without the q_f stream unless qf."""
    d = dict(base)
    d.update(low_memory=True, seq_epochs=[0, 0, 0, 0, 0])
    if not qf:
        d.update(lr_iqf=0.0, sigma_iqf=0.0)
    d.update(kw)
    return tx.TxConfig.from_dict(d)


def _compare(m_o, m_g, atol):
    for n in ALL_PARAMS:
        a, b = m_o.get(n), m_g.get(n)
        assert a.shape == b.shape, n
        if a.size:
            np.testing.assert_allclose(b, a, rtol=0, atol=atol, err_msg=n)


@pytest.mark.parametrize("qf", [False, True])
def test_tsvdx4p_tiny_matches_oracle(gpu, qf):
    """Three Tsvdx4pModel epochs on the reference's tiny fixture (its tsvdx4p
    KAT configuration, low-memory form, with and without the q_f stream):
    every table within 2e-5."""
    fx, tr, pr = tiny()
    cfg = _p_config(fx["kat"]["tsvdx4p"]["config"], qf=qf, lr_iqf=0.01 if qf else 0.0)
    m_o = O.OracleModel(cfg, tr, pr)
    m_g = tx.TxModel(tr, pr, cfg)
    m_g.set_schedule(mode=tx.SCHED_TSVDX4P)
    for e in range(1, 4):
        m_o.fit_epoch_p(e, 1)
        m_g.fit_epoch(epoch=e)
        _compare(m_o, m_g, 2e-5)
    assert m_g.calc_rmse(pr) == pytest.approx(m_o.calc_rmse(), abs=1e-6)


@pytest.mark.parametrize("qf", [False, True])
def test_tsvdx4p_synthetic_matches_oracle(gpu, qf):
    """Tsvdx4pModel on the 10k x 2k x 1M shape (f=32, nsvd_norm_exp 0.5; the
    lm set's q_f stream on or off): tables within 1e-4 and holdout RMSE within
    1e-5 after 3 epochs."""
    tr, pr = synthetic("small")
    off = {} if qf else dict(lr_iqf=0.0, sigma_iqf=0.0)
    cfg = lm_config(n_feat=32, n_epochs=3, n_time_bins=30, nsvd_norm_exp=0.5, n_threads=8,
                    seq_epochs=[0, 0, 0, 0, 0], **off)
    m_o = O.OracleModel(cfg, tr, pr)
    m_g = tx.TxModel(tr, pr, cfg)
    m_g.set_schedule(mode=tx.SCHED_TSVDX4P)
    for e in range(1, 4):
        m_o.fit_epoch_p(e, 8)
        m_g.fit_epoch(epoch=e)
    _compare(m_o, m_g, 1e-4)
    r_o, r_g = m_o.calc_rmse(), m_g.calc_rmse(pr)
    print(f"tsvdx4p 3 epochs: oracle {r_o:.6f} gpu {r_g:.6f}")
    assert abs(r_g - r_o) <= 1e-5


def test_tsvdx4p_rejects_unsupported_forms(gpu):
    tr, pr = synthetic("tiny_synth")
    for kw in (dict(low_memory=False), dict(ordinal_head=tx.OrdinalHeadConfig())):
        m = tx.TxModel(tr, pr, lm_config(n_feat=16, n_epochs=1, n_time_bins=30, **kw))
        with pytest.raises(RuntimeError, match="low_memory|ordinal"):
            m.set_schedule(mode=tx.SCHED_TSVDX4P)
        m.close()


@pytest.mark.parametrize("f", [1500, 4500])
def test_tsvdx4p_wide_rows_match_oracle(gpu, f):
    """Tsvdx4pModel epochs at f > 1024 (user and item passes on teams of 2 and
    5 waves, the lm set's q_f stream on) on the 600 x 150 x 30k set: tables
    within 1e-4 and holdout RMSE within 1e-5 after 2 epochs."""
    tr, pr = synthetic("tiny_synth")
    cfg = lm_config(n_feat=f, n_epochs=2, n_time_bins=30, nsvd_norm_exp=1.0, n_threads=8, lr_yb=0.0002,
                    seq_epochs=[0, 0, 0, 0, 0], lr_iqf=0.005, reset_u_epoch=1024)
    m_o = O.OracleModel(cfg, tr, pr)
    m_g = tx.TxModel(tr, pr, cfg)
    m_g.set_schedule(mode=tx.SCHED_TSVDX4P)
    for e in (1, 2):
        m_o.fit_epoch_p(e, 8)
        m_g.fit_epoch(epoch=e)
    _compare(m_o, m_g, 1e-4)
    assert m_g.calc_rmse(pr) == pytest.approx(m_o.calc_rmse(), abs=1e-5)


@pytest.mark.parametrize("shape,f", [("tiny_synth", 24), ("small", 24), ("tiny_synth", 1500)])
def test_fold_in_epochs_match_oracle(gpu, shape, f):
    """User-only epochs (items and y frozen) after one regular epoch: the GPU
    fold-in pass against the oracle's user pass without accumulators (f=1500:
    two-wave teams)."""
    tr, pr = synthetic(shape)
    cfg = lm_config(n_feat=f, n_epochs=3, n_time_bins=30)
    m_o = O.OracleModel(cfg, tr, pr)
    m_g = tx.TxModel(tr, pr, cfg)
    m_g.set_schedule(mode=tx.SCHED_SERIAL)
    m_o.fit_epoch(1)
    m_g.fit_epoch(epoch=1)
    frozen = {n: m_g.get(n) for n in ("ifeat", "yfeat", "ibias", "bit_bin", "ybias", "ibias_freq", "ifeat_freq")}
    for e in (2, 3):
        m_o.fit_users_epoch(e)
        m_g.fit_users_epoch(e)
    for n, v in frozen.items():
        np.testing.assert_array_equal(m_g.get(n), v, err_msg=f"{n} changed in a user-only epoch")
    _compare(m_o, m_g, 1e-5)
    assert m_g.calc_rmse(pr) == pytest.approx(m_o.calc_rmse(), abs=2e-6)
