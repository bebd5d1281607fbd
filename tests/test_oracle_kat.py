"""The oracle against the reference's own golden values and published vectors.

These pin the CPU restatement (oracle/tx_oracle.c) before anything is checked
against it.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.fixtures import tiny


def test_chacha20_published_zero_key_vector():
    # RFC 7539 / djb ChaCha20, key = 0, nonce = 0, block 0 keystream.
    expect = bytes.fromhex(
        "76b8e0ada0f13d90405d6ae55386bd28bdd219b8a08ded1aa836efcc8b770dc7"
        "da41597c5157488d7724e03fb8d84a376a43b8f41518a11cc387b669b2ee6586")
    got = O.chacha_block(np.zeros(8, np.uint32), 0, 20).astype("<u4").tobytes()
    assert got == expect


def test_ziggurat_table_matches_published_prefix():
    # Leading entries of rand_distr's ZIG_NORM_X / ZIG_NORM_F tables.
    x, f = O.zig_tables()
    assert x[:4].tolist() == [3.910757959537090045, 3.654152885361008796,
                              3.449278298560964462, 3.320244733839166074]
    assert f[:4].tolist() == [0.000477467764586655, 0.001260285930498598,
                              0.002609072746106363, 0.004037972593371872]
    assert x[256] == 0.0 and f[256] == 1.0
    assert np.all(np.diff(x[1:]) < 0)


def test_normal_fill_statistics():
    z = O.normal_fill(7, 1.0, 200_000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01
    # sigma scales the same draws; sigma = 0 still consumes the stream
    z2 = O.normal_fill(7, 0.5, 1000)
    np.testing.assert_array_equal(z2, (np.float32(0.5) * z[:1000]).astype(np.float32))
    assert np.all(O.normal_fill(7, 0.0, 10) == 0)


def test_get_users_is_a_permutation_and_seeded():
    a = O.get_users(1000, True, 42, 3)
    assert sorted(a.tolist()) == list(range(1000))
    assert not np.array_equal(a, O.get_users(1000, True, 42, 4))
    assert np.array_equal(a, O.get_users(1000, True, 42, 3))
    assert np.array_equal(O.get_users(50, False, 42, 3), np.arange(50))


def test_tsvdx4_tiny_probe_rmse_kat():
    fx, tr, pr = tiny()
    kat = fx["kat"]["tsvdx4"]
    m = O.OracleModel(kat["config"], tr, pr)
    for e in range(1, kat["config"]["n_epochs"] + 1):
        m.fit_epoch(e)
    assert m.calc_rmse() == kat["probe_rmse"]


def test_tsvdx4p_tiny_probe_rmse_kat():
    fx, tr, pr = tiny()
    kat = fx["kat"]["tsvdx4p"]
    m = O.OracleModel(kat["config"], tr, pr)
    for e in range(1, kat["config"]["n_epochs"] + 1):
        m.fit_epoch_p(e, 1)
    assert m.calc_rmse() == kat["probe_rmse"]


@pytest.mark.parametrize("which", ["tsvdx5", "tsvdx6"])
def test_neighbourhood_tiny_probe_rmse_kat(which):
    """tests/regressors_test.rs:95-155: TxModel with the global (tsvdx5) and
    the global + same-day (tsvdx6) neighbourhood terms, 5 sampled neighbours."""
    fx, tr, pr = tiny()
    kat = fx["kat"][which]
    m = O.OracleModel(kat["config"], tr, pr)
    for e in range(1, kat["config"]["n_epochs"] + 1):
        m.fit_epoch(e)
    assert m.calc_rmse() == kat["probe_rmse"]


@pytest.mark.parametrize("plant", [dict(w_nbr=1.0), dict(max_neighbors=4), dict(lambda1=20.0)])
@pytest.mark.parametrize("which", ["tsvdx5", "tsvdx6"])
def test_neighbourhood_kat_breaks_under_planted_mistake(which, plant):
    """The neighbourhood steps' weight, the sample size and the baseline's
    shrinkage each move the KAT."""
    fx, tr, pr = tiny()
    cfg = dict(fx["kat"][which]["config"], **plant)
    m = O.OracleModel(cfg, tr, pr)
    for e in range(1, cfg["n_epochs"] + 1):
        m.fit_epoch(e)
    assert m.calc_rmse() != fx["kat"][which]["probe_rmse"]


def test_neighbourhood_subscores_sum_to_predict():
    """predict = bias + mf + nsvd1 + nbr (src/tx.rs:663-724), and the nbr part
    is nonzero once the w and c tables have trained."""
    fx, tr, pr = tiny()
    cfg = fx["kat"]["tsvdx6"]["config"]
    m = O.OracleModel(cfg, tr, pr)
    for e in range(1, cfg["n_epochs"] + 1):
        m.fit_epoch(e)
    b, mf, ns, nb = m.predict_subscores(pr.user_idxs, pr.item_idxs, pr.dates)
    assert np.any(nb != 0)
    p = m.predict(pr.user_idxs, pr.item_idxs, pr.dates)
    np.testing.assert_allclose(b + mf + ns + nb, p, atol=1e-5)
    np.testing.assert_allclose(m.predict_f64(pr.user_idxs, pr.item_idxs, pr.dates), p, atol=1e-5)


@pytest.mark.parametrize("which", ["tsvdx4", "tsvdx4p"])
def test_kat_breaks_under_planted_mistake(which):
    """A dropped term must move the KAT: zeroing lr_ub changes the result."""
    fx, tr, pr = tiny()
    cfg = dict(fx["kat"][which]["config"], lr_ub=0.0)
    m = O.OracleModel(cfg, tr, pr)
    for e in range(1, cfg["n_epochs"] + 1):
        (m.fit_epoch_p(e, 1) if which == "tsvdx4p" else m.fit_epoch(e))
    assert m.calc_rmse() != fx["kat"][which]["probe_rmse"]


def test_fixture_invariants():
    # tests/lib_test.rs:6-44, 85-94
    fx, tr, pr = tiny()
    keys = list(zip(tr.user_idxs.tolist(), tr.dates.tolist()))
    assert keys == sorted(keys)
    assert pr.item_idxs.tolist() == sorted(pr.item_idxs.tolist())
    pairs = set()
    for d in (tr, pr):
        for u, i in zip(d.user_idxs.tolist(), d.item_idxs.tolist()):
            assert (u, i) not in pairs
            pairs.add((u, i))
    m = O.OracleModel(fx["kat"]["tsvdx4"]["config"], tr, pr)
    assert m.get("gbias")[0] == np.float32(np.mean(tr.residuals.astype(np.float64)))


def test_unsorted_train_is_rejected():
    # calc_user_offsets panics on unsorted input (tests/lib_test.rs:46-51)
    fx, tr, pr = tiny()
    with pytest.raises(RuntimeError, match="not user-sorted"):
        O.OracleModel(fx["kat"]["tsvdx4"]["config"], pr, tr)
