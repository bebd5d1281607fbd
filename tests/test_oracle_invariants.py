"""Pins of the oracle's low-memory trainer (low_memory = true).

low_memory = true is the mode of every full-size tsvdx4 lm job, and no
reference test covers it (tests/regressors_test.rs:28-73 runs
low_memory = false only).  The KATs pin the low_memory = false path; these
checks tie the low-memory branch to it and to what the mathematics fixes:

(i)   with the day terms and frequency bins inert, low_memory = true and false
      train bit-identical tables (src/tx.rs:606-630, 834-873, 1025-1036 are
      the only places the flag acts);
(ii)  quirk Q1: low-memory training reads and writes frequency bin 0 only
      (src/tx.rs:834, 846-853, 873), non-low-memory training the day's bin;
(iii) the n_feat = 0 bias-only model equals an FP64 brute-force SGD built
      from dictionaries (the textbook biased baseline, then every bias term of
      the lm set);
(iv)  all learning rates 0 keep every table at its initial value, and the
      epoch-end ycache is the NSVD1 sum recomputed in FP64 over training plus
      every held-out row (src/tx.rs:391-418, quirk Q5);
(v)   the initial tables are the rand_array2 stream (src/lib.rs:64-67) in
      field order, without the yfeat_day block in low-memory mode
      (src/tx.rs:606-630).

Each check is also run with a mistake planted in the oracle
(``oracle.planted``) and must then fail.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.fixtures import synthetic, tiny

TABLES = ["ubias", "ibias", "ufeat", "ufeat2", "ifeat", "yfeat", "ycache", "tu_mean", "alpha_u",
          "but_bin", "bit_bin", "ybias", "ycache_bias", "ubias_day", "ibias_freq", "ifeat_freq", "cu",
          "cut", "gbias"]


def lm(**kw):
    """The tsvdx4 lm hyper-parameters (src/bin/tsvdx4-new.rs:130-154)."""
    from netflix_prize_amd.tx import TxConfig
    return TxConfig(**kw)


def train(cfg, tr, pr, epochs, plant=0):
    with O.planted(plant):
        m = O.OracleModel(cfg, tr, pr)
        for e in range(1, epochs + 1):
            m.fit_epoch(e)
    return m


def datasets():
    fx, tr, pr = tiny()
    yield "tiny_fixture", tr, pr
    tr, pr = synthetic("tiny_synth")
    yield "tiny_synth", tr, pr


# ---------------------------------------------------------------------------
# (i) low_memory = true equals the KAT-pinned low_memory = false path when the
#     day and frequency terms are inert
# ---------------------------------------------------------------------------

def _inert_day_terms(low_memory, n_feat):
    return lm(n_feat=n_feat, n_epochs=3, n_freq_bins=1, sigma_iqf=0.0, sigma_yd=0.0, lr_ud=0.0, lr_yd=0.0,
              low_memory=low_memory, reset_u_epoch=2)


@pytest.mark.parametrize("name,tr,pr", list(datasets()), ids=lambda x: x if isinstance(x, str) else "")
def test_low_memory_equals_day_path_with_day_terms_inert(name, tr, pr):
    a = train(_inert_day_terms(True, 10), tr, pr, 3)
    b = train(_inert_day_terms(False, 10), tr, pr, 3)
    for t in TABLES:
        np.testing.assert_array_equal(a.get(t), b.get(t), err_msg=t)
    assert a.calc_rmse() == b.calc_rmse()


def test_low_memory_equivalence_catches_a_planted_yfeat_day_draw():
    tr, pr = synthetic("tiny_synth")
    cfg = _inert_day_terms(True, 10).replace(sigma_iqf=0.005)
    a = train(cfg, tr, pr, 1, plant=2)
    b = train(cfg.replace(low_memory=False), tr, pr, 1)
    # with the plant both modes draw the same stream, so q_f agrees; the
    # unplanted low-memory model must not (it skips the I x f yfeat_day draws)
    np.testing.assert_array_equal(a.get("ifeat_freq"), b.get("ifeat_freq"))
    c = O.OracleModel(cfg, tr, pr)
    assert not np.array_equal(c.get("ifeat_freq"), b.get("ifeat_freq"))


# ---------------------------------------------------------------------------
# (ii) Q1: low-memory training uses frequency bin 0 only
# ---------------------------------------------------------------------------

def _freq_only(low_memory):
    # every step off but the frequency-bin item bias b_i,f
    zero = {k: 0.0 for k in ["lr_u", "lr_ud", "lr_u2", "lr_ub", "lr_ubd", "lr_i", "lr_ib", "lr_y", "lr_yb",
                             "lr_yd", "lr_tu", "lr_ti", "lr_ta", "lr_iqf", "lr_cu"]}
    return lm(n_feat=4, n_freq_bins=4, lr_ibf=5e-3, low_memory=low_memory, **zero)


def _moved_bins(m, n_items):
    return np.nonzero(np.any(m.get("ibias_freq").reshape(n_items, 4) != 0.0, axis=0))[0].tolist()


def _day_counts_reach_bin(tr, pr, k):
    m = O.OracleModel(_freq_only(True), tr, pr)
    _, _, cnts = m.sparse_ud()
    return any(O.lib().or_freq_bin(int(c), 4) >= k for c in cnts)


def test_low_memory_training_moves_frequency_bin_zero_only():
    tr, pr = synthetic("tiny_synth")
    assert _day_counts_reach_bin(tr, pr, 2)  # the data has days in bins 1.. so Q1 is visible
    lo = train(_freq_only(True), tr, pr, 2)
    hi = train(_freq_only(False), tr, pr, 2)
    assert _moved_bins(lo, tr.n_items) == [0]
    assert len(_moved_bins(hi, tr.n_items)) > 1


def test_frequency_bin_check_catches_a_planted_real_bin():
    tr, pr = synthetic("tiny_synth")
    lo = train(_freq_only(True), tr, pr, 2, plant=1)
    assert _moved_bins(lo, tr.n_items) != [0]


# ---------------------------------------------------------------------------
# (iii) the bias-only model against an FP64 brute force
# ---------------------------------------------------------------------------

def brute_force_bias_sgd(cfg, tr, pr, epochs):
    """FP64 SGD of the n_feat = 0 timeSVD++ model (the biased baseline with
    time and NSVD1 bias terms), from dictionaries keyed by user, item, day
    and bin: no SparseUD search, no CSR.  Returns the tables and the probe RMSE."""
    U, I, B, F = tr.n_users, tr.n_items, cfg.n_time_bins, cfg.n_freq_bins
    tu, it, dt = tr.user_idxs.tolist(), tr.item_idxs.tolist(), tr.dates.tolist()
    r = tr.residuals.astype(np.float64).tolist()
    pu, pi, pd = pr.user_idxs.tolist(), pr.item_idxs.tolist(), pr.dates.tolist()
    D = 1 + max(dt)
    mu = float(np.mean(tr.residuals.astype(np.float64)))
    rows = {}
    for k, u in enumerate(tu):
        rows.setdefault(u, []).append(k)
    held = {}
    for k, u in enumerate(pu):
        held.setdefault(u, []).append(pi[k])
    tmean = {u: sum(float(dt[k]) for k in ks) / len(ks) for u, ks in rows.items()}
    day_cnt = {}
    for u, d in list(zip(tu, dt)) + list(zip(pu, pd)):
        day_cnt[(u, d)] = day_cnt.get((u, d), 0) + 1

    def tbin(d):
        return min(d * B // D, B - 1)

    def dev(u, d):
        x = d - tmean.get(u, 0.0)
        return 0.0 if x == 0 else math.copysign(abs(x) ** cfg.beta, x)

    def fbin(c):
        return 0 if c <= 1 else min(int(math.floor(math.log2(c))), F - 1)

    bu, bi, al, cu = [0.0] * U, [0.0] * I, [0.0] * U, [1.0] * U
    but, bit, bif = {}, {}, {}
    bud, yb, ycb = {}, [0.0] * I, [0.0] * U
    n_held = {u: len(v) for u, v in held.items()}
    for e in range(1, epochs + 1):
        for u in O.get_users(U, cfg.shuffle_users, cfg.seed, e).tolist():
            ks = rows.get(u)
            if not ks:
                continue
            nrm = math.sqrt(len(ks) + (n_held.get(u, 0) if cfg.full_su else 0))
            items = [it[k] for k in ks] + (held.get(u, []) if cfg.full_su else [])
            ycb[u] = sum(yb[j] for j in items) / nrm
            sum_err = 0.0
            for k in ks:
                i, d = it[k], dt[k]
                b, v = tbin(d), dev(u, d)
                bu_t = bu[u] + but.get((u, b), 0.0) + al[u] * v + bud.get((u, d), 0.0) + ycb[u]
                bi_t = bi[i] + bit.get((i, b), 0.0) + bif.get((i, 0), 0.0)
                cu_t = cu[u]
                err = mu + bu_t + bi_t * cu_t - r[k]
                bu[u] -= cfg.lr_ub * err
                but[(u, b)] = but.get((u, b), 0.0) - cfg.lr_tu * err
                al[u] -= cfg.lr_ta * err * v
                bud[(u, d)] = bud.get((u, d), 0.0) - cfg.lr_ubd * err
                bi[i] -= cfg.lr_ib * err * cu_t
                bit[(i, b)] = bit.get((i, b), 0.0) - cfg.lr_ti * err * cu_t
                bif[(i, 0)] = bif.get((i, 0), 0.0) - cfg.lr_ibf * err * cu_t
                cu[u] -= cfg.lr_cu * (err * bi_t + cfg.reg_cu * (cu[u] - 1.0))
                sum_err += err
            for k in ks:
                yb[it[k]] -= cfg.lr_yb * sum_err / nrm
        # epoch-end rebuild over every training and held-out row
        for u in range(U):
            items = [it[k] for k in rows.get(u, [])] + held.get(u, [])
            if items:
                ycb[u] = sum(yb[j] for j in items) / math.sqrt(len(items))
    sse, n = 0.0, 0
    for k in range(len(pu)):
        if pr.is_test[k]:
            continue
        u, i, d = pu[k], pi[k], pd[k]
        b, v = tbin(d), dev(u, d)
        f = fbin(day_cnt[(u, d)])
        bu_t = bu[u] + but.get((u, b), 0.0) + al[u] * v + bud.get((u, d), 0.0) + ycb[u]
        bi_t = bi[i] + bit.get((i, b), 0.0) + bif.get((i, f), 0.0)
        sse += (mu + bu_t + bi_t * cu[u] - float(pr.residuals[k])) ** 2
        n += 1
    return {"ubias": np.array(bu), "ibias": np.array(bi), "alpha_u": np.array(al), "cu": np.array(cu),
            "ybias": np.array(yb), "ycache_bias": np.array(ycb)}, math.sqrt(sse / n)


BIAS_CASES = {
    # the textbook biased baseline mu + b_u + b_i
    "textbook": dict(lr_ub=0.0031, lr_ib=0.0036, lr_ubd=0.0, lr_yb=0.0, lr_ti=0.0, lr_ta=0.0, lr_ibf=0.0,
                     lr_cu=0.0, lr_tu=0.0),
    # every bias term of the lm set (lr_tu switched on too)
    "lm_biases": dict(lr_tu=0.0025, lr_yb=2.5e-3),
}


@pytest.mark.parametrize("case", list(BIAS_CASES))
def test_bias_only_model_matches_fp64_brute_force(case):
    tr, pr = synthetic("tiny_synth")
    cfg = lm(n_feat=0, n_epochs=3, n_time_bins=30, **BIAS_CASES[case])
    m = train(cfg, tr, pr, 3)
    ref, rmse = brute_force_bias_sgd(cfg, tr, pr, 3)
    # FP32 against FP64 over 3 epochs: observed 5e-9 in RMSE, 2e-6 in the tables
    assert abs(m.calc_rmse() - rmse) < 1e-7
    for t, v in ref.items():
        np.testing.assert_allclose(m.get(t), v, atol=1e-5, rtol=0, err_msg=t)


@pytest.mark.parametrize("plant", [4, 5, 6])
def test_brute_force_catches_planted_mistakes(plant):
    # 4: NSVD1-bias step skipped; 5: day range + 1 (quirk Q4); 6: alpha step without dev
    tr, pr = synthetic("tiny_synth")
    cfg = lm(n_feat=0, n_epochs=3, n_time_bins=30, **BIAS_CASES["lm_biases"])
    m = train(cfg, tr, pr, 3, plant=plant)
    ref, rmse = brute_force_bias_sgd(cfg, tr, pr, 3)
    worst = max(float(np.max(np.abs(m.get(t) - v))) for t, v in ref.items())
    assert abs(m.calc_rmse() - rmse) > 1e-7 or worst > 1e-5


# ---------------------------------------------------------------------------
# (iv) zero learning rates, and the ycache against the NSVD1 sum
# ---------------------------------------------------------------------------

ALL_LR = ["lr_u", "lr_ud", "lr_u2", "lr_ub", "lr_ubd", "lr_i", "lr_ib", "lr_y", "lr_yb", "lr_yd", "lr_tu",
          "lr_ti", "lr_ta", "lr_ibf", "lr_iqf", "lr_cu"]


@pytest.mark.parametrize("low_memory", [True, False])
def test_zero_learning_rates_keep_tables_at_init(low_memory):
    tr, pr = synthetic("tiny_synth")
    cfg = lm(n_feat=12, low_memory=low_memory, reset_u_epoch=99, **{k: 0.0 for k in ALL_LR})
    init = O.OracleModel(cfg, tr, pr)
    m = train(cfg, tr, pr, 2)
    for t in TABLES + (["yfeat_day", "ufeat_day"] if not low_memory else []):
        if t in ("ycache", "ycache_bias"):
            continue  # written by the epoch (checked below)
        np.testing.assert_array_equal(m.get(t), init.get(t), err_msg=t)


def nsvd1_sum_fp64(m, tr, pr, f):
    y = m.get("yfeat").astype(np.float64).reshape(tr.n_items, f)
    s = np.zeros((tr.n_users, f))
    np.add.at(s, tr.user_idxs, y[tr.item_idxs])
    np.add.at(s, pr.user_idxs, y[pr.item_idxs])
    n = tr.user_cnts.astype(np.float64) + pr.user_cnts
    return s / np.sqrt(np.maximum(n, 1))[:, None]


def test_ycache_after_epoch_is_the_nsvd1_sum_of_training_and_every_heldout_row():
    tr, pr = synthetic("tiny_synth")
    f = 12
    m = train(lm(n_feat=f, n_epochs=2), tr, pr, 2)
    np.testing.assert_allclose(m.get("ycache").reshape(-1, f), nsvd1_sum_fp64(m, tr, pr, f), atol=2e-7, rtol=1e-5)


def test_ycache_check_catches_a_planted_rebuild_without_heldout_rows():
    tr, pr = synthetic("tiny_synth")
    f = 12
    m = train(lm(n_feat=f, n_epochs=1), tr, pr, 1, plant=3)
    assert not np.allclose(m.get("ycache").reshape(-1, f), nsvd1_sum_fp64(m, tr, pr, f), atol=2e-7, rtol=1e-5)


# ---------------------------------------------------------------------------
# (v) the initial draws
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("low_memory", [True, False])
def test_initial_tables_follow_the_rng_stream_in_field_order(low_memory):
    tr, pr = synthetic("tiny_synth")
    f, F = 6, 4
    cfg = lm(n_feat=f, n_freq_bins=F, low_memory=low_memory, sigma_u=1.0, sigma_i=1.0, sigma_y=1.0,
             sigma_yd=1.0, sigma_iqf=1.0)
    m = O.OracleModel(cfg, tr, pr)
    U, I = tr.n_users, tr.n_items
    blocks = [("ufeat", U * f), ("ifeat", I * f), ("yfeat", I * f)]
    if not low_memory:
        blocks.append(("yfeat_day", I * f))
    blocks.append(("ifeat_freq", I * F * f))
    z = O.normal_fill(cfg.seed, 1.0, sum(n for _, n in blocks))
    at = 0
    for name, n in blocks:
        np.testing.assert_array_equal(m.get(name), z[at:at + n], err_msg=name)
        at += n
    assert np.all(m.get("ufeat2") == 0) and np.all(m.get("cu") == 1)


def test_init_stream_check_catches_a_planted_yfeat_day_draw():
    tr, pr = synthetic("tiny_synth")
    f, F = 6, 4
    cfg = lm(n_feat=f, n_freq_bins=F, sigma_iqf=1.0)
    with O.planted(2):
        m = O.OracleModel(cfg, tr, pr)
    U, I = tr.n_users, tr.n_items
    z = O.normal_fill(cfg.seed, 1.0, U * f + 2 * I * f + I * F * f)
    assert not np.array_equal(m.get("ifeat_freq"), z[U * f + 2 * I * f:])
