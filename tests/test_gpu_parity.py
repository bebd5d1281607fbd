"""HIP path against the oracle, through the C ABI (needs an MI355X)."""
import numpy as np
import pytest

from netflix_prize_amd import tx
from oracle import oracle as O
from tests.fixtures import lm_config, synthetic, tiny

pytestmark = pytest.mark.gpu

ALL_PARAMS = [p for p in tx.PARAMS if p != "thresholds"]


def _oracle_and_gpu(cfg, tr, pr, **sched):
    m_o = O.OracleModel(cfg, tr, pr)
    m_g = tx.TxModel(tr, pr, cfg)
    if sched:
        m_g.set_schedule(**sched)
    return m_o, m_g


def _assert_params_equal(m_o, m_g, names=ALL_PARAMS, atol=0.0):
    for n in names:
        a, b = m_o.get(n), m_g.get(n)
        assert a.shape == b.shape, n
        if atol == 0.0:
            np.testing.assert_array_equal(a, b, err_msg=n)
        else:
            np.testing.assert_allclose(b, a, rtol=0, atol=atol, err_msg=n)


def test_tiny_kat_on_gpu_serial(gpu):
    """The reference's tsvdx4 KAT (tests/regressors_test.rs:75-93), trained on
    the GPU in the serial schedule, evaluated by the GPU calc_rmse."""
    fx, tr, pr = tiny()
    cfg = tx.TxConfig.from_dict(fx["kat"]["tsvdx4"]["config"])
    m = tx.TxModel(tr, pr, cfg)
    m.set_schedule(mode=tx.SCHED_SERIAL)
    for e in range(1, cfg.n_epochs + 1):
        m.fit_epoch(tr, pr, e)
    rmse = m.calc_rmse(pr)
    assert abs(rmse - fx["kat"]["tsvdx4"]["probe_rmse"]) < 1e-7, rmse


def test_tiny_serial_params_bit_identical(gpu):
    fx, tr, pr = tiny()
    cfg = tx.TxConfig.from_dict(fx["kat"]["tsvdx4"]["config"])
    m_o, m_g = _oracle_and_gpu(cfg, tr, pr, mode=tx.SCHED_SERIAL)
    _assert_params_equal(m_o, m_g)          # init draws
    for e in range(1, cfg.n_epochs + 1):
        m_o.fit_epoch(e)
        m_g.fit_epoch(epoch=e)
        _assert_params_equal(m_o, m_g)


@pytest.mark.parametrize("low_memory", [True, False])
def test_serial_schedule_matches_oracle(gpu, low_memory):
    tr, pr = synthetic("tiny_synth")
    cfg = lm_config(n_feat=24, n_epochs=2, n_time_bins=30, low_memory=low_memory, reset_u_epoch=2)
    m_o, m_g = _oracle_and_gpu(cfg, tr, pr, mode=tx.SCHED_SERIAL)
    for e in (1, 2):
        m_o.fit_epoch(e)
        m_g.fit_epoch(epoch=e)
    _assert_params_equal(m_o, m_g)
    # parameters are bit-identical; predict reduces the dot product across lanes
    assert m_g.calc_rmse(pr) == pytest.approx(m_o.calc_rmse(), abs=1e-7)


def test_predict_matches_oracle_and_fp64(gpu):
    """Identical parameters -> predictions within 1e-5 of the FP32 predictor
    and within 1e-6 RMS of the FP64 predict oracle."""
    tr, pr = synthetic("small")
    cfg = lm_config(n_feat=32, n_epochs=1, n_time_bins=30)
    m_o, m_g = _oracle_and_gpu(cfg, tr, pr)
    m_o.fit_epoch(1)
    for n in ALL_PARAMS:
        m_g.set(n, m_o.get(n))
    for ds in (pr, tr):
        u, i, d = ds.user_idxs, ds.item_idxs, ds.dates
        g = m_g.predict(u, i, d)
        o = m_o.predict(u, i, d)
        o64 = m_o.predict_f64(u, i, d)
        assert np.max(np.abs(g - o)) <= 1e-5
        assert np.sqrt(np.mean((g.astype(np.float64) - o64) ** 2)) <= 1e-6
    b, mf, ns = m_g.predict_subscores(pr.user_idxs, pr.item_idxs, pr.dates)
    ob, omf, ons = m_o.predict_subscores(pr.user_idxs, pr.item_idxs, pr.dates)
    for x, y in ((b, ob), (mf, omf), (ns, ons)):
        assert np.max(np.abs(x - y)) <= 1e-5
    assert m_g.calc_rmse(pr) == pytest.approx(m_o.calc_rmse(), abs=1e-6)


def test_full_shape_predict_fp64_oracle_on_1m_subset(gpu):
    """Parity (a) at the full Netflix shape (the f=64 configuration): after
    two epochs of the default parallel schedule, the GPU predictions for the
    first 1M held-out ratings against the FP64 predict oracle evaluated on
    the GPU model's own parameters: within 1e-5 per rating and 1e-6 RMS; the
    FP32 oracle predictor within 1e-5."""
    from tests.fixtures import full_data
    tr, pr = full_data(1)
    cfg = lm_config(n_feat=64, n_epochs=2, n_time_bins=30)
    n = 1_000_000
    u, i, d = pr.user_idxs[:n], pr.item_idxs[:n], pr.dates[:n]
    with tx.TxModel(tr, pr, cfg) as m:
        for e in (1, 2):
            m.fit_epoch(epoch=e)
        g = m.predict(u, i, d).astype(np.float64)
        o = O.OracleModel(cfg, tr, pr)
        for name in ALL_PARAMS:
            o.set(name, m.get(name))
    ref = o.predict_f64(u, i, d)
    f32 = o.predict(u, i, d).astype(np.float64)
    err = g - ref
    print(f"vs FP64 oracle on {n} ratings: max {np.abs(err).max():.2e} rms {np.sqrt(np.mean(err ** 2)):.2e}; "
          f"vs FP32 oracle max {np.abs(g - f32).max():.2e}")
    assert np.abs(err).max() <= 1e-5 and np.sqrt(np.mean(err ** 2)) <= 1e-6
    assert np.abs(g - f32).max() <= 1e-5


@pytest.mark.parametrize("n_strata", [1, 8])
def test_parallel_schedule_holdout_rmse(gpu, n_strata):
    """Parallel schedules train to the serial reference's holdout RMSE
    (within 0.001) on the 10k x 2k x 1M shape."""
    tr, pr = synthetic("small")
    cfg = lm_config(n_feat=32, n_epochs=5, n_time_bins=30)
    m_o, m_g = _oracle_and_gpu(cfg, tr, pr, mode=tx.SCHED_STRATA, n_strata=n_strata)
    for e in range(1, cfg.n_epochs + 1):
        m_o.fit_epoch(e)
        m_g.fit_epoch(epoch=e)
    r_o, r_g = m_o.calc_rmse(), m_g.calc_rmse(pr)
    print(f"strata={n_strata}: oracle {r_o:.6f} gpu {r_g:.6f}")
    assert abs(r_g - r_o) <= 1e-3


def _full_data(seed):
    from tests.fixtures import full_data
    return full_data(seed)


@pytest.mark.parametrize("f", [64, 256])
def test_full_shape_holdout_rmse_matches_serial_reference(gpu, f):
    """Parity (b) at the full Netflix shape: plain form, 20 epochs in the
    default parallel schedule vs the serial oracle's holdout RMSE
    (tests/golden/oracle_curve_full_f{64,256}.json, made by
    make_oracle_curves.py): within 1e-3 after the last epoch."""
    import json
    import os
    golden = json.load(open(os.path.join(os.path.dirname(__file__), "golden", f"oracle_curve_full_f{f}.json")))
    assert len(golden["epochs"]) == 20
    tr, pr = _full_data(golden["seed"])
    assert tr.n_ratings == golden["n_train"] and pr.n_ratings == golden["n_heldout"]
    cfg = tx.plain_form(lm_config(n_feat=f, n_epochs=20, n_time_bins=30, reset_u_epoch=10_000))
    with tx.TxModel(tr, pr, cfg) as m:
        curve = []
        for e in range(1, 21):
            m.fit_epoch(epoch=e)
            curve.append(m.calc_rmse(pr))
    ref = [x["holdout_rmse"] for x in golden["epochs"]]
    print("gpu", [round(x, 5) for x in curve])
    print("ref", [round(x, 5) for x in ref])
    assert abs(curve[-1] - ref[-1]) <= 1e-3


def test_autotune_leaves_parameters_and_picks_bounded_schedule(gpu):
    """The autotuner times candidate schedules from a snapshot of every table
    and restores it: the model's parameters are bit-identical afterwards, the
    chosen schedule stays within the fidelity bound on teams in flight, and
    training then proceeds as usual."""
    tr, pr = synthetic("small")
    cfg = lm_config(n_feat=32, n_epochs=2, n_time_bins=30)
    m = tx.TxModel(tr, pr, cfg)
    m.fit_epoch(epoch=1)
    before = {n: m.get(n) for n in ALL_PARAMS}
    best = m.autotune(trial_epochs=1)
    for n, v in before.items():
        np.testing.assert_array_equal(m.get(n), v, err_msg=n)
    teams = best["blocks"] * (1 if best["split_waves"] else 4)   # split teams: one per block
    assert best["user_order"] in (0, 2) and 0 < teams <= 3072 + 32, best
    # every candidate is reported, the stratum grids among them, and the
    # applied schedule is the one trial marked chosen (the fastest)
    trials = m.autotune_trials()
    # four orders x three team counts x two team forms, then two grids
    assert len(trials) == 26 and {t["n_strata"] for t in trials} == {1, 8, 64}
    assert {t["split_waves"] for t in trials if t["n_strata"] == 1} == {0, 1}
    chosen = [t for t in trials if t["chosen"]]
    assert len(chosen) == 1 and chosen[0]["ms"] == min(t["ms"] for t in trials)
    assert (chosen[0]["n_strata"], chosen[0]["user_order"], chosen[0]["split_waves"]) == \
        (best["n_strata"], best["user_order"], best["split_waves"])
    m.fit_epoch(epoch=2)
    assert m.calc_rmse(pr) < 1.1


def _disjoint_users(n_users=48, per_user=40, n_train=32, seed=3):
    """Users that share no item: every update of one user touches rows no
    other user reads, so any schedule of users gives the serial result up to
    float association.  Training rows sorted by (user, date, item); held-out
    items on training days and on days of their own."""
    rng = np.random.default_rng(seed)
    rows = {"tr": [], "pr": []}
    for u in range(n_users):
        items = u * per_user + rng.permutation(per_user)
        days = np.sort(rng.choice(np.arange(100, 2000), int(rng.integers(2, 9)), replace=False))
        extra = int(rng.integers(2000, 2200))  # a day with held-out ratings only
        td = rng.choice(days, n_train)
        hd = np.where(rng.random(per_user - n_train) < 0.3, extra, rng.choice(days, per_user - n_train))
        for part, its, ds in (("tr", items[:n_train], td), ("pr", items[n_train:], hd)):
            for k in np.lexsort((its, ds)):
                rows[part].append((u, its[k], int(rng.integers(1, 6)), ds[k]))
    out = []
    for part in ("tr", "pr"):
        a = np.array(rows[part], np.int64)
        out.append(tx.Dataset.from_arrays(a[:, 0], a[:, 1], a[:, 2], a[:, 3], n_users, n_users * per_user))
    return out


@pytest.mark.parametrize("f,ordinal", [(24, False), (256, False), (60, True)])
def test_parallel_day_terms_match_serial_on_disjoint_users(gpu, f, ordinal):
    """Day terms (low_memory = false) in the parallel schedule against the
    serial oracle where users share no item: every table within 1e-3 of its
    movement over two epochs (a dropped or misplaced day term moves a table
    by a large part of that).  The ordinal case is the *o jobs' head with
    fixed thresholds (src/bin/tsvdx4-new.rs:187-192)."""
    tr, pr = _disjoint_users()
    head = tx.OrdinalHeadConfig(th_init=(0.5, 1.25, 3.75, 5.5), th_gap=0.001, lr_t=0.0, reg_t=0.0) if ordinal else None
    cfg = lm_config(n_feat=f, n_epochs=2, n_time_bins=30, low_memory=False, lr_ud=0.01, lr_yd=0.01,
                    sigma_yd=0.05, lr_iqf=0.005, lr_ibf=0.005, reset_u_epoch=100, ordinal_head=head)
    m_o, m_g = _oracle_and_gpu(cfg, tr, pr)   # default schedule: parallel, day terms in the kernel
    init = {n: m_o.get(n).copy() for n in ALL_PARAMS}
    for e in (1, 2):
        m_o.fit_epoch(e)
        m_g.fit_epoch(epoch=e)
    for n in ALL_PARAMS:
        o, g = m_o.get(n), m_g.get(n)
        move = float(np.max(np.abs(o.astype(np.float64) - init[n]))) if o.size else 0.0
        err = float(np.max(np.abs(g.astype(np.float64) - o))) if o.size else 0.0
        print(f"{n}: movement {move:.3e} gpu-oracle {err:.3e}")
        assert err <= 1e-3 * move + 1e-6, n
    assert m_g.calc_rmse(pr) == pytest.approx(m_o.calc_rmse(), abs=1e-5)


@pytest.mark.parametrize("f,low_memory", [(1500, True), (1500, False), (4500, True), (4500, False)])
def test_wide_rows_match_serial_on_disjoint_users(gpu, f, low_memory):
    """f > 1024: teams of 2 and 5 waves (1,024 elements of each row per wave,
    the dot product joined through LDS) against the serial oracle on users
    that share no item: every table within 1e-3 of its movement over two
    epochs, with and without day terms."""
    tr, pr = _disjoint_users()
    cfg = lm_config(n_feat=f, n_epochs=2, n_time_bins=30, low_memory=low_memory, lr_ud=0.01, lr_yd=0.01,
                    sigma_yd=0.05, lr_iqf=0.005, lr_ibf=0.005, lr_y=0.005, reset_u_epoch=100)
    m_o, m_g = _oracle_and_gpu(cfg, tr, pr)
    init = {n: m_o.get(n).copy() for n in ALL_PARAMS}
    for e in (1, 2):
        m_o.fit_epoch(e)
        m_g.fit_epoch(epoch=e)
    for n in ALL_PARAMS:
        o, g = m_o.get(n), m_g.get(n)
        move = float(np.max(np.abs(o.astype(np.float64) - init[n]))) if o.size else 0.0
        err = float(np.max(np.abs(g.astype(np.float64) - o))) if o.size else 0.0
        print(f"{n}: movement {move:.3e} gpu-oracle {err:.3e}")
        assert err <= 1e-3 * move + 1e-6, n
    assert m_g.calc_rmse(pr) == pytest.approx(m_o.calc_rmse(), abs=1e-5)


def test_parallel_day_terms_holdout_rmse(gpu):
    """Day terms in the parallel schedule train to the serial oracle's
    holdout RMSE (within 0.001) on the 10k x 2k x 1M shape."""
    tr, pr = synthetic("small")
    cfg = lm_config(n_feat=32, n_epochs=5, n_time_bins=30, low_memory=False)
    m_o, m_g = _oracle_and_gpu(cfg, tr, pr)
    for e in range(1, cfg.n_epochs + 1):
        m_o.fit_epoch(e)
        m_g.fit_epoch(epoch=e)
    r_o, r_g = m_o.calc_rmse(), m_g.calc_rmse(pr)
    print(f"day terms: oracle {r_o:.6f} gpu {r_g:.6f}")
    assert abs(r_g - r_o) <= 1e-3


def test_day_terms_schedule_rules(gpu):
    """Day terms train in one stratum, on date-ordered ratings; other inputs
    keep the serial schedule."""
    tr, pr = _disjoint_users(n_users=8)
    cfg = lm_config(n_feat=24, n_epochs=1, n_time_bins=30, low_memory=False)
    with tx.TxModel(tr, pr, cfg) as m:
        with pytest.raises(tx.TxError):
            m.set_schedule(mode=tx.SCHED_STRATA, n_strata=8)
        m.set_schedule(mode=tx.SCHED_STRATA, n_strata=1)
    # one user's ratings out of date order: the default is the serial schedule,
    # and the parallel one is refused
    order = np.arange(tr.n_ratings)
    k = int(np.flatnonzero(np.diff(tr.dates[:32].astype(np.int64)) > 0)[0])
    order[[k, k + 1]] = order[[k + 1, k]]
    tr2 = tx.Dataset.from_arrays(tr.user_idxs[order], tr.item_idxs[order], tr.raw_ratings[order], tr.dates[order],
                                 tr.n_users, tr.n_items)
    m_o = O.OracleModel(cfg, tr2, pr)
    with tx.TxModel(tr2, pr, cfg) as m:
        with pytest.raises(tx.TxError):
            m.set_schedule(mode=tx.SCHED_STRATA, n_strata=1)
        m.fit_epoch(epoch=1)
        m_o.fit_epoch(1)
        _assert_params_equal(m_o, m)    # the default ran the serial schedule: bit-identical


def test_full_shape_day_model_matches_serial_reference(gpu):
    """The tsvdx4-60 job's model (lm set with day terms, low_memory = false,
    f=60, 29 time bins, seed 42; src/bin/tsvdx4-new.rs:18) for 20 epochs in
    the default parallel schedule at the full Netflix shape, against the serial
    oracle's holdout RMSE (tests/golden/oracle_curve_full_job60.json): within
    1e-3 after the last epoch."""
    import json
    import os
    golden = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "oracle_curve_full_job60.json")))
    assert len(golden["epochs"]) == 20
    tr, pr = _full_data(golden["seed"])
    cfg = tx.TxConfig(n_feat=60, n_epochs=20, seed=42, n_time_bins=29, low_memory=False)
    with tx.TxModel(tr, pr, cfg) as m:
        curve = []
        for e in range(1, 21):
            m.fit_epoch(epoch=e)
            curve.append(m.calc_rmse(pr))
    ref = [x["holdout_rmse"] for x in golden["epochs"]]
    print("gpu", [round(x, 5) for x in curve])
    print("ref", [round(x, 5) for x in ref])
    assert abs(curve[-1] - ref[-1]) <= 1e-3


@pytest.mark.parametrize("low_memory", [True, False])
def test_single_team_matches_serial_oracle(gpu, low_memory):
    """The parallel kernel in serial-debug order (one team: blocks = 1,
    teams_per_block = 1, the reference's user order) against the serial
    oracle on the 600 x 150 x 30k set, lm set with q_f, p2 and (second case)
    day terms: every table within 1e-3 of its movement over two epochs (float
    association is the only difference left).  The same run with item rows
    written plainly instead of atomically and the epoch captured in a
    hipGraph gives bit-identical tables (the batched machinery changes no
    result in this order)."""
    tr, pr = synthetic("tiny_synth")
    cfg = lm_config(n_feat=24, n_epochs=2, n_time_bins=30, low_memory=low_memory, lr_iqf=0.005, lr_ibf=0.005,
                    reset_u_epoch=100)
    m_o = O.OracleModel(cfg, tr, pr)
    init = {n: m_o.get(n).copy() for n in ALL_PARAMS}
    runs = []
    for q_update, graph in ((0, False), (2, True)):
        m = tx.TxModel(tr, pr, cfg)
        m.set_schedule(mode=tx.SCHED_STRATA, blocks=1, teams_per_block=1, user_order=0, q_update=q_update,
                       use_graph=graph)
        for e in (1, 2):
            m.fit_epoch(epoch=e)
        runs.append({n: m.get(n) for n in ALL_PARAMS})
        m.close()
    for e in (1, 2):
        m_o.fit_epoch(e)
    for n in ALL_PARAMS:
        np.testing.assert_array_equal(runs[0][n], runs[1][n], err_msg=n)
        o, g = m_o.get(n), runs[0][n]
        move = float(np.max(np.abs(o.astype(np.float64) - init[n]))) if o.size else 0.0
        err = float(np.max(np.abs(g.astype(np.float64) - o))) if o.size else 0.0
        print(f"{n}: movement {move:.3e} gpu-oracle {err:.3e}")
        assert err <= 1e-3 * move + 1e-6, n


def test_plain_schedule_matches_serial_oracle(gpu):
    """TX_SCHED_PLAIN, the plain per-user wavefront kernel: with one wave
    (blocks = 1) it walks the users in the reference's order one after another,
    so every table after two epochs (lm set with q_f and p2) is the serial
    oracle's up to float association (within 1e-3 of each table's movement);
    with 1,024 waves the 10k x 2k x 1M set trains to within 1e-3 of the
    oracle's holdout RMSE."""
    tr, pr = synthetic("tiny_synth")
    cfg = lm_config(n_feat=24, n_epochs=2, n_time_bins=30, lr_iqf=0.005, lr_ibf=0.005, reset_u_epoch=100)
    m_o = O.OracleModel(cfg, tr, pr)
    init = {n: m_o.get(n).copy() for n in ALL_PARAMS}
    with tx.TxModel(tr, pr, cfg) as m:
        m.set_schedule(mode=tx.SCHED_PLAIN, blocks=1)
        for e in (1, 2):
            m.fit_epoch(epoch=e)
            m_o.fit_epoch(e)
        for n in ALL_PARAMS:
            o, g = m_o.get(n), m.get(n)
            move = float(np.max(np.abs(o.astype(np.float64) - init[n]))) if o.size else 0.0
            err = float(np.max(np.abs(g.astype(np.float64) - o))) if o.size else 0.0
            assert err <= 1e-3 * move + 1e-6, (n, err, move)
    tr, pr = synthetic("small")
    cfg = lm_config(n_feat=32, n_epochs=5, n_time_bins=30)
    m_o, m_g = _oracle_and_gpu(cfg, tr, pr, mode=tx.SCHED_PLAIN, blocks=1024)
    for e in range(1, cfg.n_epochs + 1):
        m_o.fit_epoch(e)
        m_g.fit_epoch(epoch=e)
    r_o, r_g = m_o.calc_rmse(), m_g.calc_rmse(pr)
    m_g.close()
    print(f"plain 1024 waves: oracle {r_o:.6f} gpu {r_g:.6f}")
    assert abs(r_g - r_o) <= 1e-3


def test_strata_counts(gpu):
    """n_strata is 1 or a multiple of 8 up to 64 (8 cells: one XCD each); the
    old experimental bit is still accepted."""
    tr, pr = synthetic("tiny_synth")
    with tx.TxModel(tr, pr, lm_config(n_feat=16, n_epochs=1)) as m:
        for bad in (2, 12, 72):
            with pytest.raises(tx.TxError, match="n_strata"):
                m.set_schedule(mode=tx.SCHED_STRATA, n_strata=bad)
        m.set_schedule(mode=tx.SCHED_STRATA, n_strata=8)
        m.set_schedule(mode=tx.SCHED_STRATA, n_strata=64, experimental=True)
        m.fit_epoch(epoch=1)


@pytest.mark.parametrize("n_strata", [8, 64])
def test_full_shape_stratified_matches_serial_reference(gpu, n_strata):
    """The block-stratified schedule at the full Netflix shape, plain form,
    f=256, 20 epochs: every user's ratings split over n_strata item blocks, cell
    c of stratum s = user block c x item block (c + s) mod n_strata on XCD c mod
    8, the NSVD1 step gathered per item after the last stratum (n_contrib^1
    normalisation); holdout RMSE within 1e-3 of the serial oracle's after the
    last epoch (tests/golden/oracle_curve_full_f256.json)."""
    import json
    import os
    golden = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "oracle_curve_full_f256.json")))
    tr, pr = _full_data(golden["seed"])
    cfg = tx.plain_form(lm_config(n_feat=256, n_epochs=20, n_time_bins=30, reset_u_epoch=10_000))
    with tx.TxModel(tr, pr, cfg) as m:
        m.set_schedule(mode=tx.SCHED_STRATA, n_strata=n_strata)
        curve, ms = [], []
        for e in range(1, 21):
            m.fit_epoch(epoch=e)
            ms.append(m.last_epoch_stats()["ms_total"])
            curve.append(m.calc_rmse(pr))
    ref = [x["holdout_rmse"] for x in golden["epochs"]]
    print(f"{n_strata} strata: {np.median(ms):.1f} ms/epoch")
    print("gpu", [round(x, 5) for x in curve])
    print("ref", [round(x, 5) for x in ref])
    assert abs(curve[-1] - ref[-1]) <= 1e-3
