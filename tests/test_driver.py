"""The job-level driver netflix_prize_amd/tsvdx4-new (drop-in for the
reference's `tsvdx4-new <job>`, src/bin/tsvdx4-new.rs): its job table on the
CPU, and one whole job end to end on the GPU."""
import os
import re
import subprocess

import numpy as np
import pytest

from netflix_prize_amd import synth, tx
from tests.fixtures import synthetic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "netflix_prize_amd", "tsvdx4-new")


def _config(job):
    out = subprocess.run([DRIVER, job, "--print-config"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    return dict(re.findall(r"(\w+): ([^,{}]+)", out.stdout))


def test_print_config_matches_reference_job_table():
    # run_lm(400, 20, 42, 29, save_ifeat=true, low_memory=true, ...) with the lm set
    # (src/bin/tsvdx4-new.rs:19, 130-154)
    c = _config("tsvdx4-400lm")
    expect = {"n_feat": "400", "n_epochs": "20", "seed": "42", "n_time_bins": "29", "save_ifeat": "true",
              "low_memory": "true", "full_su": "true", "lr_u": "0.003", "lr_ub": "0.0031", "lr_i": "0.0036",
              "lr_y": "0.0005", "lr_yb": "2.5e-5", "lr_ti": "0.000225", "reg_u": "0.0504", "reg_i": "0.00735",
              "reg_y": "0.04", "sigma_u": "0.0015", "reset_u_epoch": "13", "ordinal_head": "None"}
    for k, v in expect.items():
        assert c[k].strip() == v, (k, c[k])
    c = _config("tsvdx4-60")           # run_lm(60, 20, 42, 29, false, false, ...) (:18)
    assert (c["n_feat"], c["low_memory"], c["save_ifeat"]) == ("60", "false", "false")


def test_print_config_tsvdx4p_job():
    """tsvdx4p-new jobs (src/bin/tsvdx4p-new.rs:14, 30-85) from the same binary
    under the reference's name: the Tsvdx4pConfig of tsvdx4p-1000lm."""
    out = subprocess.run([os.path.join(ROOT, "netflix_prize_amd", "tsvdx4p-new"), "tsvdx4p-1000lm", "--print-config"],
                         capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert out.stdout.startswith("Tsvdx4pConfig { n_feat: 1000, n_epochs: 20, seed: 42,")
    c = dict(re.findall(r"(\w+): ([^,{}\[\]]+)", out.stdout))
    expect = {"lr_yb": "0.0002", "lr_iqf": "5e-6", "sigma_iqf": "0.005", "reset_u_epoch": "1024",
              "save_ifeat": "true", "low_memory": "true", "nsvd_norm_exp": "1.0", "n_threads": "16"}
    for k, v in expect.items():
        assert c[k].strip() == v, (k, c[k])
    assert "seq_epochs: [1, 10, 20, 0, 0] }" in out.stdout


# Every job name of the reference's dispatchers: the match arms of
# src/bin/tsvdx4-new.rs:18-111 and src/bin/tsvdx4p-new.rs:14-24.
REFERENCE_JOBS = (
    ["tsvdx4-60", "tsvdx4-400lm", "tsvdx4-1600lm", "tsvdx4-2400lm", "tsvdx4-4800lm", "tsvdx4-60o", "tsvdx4-80o",
     "tsvdx4-160o", "tsvdx4-d0tb", "tsvdx4-f1", "tsvdx4-h8nb", "tsvdx4-l32nb", "tsvdx4-m1n", "tsvdx4-q16tb",
     "tsvdx4-s20fo", "tsvdx4-t8cu", "tsvdx4-16-t", "tsvdx4-n8nbo", "tsvdx4-p0tbo",
     "tsvdx4-d0tb__knn3", "tsvdx4-h8nb__knn3", "tsvdx4-n8nbo__knn3", "tsvdx4-p0tbo__knn3",
     "tsvdx4-60__knns", "tsvdx4-60o__knns", "tsvdx4-2400lm__knns", "tsvdx4-4800lm__knns",
     "tsvdx4-16-t__knnf", "tsvdx4-400lm__knnf", "tsvdx4-4800lm__knnf", "tsvdx4-4800lm__knn-d",
     "tsvdx4-60__epochs", "tsvdx4-60o__epochs", "tsvdx4-d0tb__epochs", "tsvdx4-f1__epochs",
     "tsvdx4-q16tb__epochs", "tsvdx4-400lm__epochs"],
    ["tsvdx4p-1000lm", "tsvdx4p-2000lm", "tsvdx4p-4000lm", "tsvdx4p-6000lm", "tsvdx4p-8000lm",
     "tsvdx4p-16000lm", "tsvdx4p-1000lm__knnf", "tsvdx4p-2000lm__knnf", "tsvdx4p-6000lm__knnf",
     "tsvdx4p-8000lm__knnf"],
    # src/bin/tsvdx5-new.rs:23-35 (TxModel base and ordinal jobs), :84-89 (knn3 on
    # the ordinal base), :104-128 (epoch blends); the NLPP, paropt, knnf and
    # knn-w chains of that binary belong to other model families
    ["tsvdx5-20", "tsvdx5-40", "tsvdx5-60", "tsvdx5-120", "tsvdx5-150", "tsvdx5-300", "tsvdx5-400", "tsvdx5-1000",
     "tsvdx5-1200lm", "tsvdx5-120o", "tsvdx5-150o", "tsvdx5-120o__knn3", "tsvdx5-120o__epochs",
     "tsvdx5-300__epochs"],
)


@pytest.mark.parametrize("binary,job", [("tsvdx4-new", j) for j in REFERENCE_JOBS[0]] +
                         [("tsvdx4p-new", j) for j in REFERENCE_JOBS[1]] +
                         [("tsvdx5-new", j) for j in REFERENCE_JOBS[2]])
def test_every_reference_job_name_is_in_the_table(binary, job):
    out = subprocess.run([os.path.join(ROOT, "netflix_prize_amd", binary), job, "--print-config"],
                         capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, (job, out.stderr)


def test_print_config_tsvdx5_and_tsvdx6_jobs():
    """run_base(300, 23, 100, 0.0015, false, 10, false, false, true)
    (src/bin/tsvdx5-new.rs:28, 130-161), run_ord(150, 25, 120, 1.2)
    (:33, 185-217), and the two tsvdx6 binaries, which take no job argument
    (src/bin/tsvdx6-300-new.rs:3-61, src/bin/tsvdx6-300-v2-new.rs:45-48)."""
    c = _config("tsvdx5-300")
    expect = {"n_feat": "300", "n_epochs": "23", "seed": "42", "n_time_bins": "32", "reset_u_epoch": "10",
              "max_neighbors": "100", "lr_w": "0.0015", "lr_c": "0.0015", "reg_w": "0.002", "reg_c": "0.002",
              "lr_w_day": "0.0", "w_bias": "0.8", "w_factor": "0.8", "w_nbr": "0.8", "lambda1": "25.0",
              "lambda2": "10.0", "low_memory": "false", "save_ifeat": "false", "lr_tu": "0.0", "sigma_u": "0.0015"}
    for k, v in expect.items():
        assert c[k].strip() == v, (k, c[k])
    c = _config("tsvdx5-150o")
    expect = {"n_feat": "150", "n_epochs": "25", "seed": "64", "n_time_bins": "20", "max_neighbors": "120",
              "w_bias": "1.2", "w_factor": "1.2", "w_nbr": "1.2", "lr_i": "0.0072", "sigma_u": "0.0",
              "ordinal_head": "Some(OrdinalHeadConfig"}
    for k, v in expect.items():
        assert c[k].strip().startswith(v), (k, c[k])
    for binary, w_day, w_nbr in (("tsvdx6-300-new", "0.001", "0.8"), ("tsvdx6-300-v2-new", "0.0008", "0.7")):
        out = subprocess.run([os.path.join(ROOT, "netflix_prize_amd", binary), "--print-config"],
                             capture_output=True, text=True, timeout=60)
        assert out.returncode == 0, out.stderr
        c = dict(re.findall(r"(\w+): ([^,{}]+)", out.stdout))
        assert (c["n_feat"], c["lr_w_day"], c["lr_c_day"], c["w_nbr"], c["max_neighbors"]) == ("300", w_day, w_day,
                                                                                                 w_nbr, "100")


def test_unknown_job_is_an_error():
    out = subprocess.run([DRIVER, "tsvdx4-no-such-job"], capture_output=True, text=True, timeout=60)
    assert out.returncode != 0
    assert "tsvdx4-no-such-job" in (out.stdout + out.stderr)


def test_epoch_blend_job(tmp_path):
    """tsvdx4-d0tb__epochs (run_eblend with the reference's weights and offset,
    src/bin/tsvdx4-new.rs:85-88; epoch_blend_apply, src/lib.rs:1181-1235): the
    FP64 weighted sum of the base job's per-epoch predictions plus the offset,
    stored as FP32, for the probe and qual sets, and its log."""
    w = [0.127633, 0.128381, 0.128978, 0.129290, 0.129347]
    off = 0.978328
    pd = tmp_path / "preds_new"
    pd.mkdir()
    rng = np.random.default_rng(3)
    eps = {}
    for ds in ("probex", "qual"):
        eps[ds] = [rng.uniform(1, 5, 1000).astype(np.float32) for _ in w]
        for e, a in enumerate(eps[ds], 1):
            np.save(pd / f"tsvdx4-d0tb_ep{e:02d}.{ds}.npy", a)
    out = subprocess.run([DRIVER, "tsvdx4-d0tb__epochs"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    for ds in ("probex", "qual"):
        want = np.full(1000, off, np.float64)
        for wi, a in zip(w, eps[ds]):
            want = want + wi * a.astype(np.float64)
        got = np.load(pd / f"tsvdx4-d0tb__epochs.{ds}.npy")
        assert got.dtype == np.float32
        np.testing.assert_array_equal(got, want.astype(np.float32))
    log = (pd / "tsvdx4-d0tb__epochs.out").read_text().splitlines()
    assert log[:3] == ["[tsvdx4-d0tb__epochs]", "eblend legacy mode: 5 epoch weights, constant offset=+0.978328",
                       "  epoch  1 weight = +0.127633"]
    # a missing epoch file is an error, as the reference's expect(path)
    (pd / "tsvdx4-d0tb_ep03.qual.npy").unlink()
    out = subprocess.run([DRIVER, "tsvdx4-d0tb__epochs"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert out.returncode != 0


def _write_split(root, tr, pr):
    """The reference's SPLIT_NEW sets (trainx, probex, fulltrain, qual)."""
    full_u = np.concatenate([tr.user_idxs, pr.user_idxs])
    order = np.lexsort((np.concatenate([tr.dates, pr.dates]), full_u))
    cat = lambda a, b: np.concatenate([a, b])[order]
    full = tx.Dataset.from_arrays(full_u[order], cat(tr.item_idxs, pr.item_idxs), cat(tr.raw_ratings, pr.raw_ratings),
                                  cat(tr.dates, pr.dates), tr.n_users, tr.n_items)
    for name, ds in (("trainx", tr), ("probex", pr), ("fulltrain", full), ("qual", pr)):
        synth.write_npy_dir(ds, os.path.join(root, "data", name))


@pytest.mark.gpu
def test_job_end_to_end(gpu, tmp_path):
    """tsvdx4-400lm (low-memory, parallel schedule) on a 10k x 2k x 1M
    synthetic split: every output file of the job, with the sizes of its sets,
    and the logged probe RMSE equal to the RMSE of the saved predictions."""
    tr, pr = synthetic("small")
    _write_split(str(tmp_path), tr, pr)
    env = dict(os.environ, TSVD_N_USERS=str(tr.n_users), TSVD_N_ITEMS=str(tr.n_items),
               TSVD_DATA_DIR=str(tmp_path / "data"))
    out = subprocess.run([DRIVER, "tsvdx4-400lm"], cwd=tmp_path, env=env, capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    pd = tmp_path / "preds_new"
    job = "tsvdx4-400lm"
    probe = np.load(pd / f"{job}.probex.npy")
    assert probe.shape == (pr.n_ratings,) and np.all(np.isfinite(probe))
    assert np.load(pd / f"{job}.trainx.npy").shape == (tr.n_ratings,)
    assert np.load(pd / f"{job}.qual.npy").shape == (pr.n_ratings,)
    for part in ("bias", "mf", "nsvd1"):
        assert np.load(pd / f"{job}-{part}.probex.npy").shape == (pr.n_ratings,)
    assert np.load(pd / f"{job}_ep20.probex.npy").shape == (pr.n_ratings,)
    assert np.load(pd / f"{job}.ifeat.trainx.npy").shape == (tr.n_items, 400)
    log = (pd / f"{job}.out").read_text()
    logged = float(re.search(r"probex RMSE ([0-9.]+)", log).group(1))
    rmse = float(np.sqrt(np.mean((probe.astype(np.float64) - pr.raw_ratings) ** 2)))
    assert abs(logged - rmse) < 1e-5, (logged, rmse)
    # better than the best constant prediction of the held-out ratings
    assert rmse < float(np.std(pr.raw_ratings.astype(np.float64)))

    # the job's kNN chain over its saved item factors (run_knnf, src/bin/tsvdx4-new.rs:576-588):
    # predictions = base job's + KnnfModel on the base job's residuals, equal to the oracle's
    knn_job = job + "__knnf"
    out = subprocess.run([DRIVER, knn_job], cwd=tmp_path, env=env, capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    kp = np.load(pd / f"{knn_job}.probex.npy")
    assert kp.shape == (pr.n_ratings,) and np.all(np.isfinite(kp))
    assert np.load(pd / f"{knn_job}.qual.npy").shape == (pr.n_ratings,)
    klog = (pd / f"{knn_job}.out").read_text()
    assert f'target = "1.0*{job}"' in klog and "KnnfConfig { factors:" in klog
    klogged = float(re.search(r"probex RMSE ([0-9.]+)", klog).group(1))
    krmse = float(np.sqrt(np.mean((kp.astype(np.float64) - pr.raw_ratings) ** 2)))
    assert abs(klogged - krmse) < 1e-5, (klogged, krmse)
    from oracle.knnf_ref import KnnfRef
    res_tr = tr.raw_ratings.astype(np.float32) - np.load(pd / f"{job}.trainx.npy")
    tr_k = tx.Dataset.from_arrays(tr.user_idxs, tr.item_idxs, tr.raw_ratings, tr.dates, tr.n_users, tr.n_items,
                                  residuals=res_tr)
    ref = KnnfRef(tr_k, np.load(pd / f"{job}.ifeat.trainx.npy"))
    q = np.arange(0, pr.n_ratings, 50)
    want = ref.predict(pr.user_idxs[q], pr.item_idxs[q], pr.dates[q]) + probe[q]
    clear = ref.margins(pr.user_idxs[q], pr.item_idxs[q], pr.dates[q]) > 1e-4
    assert np.max(np.abs(kp[q][clear] - want[clear])) < 1e-4


@pytest.mark.gpu
def test_tsvdx4p_job_end_to_end(gpu, tmp_path):
    """tsvdx4p-1000lm (Tsvdx4pModel epochs with the q_f stream, f=1000,
    seq_epochs 1/10/20) on the 10k x 2k x 1M split: its output files and the
    logged probe RMSE equal to the RMSE of the saved predictions."""
    tr, pr = synthetic("small")
    _write_split(str(tmp_path), tr, pr)
    env = dict(os.environ, TSVD_N_USERS=str(tr.n_users), TSVD_N_ITEMS=str(tr.n_items),
               TSVD_DATA_DIR=str(tmp_path / "data"))
    job = "tsvdx4p-1000lm"
    out = subprocess.run([DRIVER, job], cwd=tmp_path, env=env, capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    pd = tmp_path / "preds_new"
    probe = np.load(pd / f"{job}.probex.npy")
    assert probe.shape == (pr.n_ratings,) and np.all(np.isfinite(probe))
    assert np.load(pd / f"{job}.trainx.npy").shape == (tr.n_ratings,)
    assert np.load(pd / f"{job}.qual.npy").shape == (pr.n_ratings,)
    assert np.load(pd / f"{job}.ifeat.trainx.npy").shape == (tr.n_items, 1000)
    log = (pd / f"{job}.out").read_text()
    assert log.splitlines()[2].startswith("Tsvdx4pConfig {")
    logged = float(re.search(r"probex RMSE ([0-9.]+)", log).group(1))
    rmse = float(np.sqrt(np.mean((probe.astype(np.float64) - pr.raw_ratings) ** 2)))
    assert abs(logged - rmse) < 1e-5, (logged, rmse)
    assert rmse < float(np.std(pr.raw_ratings.astype(np.float64)))


@pytest.mark.gpu
def test_knns_job_end_to_end(gpu, tmp_path):
    """tsvdx4-60 (day terms, parallel schedule) then its kNN chain
    tsvdx4-60__knns (KnnsModel, support computed from each phase's training
    and held-out sets; run_knns, src/bin/tsvdx4-new.rs:49, 563-574) on the
    10k x 2k x 1M split: output files, logged RMSE, and predictions equal to
    the oracle's on the base job's residuals."""
    tr, pr = synthetic("small")
    _write_split(str(tmp_path), tr, pr)
    env = dict(os.environ, TSVD_N_USERS=str(tr.n_users), TSVD_N_ITEMS=str(tr.n_items),
               TSVD_DATA_DIR=str(tmp_path / "data"))
    for job in ("tsvdx4-60", "tsvdx4-60__knns"):
        out = subprocess.run([DRIVER, job], cwd=tmp_path, env=env, capture_output=True, text=True, timeout=600)
        assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    pd = tmp_path / "preds_new"
    base = np.load(pd / "tsvdx4-60.probex.npy")
    kp = np.load(pd / "tsvdx4-60__knns.probex.npy")
    assert kp.shape == (pr.n_ratings,) and np.all(np.isfinite(kp))
    assert np.load(pd / "tsvdx4-60__knns.qual.npy").shape == (pr.n_ratings,)
    klog = (pd / "tsvdx4-60__knns.out").read_text()
    assert "KnnsConfig { k: 25, shrinkage: 100.0, scaling: 1.0, tau: 0.025, supp_source: Compute }" in klog
    logged = float(re.search(r"probex RMSE ([0-9.]+)", klog).group(1))
    rmse = float(np.sqrt(np.mean((kp.astype(np.float64) - pr.raw_ratings) ** 2)))
    assert abs(logged - rmse) < 1e-5, (logged, rmse)
    from oracle.knnf_ref import KnnsRef
    res_tr = tr.raw_ratings.astype(np.float32) - np.load(pd / "tsvdx4-60.trainx.npy")
    tr_k = tx.Dataset.from_arrays(tr.user_idxs, tr.item_idxs, tr.raw_ratings, tr.dates, tr.n_users, tr.n_items,
                                  residuals=res_tr)
    ref = KnnsRef(tr_k, pr, tr.n_items, k=25, shrinkage=100.0, scaling=1.0, tau=0.025)
    q = np.arange(0, pr.n_ratings, 50)
    want = ref.predict(pr.user_idxs[q], pr.item_idxs[q], pr.dates[q]) + base[q]
    clear = ref.margins(pr.user_idxs[q], pr.item_idxs[q], pr.dates[q]) > 1e-4
    assert np.max(np.abs(kp[q][clear] - want[clear])) < 1e-4


@pytest.mark.gpu
def test_knn3_job_end_to_end(gpu, tmp_path):
    """tsvdx4-d0tb (bias-only timeSVD++) then its kNN chain tsvdx4-d0tb__knn3
    (Knn3Model with Knn3Config::default on the base job's residuals; run_knn3,
    src/bin/tsvdx4-new.rs:43-46, 557-561) on the 10k x 2k x 1M split: output
    files, the logged config and RMSE, and predictions equal to the oracle's."""
    tr, pr = synthetic("small")
    _write_split(str(tmp_path), tr, pr)
    env = dict(os.environ, TSVD_N_USERS=str(tr.n_users), TSVD_N_ITEMS=str(tr.n_items),
               TSVD_DATA_DIR=str(tmp_path / "data"))
    for job in ("tsvdx4-d0tb", "tsvdx4-d0tb__knn3"):
        out = subprocess.run([DRIVER, job], cwd=tmp_path, env=env, capture_output=True, text=True, timeout=600)
        assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    pd = tmp_path / "preds_new"
    base = np.load(pd / "tsvdx4-d0tb.probex.npy")
    kp = np.load(pd / "tsvdx4-d0tb__knn3.probex.npy")
    assert kp.shape == (pr.n_ratings,) and np.all(np.isfinite(kp))
    assert np.load(pd / "tsvdx4-d0tb__knn3.qual.npy").shape == (pr.n_ratings,)
    klog = (pd / "tsvdx4-d0tb__knn3.out").read_text()
    assert ("Knn3Config { threshold: 0.25, k_min: 10, k_max: 60, shrinkage: 20000.0, reg: 0.01, x: 0.8, "
            "bl_reg_m: 25.0, bl_reg_u: 10.0 }") in klog
    assert "  KNN3: computing per-item weights (2000 items, parallel)..." in klog
    logged = float(re.search(r"probex RMSE ([0-9.]+)", klog).group(1))
    rmse = float(np.sqrt(np.mean((kp.astype(np.float64) - pr.raw_ratings) ** 2)))
    assert abs(logged - rmse) < 1e-5, (logged, rmse)
    from oracle.knn3_ref import Knn3Ref
    res_tr = tr.raw_ratings.astype(np.float32) - np.load(pd / "tsvdx4-d0tb.trainx.npy")
    tr_k = tx.Dataset.from_arrays(tr.user_idxs, tr.item_idxs, tr.raw_ratings, tr.dates, tr.n_users, tr.n_items,
                                  residuals=res_tr)
    ref = Knn3Ref(tr_k)
    q = np.arange(0, pr.n_ratings, 50)
    want = ref.predict(pr.user_idxs[q], pr.item_idxs[q]) + base[q]
    assert np.max(np.abs(kp[q] - want)) < 1e-4


def _job_config(job):
    """The job's TxConfig as the driver prints it (--print-config)."""
    out = subprocess.run([DRIVER, job, "--print-config"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    d = {}
    for k, v in re.findall(r"(\w+): ([^,{}\[\]]+)", out.stdout):
        v = v.strip()
        if v in ("true", "false"):
            d[k] = v == "true"
        elif v != "None":
            d[k] = int(v) if re.fullmatch(r"-?\d+", v) else float(v)
    return tx.TxConfig.from_dict(d)


@pytest.mark.gpu
def test_transposed_job_end_to_end(gpu, tmp_path):
    """tsvdx4-16-t (Dataset::transpose, src/lib.rs:295-402: items act as users,
    training rows re-sorted by (item, date), held-out rows kept in their
    order) on the 10k x 2k x 1M split: the transposed log line, the saved
    item factors having one row per original user, probe predictions in the
    original row order, and the holdout RMSE of a TxModel trained through the
    Python API on the same transposition within 2e-3 (both parallel)."""
    tr, pr = synthetic("small")
    _write_split(str(tmp_path), tr, pr)
    env = dict(os.environ, TSVD_N_USERS=str(tr.n_users), TSVD_N_ITEMS=str(tr.n_items),
               TSVD_DATA_DIR=str(tmp_path / "data"))
    job = "tsvdx4-16-t"
    out = subprocess.run([DRIVER, job], cwd=tmp_path, env=env, capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    pd = tmp_path / "preds_new"
    log = (pd / f"{job}.out").read_text()
    assert "trainx => probex (transposed)" in log
    assert np.load(pd / f"{job}.ifeat.trainx.npy").shape == (tr.n_users, 16)
    probe = np.load(pd / f"{job}.probex.npy")
    assert probe.shape == (pr.n_ratings,) and np.all(np.isfinite(probe))
    logged = float(re.search(r"probex RMSE ([0-9.]+)", log).group(1))
    rmse = float(np.sqrt(np.mean((probe.astype(np.float64) - pr.raw_ratings) ** 2)))
    assert abs(logged - rmse) < 1e-5, (logged, rmse)
    # the same model through the Python API on a numpy transposition
    order = np.lexsort((tr.dates, tr.item_idxs))
    tt = tx.Dataset.from_arrays(tr.item_idxs[order], tr.user_idxs[order], tr.raw_ratings[order], tr.dates[order],
                                tr.n_items, tr.n_users)
    pt = tx.Dataset.from_arrays(pr.item_idxs, pr.user_idxs, pr.raw_ratings, pr.dates, tr.n_items, tr.n_users)
    cfg = _job_config(job)
    m = tx.TxModel(tt, pt, cfg)
    for e in range(1, cfg.n_epochs + 1):
        m.fit_epoch(epoch=e)
    mine = m.predict(pt.user_idxs, pt.item_idxs, pt.dates)
    m.close()
    r_mine = float(np.sqrt(np.mean((mine.astype(np.float64) - pr.raw_ratings) ** 2)))
    print(f"transposed job RMSE {rmse:.5f}, Python API {r_mine:.5f}")
    assert abs(r_mine - rmse) < 2e-3
    assert rmse < float(np.std(pr.raw_ratings.astype(np.float64)))


def reference_transpose(ds, reorder):
    """Dataset::transpose (src/lib.rs:295-402): users and items swap roles;
    with reorder, rows are sorted by the key item * n_days + date with two
    stable 16-bit radix passes, i.e. a stable sort by (item, date)."""
    order = np.arange(ds.n_ratings)
    if reorder:
        n_days = int(ds.dates.max()) + 1
        key = ds.item_idxs.astype(np.int64) * n_days + ds.dates.astype(np.int64)
        order = np.argsort(key, kind="stable")
    return tx.Dataset.from_arrays(ds.item_idxs[order], ds.user_idxs[order], ds.raw_ratings[order],
                                  ds.dates[order], ds.n_items, ds.n_users, is_test=ds.is_test[order])


@pytest.mark.gpu
def test_transposed_job_serial_matches_oracle(gpu, tmp_path):
    """tsvdx4-16-t in the serial schedule (TSVD_SCHEDULE=serial) against the
    oracle trained on the reference's transposition of the same split: the
    saved item factors (one row per original user) bit for bit, and the probe
    predictions in their original order within the predict tolerance 1e-5."""
    from oracle import oracle as O
    tr, pr = synthetic("tiny_synth")
    _write_split(str(tmp_path), tr, pr)
    env = dict(os.environ, TSVD_N_USERS=str(tr.n_users), TSVD_N_ITEMS=str(tr.n_items),
               TSVD_DATA_DIR=str(tmp_path / "data"), TSVD_SCHEDULE="serial")
    job = "tsvdx4-16-t"
    out = subprocess.run([DRIVER, job], cwd=tmp_path, env=env, capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    tt, pt = reference_transpose(tr, True), reference_transpose(pr, False)
    cfg = _job_config(job)
    o = O.OracleModel(cfg, tt, pt)
    for e in range(1, cfg.n_epochs + 1):
        o.fit_epoch(e)
    pd = tmp_path / "preds_new"
    ifeat = np.load(pd / f"{job}.ifeat.trainx.npy")
    np.testing.assert_array_equal(ifeat, o.get("ifeat").reshape(tr.n_users, cfg.n_feat))
    probe = np.load(pd / f"{job}.probex.npy")
    ref = o.predict(pt.user_idxs, pt.item_idxs, pt.dates)   # target rtg: no offset
    assert np.max(np.abs(probe - ref)) <= 1e-5


@pytest.mark.gpu
def test_knn_d_job_end_to_end(gpu, tmp_path):
    """tsvdx4-4800lm__knn-d (run_knn_d, src/bin/tsvdx4-new.rs:590-610: KnnModel
    over sim/rtg_supp.{set}.npy and sim/tsvdx4-4800lm_diff1.{set}.npy, target
    1.0*tsvdx4-4800lm) on the 10k x 2k x 1M split, with seeded base predictions
    and stat tables written beforehand: output files, the logged config and
    RMSE, and predictions equal to the oracle's."""
    from oracle.knnf_ref import KnnStatRef
    tr, pr = synthetic("small")
    _write_split(str(tmp_path), tr, pr)
    env = dict(os.environ, TSVD_N_USERS=str(tr.n_users), TSVD_N_ITEMS=str(tr.n_items),
               TSVD_DATA_DIR=str(tmp_path / "data"))
    pd, sd = tmp_path / "preds_new", tmp_path / "sim"
    pd.mkdir()
    sd.mkdir()
    rng = np.random.default_rng(11)
    base = {}
    for name, n in (("trainx", tr.n_ratings), ("probex", pr.n_ratings), ("fulltrain", tr.n_ratings + pr.n_ratings),
                    ("qual", pr.n_ratings)):
        base[name] = rng.uniform(2.5, 4.5, n).astype(np.float32)
        np.save(pd / f"tsvdx4-4800lm.{name}.npy", base[name])
    full_u = np.concatenate([tr.user_idxs, pr.user_idxs])
    full_i = np.concatenate([tr.item_idxs, pr.item_idxs])
    stats = {}
    for name, (u, i) in (("trainx", (tr.user_idxs, tr.item_idxs)), ("fulltrain", (full_u, full_i))):
        B = np.zeros((tr.n_users, tr.n_items), np.float32)
        B[u, i] = 1.0
        stats[name] = ((B.T @ B).astype(np.float32), rng.normal(0, 50, (tr.n_items, tr.n_items)).astype(np.float32))
        np.save(sd / f"rtg_supp.{name}.npy", stats[name][0])
        np.save(sd / f"tsvdx4-4800lm_diff1.{name}.npy", stats[name][1])
    job = "tsvdx4-4800lm__knn-d"
    out = subprocess.run([DRIVER, job], cwd=tmp_path, env=env, capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    kp = np.load(pd / f"{job}.probex.npy")
    assert kp.shape == (pr.n_ratings,) and np.all(np.isfinite(kp))
    assert np.load(pd / f"{job}.qual.npy").shape == (pr.n_ratings,)
    klog = (pd / f"{job}.out").read_text()
    assert ('KnnConfig { stat0: Some("rtg_supp"), stat1: Some("tsvdx4-4800lm_diff1"), stat2: None, factors: None, '
            'sim_type: Support, k: 15, sim_threshold: None, shrinkage: 100.0, scaling: 3.5, tau: 0.0, use_stat1: true, '
            'regression: false, regression_lambda: 0.1, sim_dir: "sim", preds_dir: "preds_new" }') in klog
    logged = float(re.search(r"probex RMSE ([0-9.]+)", klog).group(1))
    rmse = float(np.sqrt(np.mean((kp.astype(np.float64) - pr.raw_ratings) ** 2)))
    assert abs(logged - rmse) < 1e-5, (logged, rmse)
    res_tr = tr.raw_ratings.astype(np.float32) - base["trainx"]
    tr_k = tx.Dataset.from_arrays(tr.user_idxs, tr.item_idxs, tr.raw_ratings, tr.dates, tr.n_users, tr.n_items,
                                  residuals=res_tr)
    ref = KnnStatRef(tr_k, *stats["trainx"], k=15, shrinkage=100.0, scaling=3.5, tau=0.0)
    q = np.arange(0, pr.n_ratings, 50)
    want = ref.predict(pr.user_idxs[q], pr.item_idxs[q], pr.dates[q]) + base["probex"][q]
    clear = ref.margins(pr.user_idxs[q], pr.item_idxs[q], pr.dates[q]) > 1e-4
    assert np.max(np.abs(kp[q][clear] - want[clear])) < 1e-4
    # a missing stat file is an error, as the reference's unwrap
    (sd / "rtg_supp.fulltrain.npy").unlink()
    out = subprocess.run([DRIVER, job], cwd=tmp_path, env=env, capture_output=True, text=True, timeout=600)
    assert out.returncode != 0
