"""Benchmark: timeSVD++ (tsvdx4) SGD ratings/s per epoch on one MI355X.

Workload (default): the full synthetic shape — 480,189 users x 17,770 items,
100,480,507 ratings (~99.0M train, ~1.44M held out) — with the plain
timeSVD++ model at f=256, FP32, 30 time bins, low-memory training.  One step
is one training epoch over every training rating (SGD kernels plus the
epoch-end ycache rebuild) with all inputs resident in HBM.

Prints one JSON line; see the driver contract in the repository docs.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_per_rating(f: int, n_train: int, n_users: int, qf: bool, p2: bool, s: int = 4) -> dict:
    """Algorithmic HBM bytes per training rating of the SGD kernel and the
    rebuild kernel (minimal traffic of the reference algorithm):
      SGD:     q_i read+write, y_j gather read + update write, 5-byte rating
               record, [qf row read+write], per-user p (and p2) read+write;
      rebuild: one y_j read per train and held-out row, one ycache write per user.
    """
    upr = n_users / n_train
    sgd = f * (2 * s + 2 * s) + 5 + (2 * s * f if qf else 0) + (2 * s * f * upr if p2 else 0) + 2 * s * f * upr
    return {"sgd": sgd, "rebuild_per_row": s * f, "ycache_per_user": s * f}


def day_bytes_per_rating(f: int, n_train: int, n_heldout: int, n_days: int, s: int = 4) -> float:
    """Algorithmic HBM bytes per training rating that the day terms add to the
    SGD kernel: y_day gather of every training and (full_su) held-out row,
    y_day read and write per training row, and per training day p_day read and
    write, the day sum written and read back, the run gradient written and
    read back; plus the run norm read per rating."""
    return 3 * s * f + s * f * n_heldout / n_train + 6 * s * f * n_days / n_train + s


def nbr_bytes_per_rating(tr, pr, max_nbr: int, same_day: bool, s: int = 4) -> float:
    """Algorithmic HBM bytes per training rating that the neighbourhood terms
    add to the SGD kernel: c(i, j) read and write for every global neighbour
    j != i of the rating's user (the epoch's sample of max_nbr, or all of
    R(u) and N(u)\\R(u)), w(i, j) read and write for the rated ones; with the
    same-day terms, per same-day neighbour its list entry (item, r - b~), c_day
    read and write, and w_day read and write for the rated ones."""
    nr = np.bincount(tr.user_idxs, minlength=tr.n_users).astype(np.float64)
    ni = np.bincount(pr.user_idxs, minlength=tr.n_users).astype(np.float64)  # held-out items are unrated
    n = nr + ni
    samp = (n > max_nbr) & (nr > 0)
    rb = np.where(samp, np.clip(np.round(max_nbr * nr / np.maximum(n, 1)), 1, np.maximum(nr, 1)), nr)
    ib = np.where(samp, np.minimum(max_nbr - rb, ni), ni)
    # a rating's own item is among its rated neighbours with probability rb / nr
    rated = np.where(nr > 0, rb - rb / np.maximum(nr, 1), 0.0)
    total = float(np.sum(nr * (4 * s * rated + 2 * s * ib)))
    if same_day:
        # training rows are sorted by (user, date): each run of equal keys is one day
        key = tr.user_idxs.astype(np.int64) * 65536 + (tr.dates.astype(np.int64) + 32768)
        first = np.r_[True, key[1:] != key[:-1]]
        rid = np.cumsum(first) - 1
        run = np.bincount(rid).astype(np.float64)
        hkeys, hcnt = np.unique(pr.user_idxs.astype(np.int64) * 65536 + (pr.dates.astype(np.int64) + 32768),
                                return_counts=True)
        rk = key[first]
        at = np.minimum(np.searchsorted(hkeys, rk), max(len(hkeys) - 1, 0))
        imp = np.where((len(hkeys) > 0) & (hkeys[at] == rk), hcnt[at], 0).astype(np.float64) if len(hkeys) else \
            np.zeros(len(rk))
        # per rating: (run - 1) rated same-day neighbours, imp implicit ones
        total += float(np.sum(run * ((run - 1) * (2 * s + 4 * s) + imp * (2 * s + 2 * s))))
    return total / len(tr.user_idxs)


def profiled_traffic(f: int, args, kind: str = "") -> float | None:
    """HBM bytes per SGD launch (FETCH_SIZE + WRITE_SIZE, calibrated against a
    plain copy) from the rocprofv3 summary of this workload in profiles/, or
    None for a workload that has no such summary."""
    path = os.path.join(ROOT, "profiles", f"k_sgd_par_{kind}{args.shape}_f{f}.md")
    if not os.path.exists(path) or args.strata != 1:
        return None
    text = open(path).read()
    if "```json" not in text:
        return None
    block = text.split("```json", 1)[1].split("```", 1)[0]
    return float(json.loads(block)["traffic_bytes_per_launch"])


_ATOMIC_CEILING = None


def atomic_ceiling() -> dict | None:
    """The device's rate for the SGD kernel's item-row update pattern
    (netflix_prize_amd/atomic-ceiling: 1 KB rows of a 17,770-row table updated
    by 64-lane FP32 atomic adds, by 64-bit integer atomics and by plain
    load + store; a 2-lane FP32 atomic add on a 128-byte bias row; and the
    kernel's whole per-rating item update, row + bias row), as the best rate
    over its wave counts per form."""
    import subprocess
    global _ATOMIC_CEILING
    if _ATOMIC_CEILING is not None:
        return _ATOMIC_CEILING
    exe = os.path.join(ROOT, "netflix_prize_amd", "atomic-ceiling")
    if not os.path.exists(exe):
        return None
    out = subprocess.run([exe, "1000"], capture_output=True, text=True, timeout=120)
    if out.returncode != 0:
        log(f"bench: atomic-ceiling failed: {out.stderr[-500:]}")
        return None
    best = {}
    for line in out.stdout.splitlines():
        r = json.loads(line)
        if r["lanes_per_s"] > best.get(r["form"], {}).get("lanes_per_s", 0.0):
            best[r["form"]] = r
    _ATOMIC_CEILING = {k: {"rows_per_s": v["rows_per_s"], "lanes_per_s": v["lanes_per_s"],
                           "atomic_instr_per_s": v.get("atomic_instr_per_s"), "waves": v["waves"]}
                       for k, v in best.items()}
    return _ATOMIC_CEILING


def atomic_cost_model(ceil: dict) -> dict | None:
    """Time per FP32 atomic instruction as a x instructions + b x 128-byte
    lines + c x dwords, solved from three probe forms (scatter-atomic: 1,
    64, 64; f32-atomic row: 4, 8, 256; bias-atomic: 1, 1, 2) and checked on a
    fourth (row+bias-atomic: 5, 9, 258)."""
    need = ("scatter-atomic", "f32-atomic", "bias-atomic", "row+bias-atomic")
    if not ceil or any(k not in ceil for k in need):
        return None
    A = np.array([[1.0, 64.0, 64.0], [4.0, 8.0, 256.0], [1.0, 1.0, 2.0]])
    t = np.array([1.0 / ceil[k]["rows_per_s"] for k in need[:3]])
    a, b, c = np.linalg.solve(A, t)
    pred = 5 * a + 9 * b + 258 * c
    meas = 1.0 / ceil["row+bias-atomic"]["rows_per_s"]
    return {"s_per_instr": a, "s_per_line": b, "s_per_dword": c, "check_rel_err": (pred - meas) / meas}


def nbr_atomic_work(tr, pr, max_nbr: int, seed: int = 0) -> dict:
    """Per training rating, the atomic updates of the global neighbourhood
    terms (src/tx.rs:1041-1070): w(i, j) for the sampled rated neighbours,
    c(i, j) for every sampled neighbour, all in row i of their table.  The
    sample follows the reference's budget (src/tx.rs:763-781), drawn here
    uniformly with a fixed seed; lines = distinct 32-element (128-byte)
    column blocks of the row; instructions = two (w, c) per 64 neighbours."""
    U = tr.n_users
    toff = np.searchsorted(tr.user_idxs, np.arange(U + 1))
    poff = np.searchsorted(pr.user_idxs, np.arange(U + 1))
    nr = np.diff(toff).astype(np.int64)
    rng = np.random.default_rng(seed)
    lines = dwords = instr = 0.0
    for u in np.nonzero(nr)[0]:
        R = tr.item_idxs[toff[u]:toff[u + 1]]
        P = np.setdiff1d(pr.item_idxs[poff[u]:poff[u + 1]], R)
        n = len(R) + len(P)
        if n > max_nbr:
            rb = min(max(int(round(max_nbr * len(R) / n)), 1), len(R))
            ib = min(max_nbr - rb, len(P))
            R = rng.choice(R, rb, replace=False)
            P = rng.choice(P, ib, replace=False) if ib else P[:0]
        k = nr[u]
        lines += k * (len(np.unique(R >> 5)) + len(np.unique(np.concatenate([R, P]) >> 5)))
        dwords += k * (2 * len(R) + len(P))
        instr += k * 2 * (1 + (len(R) + len(P) > 64))
    N = float(nr.sum())
    return {"lines": lines / N, "dwords": dwords / N, "instructions": instr / N}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(tr, pr, cfg, target_s: float, workers: int) -> dict:
    """Serial oracle restatement of TxModel::fit_epoch on one core, one epoch
    over a bounded sample of users (first users of the training set)."""
    from types import SimpleNamespace

    from oracle import oracle as O

    # size the sample to ~target_s of CPU work from a quick probe rate
    def subset(n_users):
        a, b = 0, int(np.searchsorted(tr.user_idxs, n_users))
        c = int(np.searchsorted(pr.user_idxs, n_users)) if len(pr.user_idxs) else 0

        def sl(ds, hi):
            d = SimpleNamespace()
            d.user_idxs = ds.user_idxs[:hi]
            d.item_idxs = ds.item_idxs[:hi]
            d.dates = ds.dates[:hi]
            d.residuals = ds.residuals[:hi]
            d.raw_ratings = ds.raw_ratings[:hi]
            d.is_test = ds.is_test[:hi]
            d.user_cnts = ds.user_cnts[:n_users] if ds is tr else np.bincount(ds.user_idxs[:hi], minlength=n_users).astype(np.int32)
            d.n_users = n_users
            d.n_items = ds.n_items
            return d
        return sl(tr, b), sl(pr, c)

    probe_users = 2000
    t_tr, t_pr = subset(probe_users)
    m = O.OracleModel(cfg, t_tr, t_pr)
    t0 = time.perf_counter()
    m.fit_epoch(1)
    rate = len(t_tr.user_idxs) / max(time.perf_counter() - t0, 1e-6)
    want = int(rate * target_s)
    n_users = int(np.searchsorted(np.cumsum(tr.user_cnts), want)) + 1
    n_users = max(probe_users, min(n_users, tr.n_users))
    s_tr, s_pr = subset(n_users)
    m = O.OracleModel(cfg, s_tr, s_pr)
    t0 = time.perf_counter()
    m.fit_epoch(1)
    dt = time.perf_counter() - t0
    out = {"value": len(s_tr.user_idxs) / dt, "unit": "ratings/s/epoch", "cores": 1, "kind": "port",
           "sample": f"one serial TxModel::fit_epoch (oracle restatement, incl. ycache rebuild) over the first "
                     f"{n_users} users = {len(s_tr.user_idxs)} training ratings, same model/config; {dt:.1f} s"}
    if workers > 1:
        cfgp = dict(vars(cfg)) if not isinstance(cfg, dict) else dict(cfg)
        cfgp.update(nsvd_norm_exp=1.0, n_threads=workers, seq_epochs=[0, 0, 0, 0, 0])
        mp = O.OracleModel(cfgp, s_tr, s_pr)
        t0 = time.perf_counter()
        mp.fit_epoch_p(1, workers)
        dtp = time.perf_counter() - t0
        out["parallel_variant"] = {"value": len(s_tr.user_idxs) / dtp, "cores": workers,
                                   "what": "Tsvdx4pModel user+item passes (restated), same sample"}
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--f", type=int, default=256)
    ap.add_argument("--shape", choices=["full", "small"], default="full")
    ap.add_argument("--strata", type=int, default=1)
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--blocks", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-workers", type=int, default=0, help="threads for the parallel CPU variant (0 = auto)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--ttr-epochs", type=int, default=20, help="epochs of the time-to-RMSE run (0 = skip)")
    ap.add_argument("--autotune", type=int, default=1, help="autotune the batched form of the time-to-RMSE run")
    ap.add_argument("--fold-in-users", type=int, default=80_189,
                    help="users held out of training and folded in afterwards (0 = skip)")
    ap.add_argument("--knnf", type=int, default=1,
                    help="time KnnfModel (similarity matrix + held-out predictions) over the model's item factors")
    ap.add_argument("--knnf-f", type=int, default=0,
                    help="use random item factors of this width instead of the trained model's (0 = trained)")
    ap.add_argument("--frozen-epochs", type=int, default=4,
                    help="epochs of the tsvdx4p schedule with FP32 and with BF16 frozen sides (0 = skip)")
    ap.add_argument("--bucket-epochs", type=int, default=20,
                    help="epochs of the length-bucketed launches (user_order 3) run (0 = skip)")
    ap.add_argument("--nbr-epochs", type=int, default=3,
                    help="epochs of each tsvdx5/6 neighbourhood model (0 = skip)")
    ap.add_argument("--day-epochs", type=int, default=3,
                    help="epochs of the tsvdx4-60 job's model (day terms, f=60) in the parallel schedule (0 = skip)")
    args = ap.parse_args(argv)

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on exactly one GPU; multi-GPU runs are not part of it "
              "(see the sharding design notes)", file=sys.stderr)
        return 2

    import torch  # first: the libraries then share torch's HIP runtime
    from netflix_prize_amd import synth, tx

    t0 = time.time()
    spec = synth.FULL if args.shape == "full" else synth.SMALL
    tr, pr = synth.generate(spec, args.seed)
    log(f"bench: data {spec.n_users}x{spec.n_items}, {tr.n_ratings} train / {pr.n_ratings} held out "
        f"({time.time() - t0:.1f}s)")
    cfg = tx.plain_form(tx.TxConfig(n_feat=args.f, n_epochs=20, n_time_bins=30, low_memory=True,
                                    reset_u_epoch=10_000))
    t0 = time.time()
    model = tx.TxModel(tr, pr, cfg)
    t_create = time.time() - t0
    model.set_schedule(mode=tx.SCHED_STRATA, n_strata=args.strata, use_graph=bool(args.graph), blocks=args.blocks,
                       experimental=args.strata > 1)
    team_split = bool(model.schedule()["split_waves"])
    log(f"bench: model created ({t_create:.1f}s); schedule {model.schedule()}")

    have_torch_gpu = torch.cuda.is_available()  # synchronisation of the device around the timed region

    epoch = 0
    for _ in range(args.warmup):
        epoch += 1
        model.fit_epoch(epoch=epoch)
        log(f"bench: warmup epoch {epoch} {model.last_epoch_stats()['ms_total']:.1f} ms")
    if have_torch_gpu:
        torch.cuda.synchronize()
    t_start = time.perf_counter()
    stats = []
    for _ in range(args.steps):
        epoch += 1
        model.fit_epoch(epoch=epoch)
        stats.append(model.last_epoch_stats())
    if have_torch_gpu:
        torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    for k, s in enumerate(stats):
        log(f"bench: epoch {k + 1}: total {s['ms_total']:.2f} ms, sgd {s['ms_sgd']:.2f} ms, "
            f"rebuild {s['ms_rebuild']:.2f} ms")
    ms_per_step = wall * 1000.0 / max(args.steps, 1)
    n_train = tr.n_ratings
    value = n_train / (ms_per_step / 1000.0)
    rmse = model.calc_rmse(pr)
    log(f"bench: holdout RMSE after {epoch} epochs: {rmse:.5f}")

    # time-to-holdout-RMSE: fresh models trained for the full epoch budget in
    # the batched form (length buckets, autotuned teams, hipGraph epoch), in the
    # plain GPU form (TX_SCHED_PLAIN: one wave per user over the epoch's user
    # order, no prefetch, queue, buckets or graph, 1,024 waves) and in the
    # tsvdx4p schedule with BF16 frozen sides; GPU epoch time until the holdout
    # RMSE is within 1e-3 of the serial trainer's after the same epochs (or of
    # the form's own final RMSE); tx_create time reported beside it
    ttr = None
    if args.ttr_epochs > 0:
        ref = None
        gpath = os.path.join(ROOT, "tests", "golden", f"oracle_curve_{args.shape}_f{cfg.n_feat}.json")
        if os.path.exists(gpath):
            g = json.load(open(gpath))
            if g.get("seed") == args.seed and len(g["epochs"]) >= args.ttr_epochs:
                ref = g["epochs"][args.ttr_epochs - 1]["holdout_rmse"]

        def curve_of(form):
            t0 = time.time()
            fcfg = cfg.replace(nsvd_norm_exp=1.0) if form == "tsvdx4p_bf16" else cfg
            m2 = tx.TxModel(tr, pr, fcfg)
            created = time.time() - t0
            tuned = None
            if form == "batched":
                m2.set_schedule(mode=tx.SCHED_STRATA, n_strata=args.strata, use_graph=bool(args.graph),
                                blocks=args.blocks, experimental=args.strata > 1)
                if args.autotune:
                    tuned = m2.autotune(1)
                    tuned["trials"] = [{k: (round(v, 2) if k == "ms" else v) for k, v in t.items()
                                        if k not in ("bucket", "lanes_per_user", "waves_per_user")}
                                       for t in m2.autotune_trials()]
            elif form == "plain":
                m2.set_schedule(mode=tx.SCHED_PLAIN, blocks=1024)
            else:
                m2.set_schedule(mode=tx.SCHED_TSVDX4P)
                m2.set_frozen_precision("bf16")
            curve, secs = [], []
            for e in range(1, args.ttr_epochs + 1):
                m2.fit_epoch(epoch=e)
                secs.append(m2.last_epoch_stats()["ms_total"] / 1000.0)
                curve.append(m2.calc_rmse(pr))
            m2.close()
            return curve, secs, tuned, created

        ttr = {"epochs": args.ttr_epochs, "serial_reference_rmse": None if ref is None else round(ref, 5)}
        for form in ("batched", "plain", "tsvdx4p_bf16"):
            curve, secs, tuned, created = curve_of(form)
            target = (ref if ref is not None else curve[-1]) + 1e-3
            hit = next((k for k, r in enumerate(curve) if r <= target), None)
            ttr[form] = {"final_holdout_rmse": round(curve[-1], 5), "target_rmse": round(target, 5),
                         "epochs_to_target": None if hit is None else hit + 1,
                         "seconds_to_target": None if hit is None else round(sum(secs[:hit + 1]), 3),
                         "seconds_total": round(sum(secs), 3), "tx_create_s": round(created, 2),
                         "ms_per_epoch": round(statistics.median(secs[1:] or secs) * 1000.0, 2),
                         "curve": [round(r, 5) for r in curve]}
            if tuned is not None:
                ttr[form]["autotuned_schedule"] = {k: tuned[k] for k in ("blocks", "user_order", "heavy_min",
                                                                          "n_strata", "split_waves", "trials")}
            log(f"bench: time-to-RMSE {form}: {ttr[form]}")

    # length-bucketed launches (user_order 3): the users of 1-16, 17-256,
    # 257-4096 and 4097+ training ratings in four queues and launches, each with
    # its own team shape and teams in flight, autotuned per bucket after the
    # first epoch; epoch time, holdout RMSE after the epoch budget and the
    # per-bucket choice against the one-queue default
    buckets = None
    if args.bucket_epochs > 0:
        bm = tx.TxModel(tr, pr, cfg)
        bm.set_schedule(mode=tx.SCHED_STRATA, user_order=tx.USER_ORDER_BUCKETS)
        bms = []
        for e in range(1, args.bucket_epochs + 1):
            bm.fit_epoch(epoch=e)
            if e == 1:
                bm.autotune(1)
                bucket_trials = [{k: (round(v, 2) if k == "ms" else v) for k, v in t.items()
                                  if k not in ("n_strata", "blocks", "user_order", "heavy_min")}
                                 for t in bm.autotune_trials()]
            bms.append(bm.last_epoch_stats()["ms_total"])
        r_b = bm.calc_rmse(pr)
        buckets = {"epochs": args.bucket_epochs, "ms_per_epoch": round(statistics.median(bms[1:] or bms), 2),
                   "final_holdout_rmse": round(r_b, 5),
                   "buckets": [{k: (round(v, 3) if isinstance(v, float) else v) for k, v in b.items()}
                               for b in bm.buckets()],
                   "autotune_trials": bucket_trials}
        bm.close()
        log(f"bench: length-bucketed launches: {buckets}")

    # new users folded into a model trained on the other users: the encoder's
    # warm start against the reference's random start
    fold = None
    if args.fold_in_users > 0:
        from netflix_prize_amd import encoder as E
        n_new = args.fold_in_users if args.fold_in_users < spec.n_users else spec.n_users // 6
        log(f"bench: fold-in of {n_new} new users (encoder vs random start)")
        fold = E.evaluate_new_users(tr, pr, cfg, spec.n_users - n_new, args.ttr_epochs or 20, 5, 20, log=log)
        fold["gemm"] = E.gemm_benchmark(E.n_features(cfg.n_feat), cfg.n_feat + 1)
        log(f"bench: encoder step products: {fold['gemm']}")

    # KnnfModel (the *__knnf jobs) over the trained model's item factors: the
    # cosine similarity matrix by FP32 MFMA at create, then every held-out query
    knnf = None
    if args.knnf:
        from netflix_prize_amd import knn
        if args.knnf_f > 0:
            F = np.random.default_rng(0).normal(size=(tr.n_items, args.knnf_f)).astype(np.float32)
        else:
            F = model.get("ifeat").reshape(tr.n_items, cfg.n_feat)
        t0 = time.perf_counter()
        km = knn.KnnfModel(tr, F)
        t_knn = time.perf_counter() - t0
        km.predict(pr.user_idxs, pr.item_idxs, pr.dates)
        t0 = time.perf_counter()
        km.predict(pr.user_idxs, pr.item_idxs, pr.dates)
        t_pred = time.perf_counter() - t0
        km.close()
        ip, fp = (tr.n_items + 63) // 64 * 64, (F.shape[1] + 31) // 32 * 32
        knnf = {"workload": f"KnnfModel k=25 over {tr.n_items} x {F.shape[1]} item factors "
                            f"({'random' if args.knnf_f else 'trained'}), {pr.n_ratings} held-out queries",
                "create_s": round(t_knn, 4), "cosine_gflop": round(2.0 * ip * ip * fp / 1e9, 1),
                "predict_s": round(t_pred, 4), "queries_per_s": round(pr.n_ratings / t_pred, 1)}
        # KnnsModel (the *__knns jobs): support of every user's training and
        # held-out item pairs, then the same held-out queries
        t0 = time.perf_counter()
        km = knn.KnnsModel(tr, pr, knn.KnnsConfig())
        t_knn = time.perf_counter() - t0
        km.predict(pr.user_idxs, pr.item_idxs, pr.dates)
        t0 = time.perf_counter()
        km.predict(pr.user_idxs, pr.item_idxs, pr.dates)
        t_pred = time.perf_counter() - t0
        km.close()
        nu = np.bincount(tr.user_idxs, minlength=tr.n_users) + np.bincount(pr.user_idxs, minlength=tr.n_users)
        knnf["knns"] = {"create_s": round(t_knn, 4), "support_pairs": int(np.sum(nu.astype(np.int64) ** 2)),
                        "predict_s": round(t_pred, 4), "queries_per_s": round(pr.n_ratings / t_pred, 1)}
        # Knn3Model (the *__knn3 jobs, Knn3Config::default) on the training
        # set's residual targets: co-rating counts, candidates, the FP64 normal
        # equations and NNLS of every item, then the held-out queries
        t0 = time.perf_counter()
        km = knn.Knn3Model(tr, knn.Knn3Config())
        t_knn = time.perf_counter() - t0
        km.predict(pr.user_idxs, pr.item_idxs)
        t0 = time.perf_counter()
        km.predict(pr.user_idxs, pr.item_idxs)
        t_pred = time.perf_counter() - t0
        n_w = int(sum(len(x) for x in km.neighbors()))
        km.close()
        nt = np.bincount(tr.user_idxs, minlength=tr.n_users).astype(np.int64)
        knnf["knn3"] = {"create_s": round(t_knn, 4), "pair_lookups": int(np.sum(nt ** 2)), "weights": n_w,
                        "predict_s": round(t_pred, 4), "queries_per_s": round(pr.n_ratings / t_pred, 1)}
        log(f"bench: knnf {knnf}")

    # the tsvdx4p schedule (Tsvdx4pModel: user pass, NSVD1 gather, item pass)
    # with the frozen side of each pass read in FP32 and from BF16 copies
    frozen = None
    if args.frozen_epochs > 0:
        pcfg = tx.TxConfig(n_feat=cfg.n_feat, n_epochs=20, n_time_bins=30, nsvd_norm_exp=1.0, lr_yb=2e-4,
                           reset_u_epoch=10_000)
        frozen = {"workload": f"tsvdx4p schedule, lm set, f={cfg.n_feat}, nsvd_norm_exp 1, "
                              f"{args.frozen_epochs} epochs (the first builds the by-item index)"}
        for prec in ("f32", "bf16"):
            pm = tx.TxModel(tr, pr, pcfg)
            pm.set_schedule(mode=tx.SCHED_TSVDX4P)
            pm.set_frozen_precision(prec)
            st = []
            for e in range(1, args.frozen_epochs + 1):
                pm.fit_epoch(epoch=e)
                st.append(pm.last_epoch_stats())
            later = st[1:] or st
            frozen[prec] = {"ms_per_epoch": round(statistics.median(x["ms_total"] for x in later), 2),
                            "ms_user_pass_and_nsvd1": round(statistics.median(x["ms_sgd"] for x in later), 2),
                            "holdout_rmse": round(pm.calc_rmse(pr), 6)}
            pm.close()
            log(f"bench: tsvdx4p {prec} frozen sides: {frozen[prec]}")

    # the tsvdx4-60 job's model (lm set with day terms, low_memory false, f=60,
    # 29 time bins, seed 42) in the default parallel schedule: epoch times and
    # the holdout RMSE against the serial oracle's after the same epochs
    day = None
    if args.day_epochs > 0:
        dcfg = tx.TxConfig(n_feat=60, n_epochs=20, seed=42, n_time_bins=29, low_memory=False)
        t0 = time.time()
        dm = tx.TxModel(tr, pr, dcfg)
        dm.set_schedule(mode=tx.SCHED_STRATA)   # stream launches: ms_sgd is the SGD launch alone
        log(f"bench: day-term model created ({time.time() - t0:.1f}s)")
        dst, dcurve = [], []
        for e in range(1, args.day_epochs + 1):
            dm.fit_epoch(epoch=e)
            dst.append(dm.last_epoch_stats())
            dcurve.append(dm.calc_rmse(pr))
            log(f"bench: day-term epoch {e}: {dst[-1]['ms_total']:.1f} ms, holdout RMSE {dcurve[-1]:.5f}")
        dm.close()
        gpath = os.path.join(ROOT, "tests", "golden", f"oracle_curve_{args.shape}_job60.json")
        dref = None
        if os.path.exists(gpath):
            g = json.load(open(gpath))
            if g.get("seed") == args.seed and len(g["epochs"]) >= args.day_epochs:
                dref = [x["holdout_rmse"] for x in g["epochs"][:args.day_epochs]]
        ms = statistics.median(x["ms_total"] for x in dst[1:] or dst)
        sgd = statistics.median(x["ms_sgd"] for x in dst[1:] or dst)
        # training days: the rows are sorted by (user, date)
        u, d = tr.user_idxs, tr.dates
        n_days = int(np.count_nonzero((u[1:] != u[:-1]) | (d[1:] != d[:-1]))) + 1
        qf = dcfg.lr_iqf != 0.0 or dcfg.sigma_iqf != 0.0
        bpr = (algorithmic_bytes_per_rating(60, n_train, tr.n_users, qf, dcfg.lr_u2 != 0.0)["sgd"]
               + day_bytes_per_rating(60, n_train, pr.n_ratings, n_days))
        ach = bpr * n_train / (sgd / 1000.0) / 1e9
        day = {"workload": "tsvdx4-60 model: lm set, day terms, f=60, 29 time bins, 1 stratum",
               "ms_per_epoch": round(ms, 2), "ratings_per_s": round(n_train / (ms / 1000.0), 1),
               "holdout_rmse": [round(r, 5) for r in dcurve],
               "serial_reference_rmse": None if dref is None else [round(r, 5) for r in dref],
               "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": profiled_traffic(60, args, "day_"),
                            "kernel": "k_sgd_par (day terms)", "bytes_per_rating": round(bpr, 1),
                            "avg_launch_ms": round(sgd, 3), "training_days": n_days}}

    # the tsvdx5 neighbourhood model (tsvdx5-150: global w/c terms, f=150,
    # 100 sampled neighbours) in the default parallel schedule, with atomic and
    # with plain updates of w and c: epoch times, the holdout RMSE against the
    # same model without the neighbourhood terms, and the SGD kernel's
    # algorithmic bytes.  (The same-day tsvdx6 model is not run at this shape
    # here; see the design notes.)
    nbr = None
    if args.nbr_epochs > 0:
        nbr = {}
        u, d = tr.user_idxs, tr.dates
        n_days = int(np.count_nonzero((u[1:] != u[:-1]) | (d[1:] != d[:-1]))) + 1
        for name, f, same_day in (("tsvdx5-150", 150, False),):
            ncfg = tx.TxConfig(n_feat=f, n_epochs=23, seed=42, n_time_bins=32, low_memory=False, reset_u_epoch=1024,
                               max_neighbors=100, lr_w=0.0015, lr_c=0.0015, reg_w=0.002, reg_c=0.002,
                               w_bias=0.8, w_factor=0.8, w_nbr=0.8, lambda1=25.0, lambda2=10.0)
            if same_day:
                ncfg = ncfg.replace(lr_w_day=0.001, lr_c_day=0.001, reg_w_day=0.002, reg_c_day=0.002)
            res = {}
            forms = [("nbr", ncfg, 0), ("nbr_plain_updates", ncfg, 2),
                     ("without_nbr", ncfg.replace(lr_w=0.0, lr_c=0.0, lr_w_day=0.0, lr_c_day=0.0), 0)]
            for form, c, nb_upd in forms:
                t0 = time.time()
                nm = tx.TxModel(tr, pr, c)
                t_c = time.time() - t0
                # stream launches: ms_sgd is the SGD launch alone; nbr_plain_updates:
                # the neighbourhood tables by plain read-modify-write instead of atomics
                nm.set_schedule(mode=tx.SCHED_STRATA, nbr_update=nb_upd)
                nst, ncurve = [], []
                for e in range(1, args.nbr_epochs + 1):
                    nm.fit_epoch(epoch=e)
                    nst.append(nm.last_epoch_stats())
                    ncurve.append(nm.calc_rmse(pr))
                nm.close()
                res[form] = {"ms_per_epoch": round(statistics.median(x["ms_total"] for x in nst[1:] or nst), 2),
                             "ms_sgd": round(statistics.median(x["ms_sgd"] for x in nst[1:] or nst), 3),
                             "holdout_rmse": [round(r, 5) for r in ncurve], "tx_create_s": round(t_c, 2)}
                log(f"bench: {name} {form}: {res[form]}")
            bpr = (algorithmic_bytes_per_rating(f, n_train, tr.n_users, True, True)["sgd"]
                   + day_bytes_per_rating(f, n_train, pr.n_ratings, n_days)
                   + nbr_bytes_per_rating(tr, pr, 100, same_day))
            sgd = res["nbr"]["ms_sgd"]
            ach = bpr * n_train / (sgd / 1000.0) / 1e9
            res["ratings_per_s"] = round(n_train / (res["nbr"]["ms_per_epoch"] / 1000.0), 1)
            res["roofline"] = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": round(ach / HBM_PEAK_GBS, 4), "kernel": "k_sgd_par (neighbourhood)",
                               "bytes_per_rating": round(bpr, 1), "avg_launch_ms": sgd}
            # the neighbourhood terms' increment of the SGD launch against the
            # atomic floor of their element updates (the probe's cost model)
            cost = atomic_cost_model(atomic_ceiling())
            if cost is not None:
                t0 = time.time()
                wk = nbr_atomic_work(tr, pr, 100)
                per = (cost["s_per_instr"] * wk["instructions"] + cost["s_per_line"] * wk["lines"]
                       + cost["s_per_dword"] * wk["dwords"])
                inc = res["nbr"]["ms_sgd"] - res["without_nbr"]["ms_sgd"]
                floor = per * n_train * 1000.0
                res["atomic"] = {"per_rating": {k: round(float(v), 2) for k, v in wk.items()},
                                 "ns_per_rating": round(float(per) * 1e9, 3), "floor_ms": round(float(floor), 1),
                                 "increment_ms": round(inc, 1),
                                 "frac": round(float(floor / inc), 4) if inc > 0 else None,
                                 "model": {k: float(f"{v:.4g}") for k, v in cost.items()}}
                log(f"bench: {name} neighbourhood atomics ({time.time() - t0:.1f}s): {res['atomic']}")
            res["workload"] = (f"{name} model: lm set, day terms, f={f}, 32 time bins, max_neighbors 100"
                               + (", same-day terms" if same_day else "") + ", 1 stratum")
            nbr[name] = res

    # roofline of the dominant kernel (SGD), from the library's events on its stream
    qf = cfg.lr_iqf != 0.0 or cfg.sigma_iqf != 0.0
    p2 = cfg.lr_u2 != 0.0
    ab = algorithmic_bytes_per_rating(cfg.n_feat, n_train, tr.n_users, qf, p2)
    graph = bool(stats[0]["graph"]) if stats else False
    sgd_ms = statistics.median(s["ms_sgd"] for s in stats)
    launches = args.strata
    if graph:
        # in graph mode ms_sgd spans the whole captured epoch; take out the rebuild
        # share measured in a non-graph epoch
        model.set_schedule(mode=tx.SCHED_STRATA, n_strata=args.strata, use_graph=False, blocks=args.blocks,
                           experimental=args.strata > 1)
        epoch += 1
        model.fit_epoch(epoch=epoch)
        s = model.last_epoch_stats()
        sgd_ms = s["ms_sgd"]
    avg_launch_ms = sgd_ms / launches
    bytes_per_launch = ab["sgd"] * n_train / launches
    achieved = bytes_per_launch / (avg_launch_ms / 1000.0) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": profiled_traffic(cfg.n_feat, args),
                "kernel": "k_sgd_par", "bytes_per_rating": round(ab["sgd"], 1),
                "avg_launch_ms": round(avg_launch_ms, 3), "launches_per_epoch": launches}
    # the item-row updates are FP32 atomic adds, which the device performs at a
    # rate set by the 128-byte lines they touch, not by their lanes (a 2-lane
    # add costs about as much as a 64-lane add on one line): the lines one
    # launch updates against the device's measured line rate for the kernel's
    # own per-rating pattern (q row + item-bias row, row+bias-atomic)
    ceil = atomic_ceiling()
    if ceil and "row+bias-atomic" in ceil:
        row_lines = -(-cfg.n_feat * 4 // 128)                   # q row (and q_f row) in 128-byte lines
        lines = row_lines * (2 if qf else 1) + 1                 # + the item-bias row (b_i, b_i,bin, b_i,f)
        rate = ceil["row+bias-atomic"]["rows_per_s"] * (256 * 4 // 128 + 1)
        got = lines * n_train / launches / (avg_launch_ms / 1000.0)
        roofline["atomic"] = {"unit": "atomic 128-byte lines/s", "achieved": round(got, 1), "ceiling": round(rate, 1),
                              "frac": round(got / rate, 4), "lines_per_rating": lines,
                              "floor_ms": round(lines * n_train / launches / rate * 1000.0, 3),
                              "measured_forms": ceil}
        log(f"bench: atomic roofline {roofline['atomic']}")

    cpu = None
    # host cores: the process's affinity; worker pools stay within the 16-core
    # share the GPU box allots one GPU
    n_aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    if not args.no_cpu:
        workers = args.cpu_workers or min(16, n_aff)
        log(f"bench: CPU baseline (~{args.cpu_seconds:.0f}s serial + parallel variant on {workers} threads)")
        cpu = cpu_baseline(tr, pr, cfg, args.cpu_seconds, workers)
        cpu["host_cpu_count"] = os.cpu_count()
        cpu["host_affinity_cores"] = n_aff

    line = {
        "metric": "timeSVD++ SGD ratings/s/epoch (f=256); time-to-holdout-RMSE vs plain GPU form",
        "value": round(value, 1),
        "unit": "ratings/s/epoch",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded planted rank-16 model, Netflix-shaped counts)",
        "config": {"workload": f"plain timeSVD++ {spec.n_users}x{spec.n_items}, {spec.n_ratings} ratings, "
                               f"f={cfg.n_feat}, FP32, 30 time bins, low-memory, "
                               f"{args.strata} strat{'um' if args.strata == 1 else 'a'}, "
                               f"{'split teams (rating wave + streaming wave)' if team_split else 'one-wave teams'}, "
                               f"{'hipGraph epoch' if args.graph else 'stream launches'}",
                   "n_train": n_train, "n_heldout": pr.n_ratings, "holdout_rmse": round(rmse, 5),
                   "epochs_trained": epoch, "tx_create_s": round(t_create, 2)},
        "roofline": roofline,
        "time_to_rmse": ttr,
        "fold_in": fold,
        "day_terms": day,
        "neighbourhood": nbr,
        "length_buckets": buckets,
        "tsvdx4p_frozen": frozen,
        "knnf": knnf,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    model.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
