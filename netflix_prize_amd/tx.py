"""Python mirror of the reference's tsvdx4 interface over libtsvd_tx.so.

Names and argument meanings follow the reference:
  * :class:`TxConfig`      — TxConfig (src/tx.rs:128-204)
  * :class:`OrdinalHeadConfig` — src/lib.rs:502-507
  * :class:`Dataset`       — Dataset / Dataset::load (src/lib.rs:227-290)
  * :class:`TxModel`       — impl Regressor for TxModel (src/tx.rs:457-1116):
    ``TxModel(tr, pr, cfg)``, ``n_epochs()``, ``fit_epoch(tr, pr, epoch)``,
    ``predict(u, i, day)``, ``n_subscores()``, ``subscore_names()``,
    ``predict_subscores(...)``, ``save_artifacts(model_name, tr_set, preds_dir)``
  * :func:`calc_rmse`, :func:`gen_preds`, :func:`rmse_from_preds`,
    :func:`calc_gbias`, :func:`calc_user_offsets`, :func:`get_users`
    (src/lib.rs:463-494, 659-756)

Every compute call runs on the GPU through the C ABI of include/tsvd_tx.h.
There is no CPU fallback: if the library is missing, loading fails loudly.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libtsvd_tx.so")

PARAMS = ["ubias", "ibias", "ufeat", "ufeat2", "ifeat", "yfeat", "ycache", "yfeat_day",
          "tu_mean", "alpha_u", "but_bin", "bit_bin", "ybias", "ycache_bias", "ubias_day",
          "ufeat_day", "ycache_day", "ibias_freq", "ifeat_freq", "cu", "cut", "gbias",
          "thresholds", "w", "c", "w_day", "c_day"]
# the dense item x item neighbourhood tables (empty unless the model has them)
NBR_PARAMS = ("w", "c", "w_day", "c_day")
PARAM_ID = {n: k for k, n in enumerate(PARAMS)}

SCHED_SERIAL = 0
SCHED_STRATA = 1
SCHED_TSVDX4P = 2  # Tsvdx4pModel epochs (src/tsvdx4p.rs:798-805)
SCHED_PLAIN = 3    # the plain per-user wavefront kernel
NOALLOC_MAX_ROWS = 1 << 20  # TX_NOALLOC_MAX_ROWS


class TxError(RuntimeError):
    """A failed library call (the reference panics in the same situations)."""


@dataclass
class OrdinalHeadConfig:
    th_init: Sequence[float] = (0.5, 1.25, 3.75, 5.5)
    th_gap: float = 0.001
    lr_t: float = 0.0
    reg_t: float = 0.0


@dataclass
class TxConfig:
    """TxConfig (src/tx.rs:128-204); defaults are the tsvdx4 "lm" set
    (src/bin/tsvdx4-new.rs:130-154) with tx4-mode neighbourhood knobs."""
    n_feat: int = 60
    n_epochs: int = 20
    seed: int = 42
    shuffle_users: bool = True
    n_time_bins: int = 29
    beta: float = 0.3
    n_freq_bins: int = 16
    lr_u: float = 0.003
    lr_ud: float = 0.00125
    lr_u2: float = 7e-6
    lr_ub: float = 0.0031
    lr_ubd: float = 0.003
    lr_i: float = 0.0036
    lr_ib: float = 0.0036
    lr_y: float = 0.0005
    lr_yb: float = 2.5e-5
    lr_yd: float = 0.000267
    lr_tu: float = 0.0
    lr_ti: float = 0.000225
    lr_ta: float = 2.25e-5
    lr_ibf: float = 5e-5
    lr_iqf: float = 5e-6
    lr_cu: float = 0.002
    reg_iqf: float = 0.007
    reg_cu: float = 0.01
    reg_u: float = 0.0504
    reg_u2: float = 0.4
    reg_ud: float = 0.04
    reg_i: float = 0.00735
    reg_y: float = 0.04
    reg_yd: float = 0.02667
    sigma_iqf: float = 0.005
    sigma_u: float = 0.0015
    sigma_i: float = 0.005
    sigma_y: float = 0.00333
    sigma_yd: float = 0.009
    reset_u_epoch: int = 13
    max_neighbors: int = 0
    lr_w: float = 0.0
    lr_c: float = 0.0
    reg_w: float = 0.0
    reg_c: float = 0.0
    lr_w_day: float = 0.0
    lr_c_day: float = 0.0
    reg_w_day: float = 0.0
    reg_c_day: float = 0.0
    w_bias: float = 1.0
    w_factor: float = 1.0
    w_nbr: float = 1.0
    sum_err_bug: bool = False
    lambda1: float = 0.0
    lambda2: float = 0.0
    ordinal_head: Optional[OrdinalHeadConfig] = None
    save_ifeat: bool = False
    low_memory: bool = True
    full_su: bool = True
    # Tsvdx4pConfig extras (src/tsvdx4p.rs:175-186), used by SCHED_TSVDX4P
    nsvd_norm_exp: float = 0.0
    n_threads: int = 0
    seq_epochs: tuple = (0, 0, 0, 0, 0)

    def replace(self, **kw) -> "TxConfig":
        return dataclasses.replace(self, **kw)

    @classmethod
    def from_dict(cls, d: dict) -> "TxConfig":
        kw = {k: v for k, v in d.items() if k in {f.name for f in dataclasses.fields(cls)}}
        oh = kw.get("ordinal_head")
        if isinstance(oh, dict):
            kw["ordinal_head"] = OrdinalHeadConfig(**oh)
        return cls(**kw)


def plain_form(cfg: TxConfig) -> TxConfig:
    """The "plain" timeSVD++ of the north-star configs: the lm set with the
    p2, cu, frequency, NSVD1-bias and per-user-bin terms held inert."""
    return cfg.replace(lr_u2=0.0, lr_cu=0.0, lr_ibf=0.0, lr_iqf=0.0, lr_yb=0.0, lr_tu=0.0, sigma_iqf=0.0)


# --------------------------------------------------------------------------
# ctypes structures (include/tsvd_tx.h)
# --------------------------------------------------------------------------

_F32 = ["beta", "lr_u", "lr_ud", "lr_u2", "lr_ub", "lr_ubd", "lr_i", "lr_ib", "lr_y", "lr_yb", "lr_yd",
        "lr_tu", "lr_ti", "lr_ta", "lr_ibf", "lr_iqf", "lr_cu", "reg_iqf", "reg_cu", "reg_u", "reg_u2",
        "reg_ud", "reg_i", "reg_y", "reg_yd", "sigma_iqf", "sigma_u", "sigma_i", "sigma_y", "sigma_yd",
        "lr_w", "lr_c", "reg_w", "reg_c", "lr_w_day", "lr_c_day", "reg_w_day", "reg_c_day",
        "w_bias", "w_factor", "w_nbr", "lambda1", "lambda2"]
_I32 = ["n_feat", "n_epochs", "n_time_bins", "n_freq_bins", "reset_u_epoch", "max_neighbors",
        "shuffle_users", "sum_err_bug", "low_memory", "full_su", "save_ifeat", "has_ordinal"]


class CConfig(C.Structure):
    _fields_ = ([("seed", C.c_uint64)] + [(n, C.c_int32) for n in _I32] + [(n, C.c_float) for n in _F32]
                + [("th_init", C.c_float * 4), ("th_gap", C.c_float), ("lr_t", C.c_float),
                   ("reg_t", C.c_float), ("nsvd_norm_exp", C.c_float), ("n_threads", C.c_int32),
                   ("seq_epochs", C.c_int32 * 5)])


class CRatings(C.Structure):
    _fields_ = [("n_ratings", C.c_int64), ("n_users", C.c_int32), ("n_items", C.c_int32),
                ("user_idxs", C.c_void_p), ("item_idxs", C.c_void_p), ("dates", C.c_void_p),
                ("residuals", C.c_void_p), ("raw_ratings", C.c_void_p), ("is_test", C.c_void_p),
                ("user_cnts", C.c_void_p)]


class CSchedule(C.Structure):
    _fields_ = [("mode", C.c_int32), ("n_strata", C.c_int32), ("teams_per_block", C.c_int32),
                ("blocks", C.c_int32), ("use_graph", C.c_int32), ("update_flags", C.c_int32),
                ("user_order", C.c_int32), ("heavy_min", C.c_int32), ("split_waves", C.c_int32)]


class CEpochStats(C.Structure):
    _fields_ = [("ms_total", C.c_double), ("ms_sgd", C.c_double), ("ms_rebuild", C.c_double),
                ("ms_other", C.c_double), ("n_ratings", C.c_int64), ("n_launches", C.c_int32),
                ("graph", C.c_int32)]


N_BUCKETS = 4  # TX_N_BUCKETS
USER_ORDER_BUCKETS = 3  # length buckets launched separately


class CBucketInfo(C.Structure):
    _fields_ = [("lo", C.c_int32), ("hi", C.c_int32), ("users", C.c_int64), ("ratings", C.c_int64),
                ("lanes_per_user", C.c_int32), ("waves_per_user", C.c_int32), ("teams", C.c_int32),
                ("one_launch", C.c_int32), ("ms", C.c_double)]


class CAutotuneTrial(C.Structure):
    _fields_ = [("bucket", C.c_int32), ("n_strata", C.c_int32), ("blocks", C.c_int32), ("user_order", C.c_int32),
                ("heavy_min", C.c_int32), ("lanes_per_user", C.c_int32), ("waves_per_user", C.c_int32),
                ("teams", C.c_int32), ("chosen", C.c_int32), ("split_waves", C.c_int32), ("ms", C.c_double)]


def config_struct(cfg: TxConfig) -> CConfig:
    c = CConfig()
    c.seed = int(cfg.seed)
    for n in _I32:
        if n == "has_ordinal":
            continue
        setattr(c, n, int(getattr(cfg, n)))
    c.has_ordinal = int(cfg.ordinal_head is not None)
    for n in _F32:
        setattr(c, n, float(getattr(cfg, n)))
    c.nsvd_norm_exp = float(cfg.nsvd_norm_exp)
    c.n_threads = int(cfg.n_threads)
    se = list(cfg.seq_epochs or ()) + [0] * 5
    for k in range(5):
        c.seq_epochs[k] = int(se[k])
    if cfg.ordinal_head is not None:
        for k in range(4):
            c.th_init[k] = float(cfg.ordinal_head.th_init[k])
        c.th_gap = float(cfg.ordinal_head.th_gap)
        c.lr_t = float(cfg.ordinal_head.lr_t)
        c.reg_t = float(cfg.ordinal_head.reg_t)
    return c


_lib = None


def lib() -> C.CDLL:
    """Load libtsvd_tx.so (built by netflix_prize_amd.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TxError(f"{LIB_PATH} is missing: build it with `python -m netflix_prize_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp, i64, i32 = C.c_void_p, C.c_int64, C.c_int32
        L.tx_last_error.restype = C.c_char_p
        L.tx_create.argtypes = [vp, vp, vp, C.POINTER(vp)]
        L.tx_destroy.argtypes = [vp]
        L.tx_set_schedule.argtypes = [vp, vp]
        L.tx_fit_epoch.argtypes = [vp, i32]
        L.tx_fit_users_epoch.argtypes = [vp, i32]
        L.tx_rebuild_ycache.argtypes = [vp]
        L.tx_set_table_precision.argtypes = [vp, i32]
        L.tx_autotune.argtypes = [vp, i32, vp]
        L.tx_last_epoch_stats.argtypes = [vp, vp]
        L.tx_predict.argtypes = [vp, i64, vp, vp, vp, vp]
        L.tx_predict_noalloc.argtypes = [vp, i64, vp, vp, vp, vp]
        L.tx_set_frozen_precision.argtypes = [vp, i32]
        L.tx_predict_device.argtypes = [vp, i64, vp, vp, vp, vp, vp]
        L.tx_predict_subscores.argtypes = [vp, i64, vp, vp, vp, vp, vp, vp, vp]
        L.tx_n_subscores.argtypes = [vp, C.POINTER(i32)]
        L.tx_get_buckets.argtypes = [vp, vp]
        L.tx_get_autotune_trials.argtypes = [vp, vp, i32, vp]
        L.tx_get_schedule.argtypes = [vp, vp]
        L.tx_set_bucket_shape.argtypes = [vp, i32, i32, i32, i32]
        L.tx_calc_rmse.argtypes = [vp, vp, C.POINTER(C.c_double)]
        L.tx_param_len.argtypes = [vp, i32, C.POINTER(i64)]
        L.tx_get_param.argtypes = [vp, i32, vp, i64]
        L.tx_set_param.argtypes = [vp, i32, vp, i64]
        L.tx_set_param_device.argtypes = [vp, i32, vp, i64, i64, vp]
        L.tx_get_users.argtypes = [i64, i32, C.c_uint64, i64, vp]
        L.tx_rand_normal.argtypes = [C.c_uint64, C.c_float, i64, vp]
        for name in ["tx_create", "tx_destroy", "tx_set_schedule", "tx_fit_epoch", "tx_fit_users_epoch",
                     "tx_rebuild_ycache", "tx_set_table_precision", "tx_autotune", "tx_get_autotune_trials",
                     "tx_get_schedule",
                     "tx_last_epoch_stats",
                     "tx_predict", "tx_predict_device", "tx_predict_subscores", "tx_n_subscores", "tx_calc_rmse",
                     "tx_get_buckets", "tx_set_bucket_shape",
                     "tx_param_len",
                     "tx_get_param", "tx_set_param", "tx_set_param_device", "tx_get_users", "tx_rand_normal"]:
            getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise TxError(f"{what} failed ({rc}): {lib().tx_last_error().decode()}")


# --------------------------------------------------------------------------
# Dataset (src/lib.rs:227-290)
# --------------------------------------------------------------------------

N_USERS = 480_189
N_ITEMS = 17_770


@dataclass
class Dataset:
    user_idxs: np.ndarray
    user_cnts: np.ndarray
    item_idxs: np.ndarray
    item_cnts: np.ndarray
    item_years: np.ndarray
    raw_ratings: np.ndarray
    residuals: np.ndarray
    dates: np.ndarray
    is_test: np.ndarray
    name: str = "ds"
    n_users: int = 0
    n_items: int = 0
    transposed: bool = False

    @property
    def n_ratings(self) -> int:
        return len(self.user_idxs)

    def contains_test_examples(self) -> bool:
        return bool(np.any(self.is_test != 0))

    @classmethod
    def load(cls, name: str, target_spec: str = "rtg", preds_dir: str = "preds_new", data_dir: str = "data",
             n_users: int = N_USERS, n_items: int = N_ITEMS) -> "Dataset":
        """Dataset::load (src/lib.rs:252-290): reads data/<name>/*.npy and
        builds residuals from a target spec ("rtg" or "w*model + ...")."""
        def rd(fname):
            path = os.path.join(data_dir, name, fname)
            try:
                return np.load(path, allow_pickle=False)
            except Exception as e:  # the reference panics with the path
                raise TxError(f"{path}: {e}") from e
        raw = rd("ratings.npy").astype(np.int8, copy=False)
        residuals = raw.astype(np.float32)
        if target_spec != "rtg":
            for term in target_spec.split(" + "):
                w, model = (term.split("*", 1) if "*" in term else ("1.0", term))
                preds = np.load(os.path.join(preds_dir, f"{model.strip()}.{name}.npy"), allow_pickle=False)
                residuals = (residuals - preds.astype(np.float32) * np.float32(float(w))).astype(np.float32)
        return cls(user_idxs=rd("user_idxs.npy").astype(np.int32, copy=False),
                   user_cnts=rd("user_cnts.npy").astype(np.int32, copy=False),
                   item_idxs=rd("item_idxs.npy").astype(np.int32),
                   item_cnts=rd("item_cnts.npy").astype(np.int32, copy=False),
                   item_years=rd("item_years.npy").astype(np.int32, copy=False),
                   raw_ratings=raw, residuals=residuals,
                   dates=rd("dates.npy").astype(np.int16, copy=False),
                   is_test=rd("is_test.npy").astype(np.int8, copy=False),
                   name=name, n_users=n_users, n_items=n_items)

    @classmethod
    def from_arrays(cls, user_idxs, item_idxs, ratings, dates, n_users, n_items, is_test=None, name="ds",
                    residuals=None) -> "Dataset":
        u = np.ascontiguousarray(user_idxs, np.int32)
        i = np.ascontiguousarray(item_idxs, np.int32)
        r = np.ascontiguousarray(ratings, np.int8)
        return cls(user_idxs=u, user_cnts=np.bincount(u, minlength=n_users).astype(np.int32),
                   item_idxs=i, item_cnts=np.bincount(i, minlength=n_items).astype(np.int32),
                   item_years=np.zeros(n_items, np.int32), raw_ratings=r,
                   residuals=(r.astype(np.float32) if residuals is None
                              else np.ascontiguousarray(residuals, np.float32)),
                   dates=np.ascontiguousarray(dates, np.int16),
                   is_test=(np.zeros(len(u), np.int8) if is_test is None else np.ascontiguousarray(is_test, np.int8)),
                   name=name, n_users=n_users, n_items=n_items)


class _RatingsRef:
    """C view of a Dataset; holds the arrays alive. masked=True hides the
    targets (MaskedDataset, src/lib.rs:421-435)."""

    def __init__(self, ds, masked: bool = False):
        self.u = np.ascontiguousarray(ds.user_idxs, np.int32)
        self.i = np.ascontiguousarray(ds.item_idxs, np.int32)
        self.d = np.ascontiguousarray(ds.dates, np.int16)
        self.cnts = np.ascontiguousarray(ds.user_cnts, np.int32)
        self.r = None if masked else np.ascontiguousarray(ds.residuals, np.float32)
        raw = getattr(ds, "raw_ratings", None)
        self.raw = None if (masked or raw is None) else np.ascontiguousarray(raw, np.int8)
        it = getattr(ds, "is_test", None)
        self.t = None if it is None else np.ascontiguousarray(it, np.int8)
        s = CRatings()
        s.n_ratings = len(self.u)
        s.n_users = int(ds.n_users)
        s.n_items = int(ds.n_items)
        s.user_idxs = self.u.ctypes.data
        s.item_idxs = self.i.ctypes.data
        s.dates = self.d.ctypes.data
        s.residuals = None if self.r is None else self.r.ctypes.data
        s.raw_ratings = None if self.raw is None else self.raw.ctypes.data
        s.is_test = None if self.t is None else self.t.ctypes.data
        s.user_cnts = self.cnts.ctypes.data
        self.s = s


# --------------------------------------------------------------------------
# TxModel
# --------------------------------------------------------------------------

class TxModel:
    """impl Regressor for TxModel (src/tx.rs:457-1116), trained on the GPU."""

    def __init__(self, tr, pr, cfg: TxConfig, schedule: Optional[dict] = None):
        self.cfg = cfg
        self.n_users = int(tr.n_users)
        self.n_items = int(tr.n_items)
        self._cc = config_struct(cfg)
        tref, pref = _RatingsRef(tr), _RatingsRef(pr, masked=True)
        h = C.c_void_p()
        _check(lib().tx_create(C.byref(self._cc), C.byref(tref.s), C.byref(pref.s), C.byref(h)), "tx_create")
        self._h = h
        self._default_split = bool(self.schedule()["split_waves"])
        if schedule is not None:
            self.set_schedule(**schedule)

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().tx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # Regressor
    def n_epochs(self) -> int:
        return int(self.cfg.n_epochs)

    def n_subscores(self) -> int:
        """4 with the neighbourhood terms (src/tx.rs:653), else 3."""
        n = C.c_int32()
        _check(lib().tx_n_subscores(self._h, C.byref(n)), "tx_n_subscores")
        return int(n.value)

    def subscore_names(self):
        return ["bias", "mf", "nsvd1", "nbr"][: self.n_subscores()]

    def set_schedule(self, mode: int = SCHED_STRATA, n_strata: int = 1, blocks: int = 0, use_graph: bool = False,
                     teams_per_block: int = 0, q_update: int = 0, y_update: int = 2,
                     user_order: int = 2, heavy_min: int = 1024, experimental: bool = False,
                     nbr_update: int = 0, split_waves: bool | None = None) -> None:
        """tx_set_schedule.  q_update / y_update / nbr_update: 0 atomic, 2 plain
        updates of the item rows, the NSVD1 rows and the neighbourhood tables;
        experimental is accepted and ignored (n_strata > 1 no longer needs it);
        split_waves: a rating wave and a streaming wave per team (blocks =
        teams); None keeps the model's default where the schedule takes them
        (one stratum, user order 0-2, plain NSVD1 rows)."""
        s = CSchedule()
        if split_waves is None:
            split_waves = (self._default_split and mode == SCHED_STRATA and n_strata == 1
                           and user_order != USER_ORDER_BUCKETS and y_update == 2)
        s.split_waves = int(bool(split_waves))
        s.user_order = int(user_order)
        s.heavy_min = int(heavy_min)
        s.update_flags = int(q_update) | (int(y_update) << 2) | (16 if experimental else 0) | (int(nbr_update) << 5)
        s.mode = int(mode)
        s.n_strata = int(n_strata)
        s.blocks = int(blocks)
        s.use_graph = int(bool(use_graph))
        s.teams_per_block = int(teams_per_block)
        _check(lib().tx_set_schedule(self._h, C.byref(s)), "tx_set_schedule")

    def fit_epoch(self, tr=None, pr=None, epoch: int = 1) -> None:
        """Regressor::fit_epoch; the training sets were fixed at creation, so
        tr / pr are accepted for interface parity only."""
        _check(lib().tx_fit_epoch(self._h, int(epoch)), "tx_fit_epoch")

    def fit_users_epoch(self, epoch: int = 1) -> None:
        """One user-only epoch with item tables and y frozen (the user pass of
        Tsvdx4pModel, src/tsvdx4p.rs:390-639, without its NSVD1 accumulators):
        folds users into fixed item tables."""
        _check(lib().tx_fit_users_epoch(self._h, int(epoch)), "tx_fit_users_epoch")

    def autotune(self, trial_epochs: int = 1) -> dict:
        """Pick teams in flight and the length-bucket bound by timing trial
        epochs from a snapshot of the parameters (restored afterwards); the
        chosen schedule is applied and returned."""
        s = CSchedule()
        _check(lib().tx_autotune(self._h, int(trial_epochs), C.byref(s)), "tx_autotune")
        return {k: getattr(s, k) for k, _ in CSchedule._fields_}

    def schedule(self) -> dict:
        """The current schedule (tx_get_schedule)."""
        s = CSchedule()
        _check(lib().tx_get_schedule(self._h, C.byref(s)), "tx_get_schedule")
        return {k: getattr(s, k) for k, _ in CSchedule._fields_}

    def autotune_trials(self) -> list:
        """The last autotune's trials (tx_get_autotune_trials): per trial the
        bucket (-1: whole schedule), stratum grid, grid size, user order,
        heavy_min, lanes and waves per user, teams, whether it was chosen,
        and its epoch time in ms."""
        n = C.c_int32(0)
        _check(lib().tx_get_autotune_trials(self._h, None, 0, C.byref(n)), "tx_get_autotune_trials")
        arr = (CAutotuneTrial * max(n.value, 1))()
        _check(lib().tx_get_autotune_trials(self._h, arr, n.value, C.byref(n)), "tx_get_autotune_trials")
        return [{k: getattr(t, k) for k, _ in CAutotuneTrial._fields_} for t in arr[:n.value]]

    def buckets(self) -> list:
        """The length buckets of user_order 3 (tx_get_buckets): bounds, users,
        ratings, lanes and waves per user, teams in flight, whether the last
        epoch ran every bucket in one launch, last launch ms."""
        arr = (CBucketInfo * N_BUCKETS)()
        _check(lib().tx_get_buckets(self._h, arr), "tx_get_buckets")
        return [{k: getattr(b, k) for k, _ in CBucketInfo._fields_} for b in arr]

    def set_bucket_shape(self, bucket: int, lanes_per_user: int, waves_per_user: int, teams: int) -> None:
        _check(lib().tx_set_bucket_shape(self._h, int(bucket), int(lanes_per_user), int(waves_per_user), int(teams)),
               "tx_set_bucket_shape")

    def set_table_precision(self, precision: str) -> None:
        """"f32" (default) or "bf16": the factor tables predictions read
        (BF16 copies of the FP32 masters, refreshed after every change)."""
        code = {"f32": 0, "bf16": 1}[precision]
        _check(lib().tx_set_table_precision(self._h, code), "tx_set_table_precision")

    def set_frozen_precision(self, precision: str) -> None:
        """"f32" (default) or "bf16": the side a user or item pass holds
        frozen (q, q_f in the user passes; p, p2, ycache in the item pass) is
        read from BF16 copies of the FP32 masters."""
        code = {"f32": 0, "bf16": 1}[precision]
        _check(lib().tx_set_frozen_precision(self._h, code), "tx_set_frozen_precision")

    def rebuild_ycache(self) -> None:
        """rebuild_ycache (src/tx.rs:391-418) from the current y tables."""
        _check(lib().tx_rebuild_ycache(self._h), "tx_rebuild_ycache")

    def last_epoch_stats(self) -> dict:
        s = CEpochStats()
        _check(lib().tx_last_epoch_stats(self._h, C.byref(s)), "tx_last_epoch_stats")
        return {k: getattr(s, k) for k, _ in CEpochStats._fields_}

    def predict(self, u, i, day):
        """Batched Regressor::predict over arrays (a scalar triple also works)."""
        scalar = np.isscalar(u)
        u = np.ascontiguousarray(np.atleast_1d(u), np.int32)
        i = np.ascontiguousarray(np.atleast_1d(i), np.int32)
        d = np.ascontiguousarray(np.atleast_1d(day), np.int16)
        out = np.empty(len(u), np.float32)
        _check(lib().tx_predict(self._h, len(u), u.ctypes.data, i.ctypes.data, d.ctypes.data, out.ctypes.data),
               "tx_predict")
        return float(out[0]) if scalar else out

    def predict_noalloc(self, u, i, day):
        """predict through the library's fixed scratch array: no device memory
        is allocated; at most NOALLOC_MAX_ROWS queries per call."""
        u = np.ascontiguousarray(np.atleast_1d(u), np.int32)
        i = np.ascontiguousarray(np.atleast_1d(i), np.int32)
        d = np.ascontiguousarray(np.atleast_1d(day), np.int16)
        out = np.empty(len(u), np.float32)
        _check(lib().tx_predict_noalloc(self._h, len(u), u.ctypes.data, i.ctypes.data, d.ctypes.data,
                                        out.ctypes.data), "tx_predict_noalloc")
        return out

    def predict_device(self, n, d_users, d_items, d_days, d_out, stream=None) -> None:
        """Device-pointer predict (ints are raw device addresses)."""
        _check(lib().tx_predict_device(self._h, int(n), d_users, d_items, d_days, d_out, stream),
               "tx_predict_device")

    def predict_subscores(self, u, i, day):
        u = np.ascontiguousarray(np.atleast_1d(u), np.int32)
        i = np.ascontiguousarray(np.atleast_1d(i), np.int32)
        d = np.ascontiguousarray(np.atleast_1d(day), np.int16)
        k = self.n_subscores()
        outs = [np.empty(len(u), np.float32) for _ in range(k)]
        ptrs = [o.ctypes.data for o in outs] + [None] * (4 - k)
        _check(lib().tx_predict_subscores(self._h, len(u), u.ctypes.data, i.ctypes.data, d.ctypes.data, *ptrs),
               "tx_predict_subscores")
        return outs

    def calc_rmse(self, ds) -> float:
        ref = _RatingsRef(ds)
        out = C.c_double()
        _check(lib().tx_calc_rmse(self._h, C.byref(ref.s), C.byref(out)), "tx_calc_rmse")
        return float(out.value)

    def param_len(self, name: str) -> int:
        n = C.c_int64()
        _check(lib().tx_param_len(self._h, PARAM_ID[name], C.byref(n)), "tx_param_len")
        return int(n.value)

    def get(self, name: str) -> np.ndarray:
        n = self.param_len(name)
        out = np.empty(n, np.float32)
        _check(lib().tx_get_param(self._h, PARAM_ID[name], out.ctypes.data, n), f"tx_get_param({name})")
        return out

    def set(self, name: str, value) -> None:
        v = np.ascontiguousarray(value, np.float32).ravel()
        _check(lib().tx_set_param(self._h, PARAM_ID[name], v.ctypes.data, len(v)), f"tx_set_param({name})")

    def set_device(self, name: str, ptr: int, rows: int, ld: int, stream=None) -> None:
        """tx_set_param_device: `rows` rows read from device address `ptr`
        with a pitch of `ld` floats, after the work queued on `stream`."""
        _check(lib().tx_set_param_device(self._h, PARAM_ID[name], ptr, rows, ld, stream),
               f"tx_set_param_device({name})")

    def save_artifacts(self, model_name: str, tr_set: str, preds_dir: str) -> Optional[str]:
        """save_artifacts (src/tx.rs:726-732): the item factors as .ifeat."""
        if not self.cfg.save_ifeat:
            return None
        path = os.path.join(preds_dir, f"{model_name}.ifeat.{tr_set}.npy")
        np.save(path, self.get("ifeat").reshape(self.n_items, self.cfg.n_feat))
        return path


# --------------------------------------------------------------------------
# evaluation helpers (src/lib.rs:463-494, 659-756)
# --------------------------------------------------------------------------

def calc_gbias(tr) -> np.float32:
    return np.float32(np.sum(np.asarray(tr.residuals, np.float64)) / len(tr.residuals))


def calc_user_offsets(tr) -> np.ndarray:
    u = np.asarray(tr.user_idxs)
    bad = np.nonzero(u[1:] < u[:-1])[0]
    if len(bad):
        t = int(bad[0]) + 1
        raise TxError(f"dataset '{getattr(tr, 'name', 'ds')}' is not user-sorted at index {t}: "
                      f"user {u[t]} follows user {u[t - 1]}")
    return np.concatenate([[0], np.cumsum(np.asarray(tr.user_cnts, np.int64))])


def get_users(n: int, shuffle: bool, seed: int, epoch: int) -> np.ndarray:
    out = np.empty(n, np.int64)
    _check(lib().tx_get_users(int(n), int(bool(shuffle)), int(seed), int(epoch), out.ctypes.data), "tx_get_users")
    return out


def rand_array(seed: int, sigma: float, n: int) -> np.ndarray:
    out = np.empty(n, np.float32)
    _check(lib().tx_rand_normal(int(seed), float(sigma), int(n), out.ctypes.data), "tx_rand_normal")
    return out


def calc_rmse(model: TxModel, ds) -> float:
    return model.calc_rmse(ds)


def gen_preds(model: TxModel, ds) -> np.ndarray:
    """gen_preds_parallel (src/lib.rs:710-722): predict + (raw - residual)."""
    p = model.predict(ds.user_idxs, ds.item_idxs, ds.dates)
    offset = np.asarray(ds.raw_ratings, np.float32) - np.asarray(ds.residuals, np.float32)
    return (p + offset).astype(np.float32)


def rmse_from_preds(preds: np.ndarray, ds) -> float:
    """rmse_from_preds (src/lib.rs:725-735)."""
    keep = np.asarray(ds.is_test) == 0
    err = (preds[keep].astype(np.float32) - np.asarray(ds.raw_ratings, np.float32)[keep]).astype(np.float64)
    return float(np.sqrt(np.mean(err * err)))
