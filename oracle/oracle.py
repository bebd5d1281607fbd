"""ORACLE binding — test infrastructure only.

ctypes wrapper over ``oracle/libtxoracle.so`` (the plain-C restatement of the
reference's TxModel / Tsvdx4pModel in ``oracle/tx_oracle.c``).  Only the tests,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import this
module, and only as the checker; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libtxoracle.so")

# Parameter ids (same numbering as the product's C-ABI).
PARAMS = ["ubias", "ibias", "ufeat", "ufeat2", "ifeat", "yfeat", "ycache", "yfeat_day",
          "tu_mean", "alpha_u", "but_bin", "bit_bin", "ybias", "ycache_bias", "ubias_day",
          "ufeat_day", "ycache_day", "ibias_freq", "ifeat_freq", "cu", "cut", "gbias",
          "thresholds", "w", "c", "w_day", "c_day"]
PARAM_ID = {n: k for k, n in enumerate(PARAMS)}


class _Dataset(C.Structure):
    _fields_ = [("n_ratings", C.c_int64), ("n_users", C.c_int32), ("n_items", C.c_int32),
                ("user_idxs", C.c_void_p), ("item_idxs", C.c_void_p), ("dates", C.c_void_p),
                ("residuals", C.c_void_p), ("raw_ratings", C.c_void_p), ("is_test", C.c_void_p),
                ("user_cnts", C.c_void_p)]


_F32_FIELDS = ["beta", "lr_u", "lr_ud", "lr_u2", "lr_ub", "lr_ubd", "lr_i", "lr_ib", "lr_y", "lr_yb",
               "lr_yd", "lr_tu", "lr_ti", "lr_ta", "lr_ibf", "lr_iqf", "lr_cu", "reg_iqf", "reg_cu",
               "reg_u", "reg_u2", "reg_ud", "reg_i", "reg_y", "reg_yd", "sigma_iqf", "sigma_u",
               "sigma_i", "sigma_y", "sigma_yd", "lr_w", "lr_c", "reg_w", "reg_c", "lr_w_day",
               "lr_c_day", "reg_w_day", "reg_c_day", "w_bias", "w_factor", "w_nbr", "lambda1",
               "lambda2"]


class _Config(C.Structure):
    _fields_ = ([("seed", C.c_uint64)]
                + [(n, C.c_int32) for n in ["n_feat", "n_epochs", "n_time_bins", "n_freq_bins",
                                            "reset_u_epoch", "max_neighbors", "shuffle_users",
                                            "sum_err_bug", "low_memory", "full_su", "save_ifeat",
                                            "has_ordinal"]]
                + [(n, C.c_float) for n in _F32_FIELDS]
                + [("th_init", C.c_float * 4), ("th_gap", C.c_float), ("lr_t", C.c_float),
                   ("reg_t", C.c_float), ("nsvd_norm_exp", C.c_float), ("n_threads", C.c_int32),
                   ("seq_epochs", C.c_int32 * 5)])


def _get(cfg, name, default=None):
    if isinstance(cfg, dict):
        return cfg.get(name, default)
    return getattr(cfg, name, default)


def make_config(cfg) -> _Config:
    """Build the C struct from a dict or an object with TxConfig field names."""
    c = _Config()
    c.seed = int(_get(cfg, "seed"))
    for n in ["n_feat", "n_epochs", "n_time_bins", "n_freq_bins", "reset_u_epoch"]:
        setattr(c, n, int(_get(cfg, n)))
    c.max_neighbors = int(_get(cfg, "max_neighbors", 0) or 0)
    c.shuffle_users = int(bool(_get(cfg, "shuffle_users", True)))
    c.sum_err_bug = int(bool(_get(cfg, "sum_err_bug", False)))
    c.low_memory = int(bool(_get(cfg, "low_memory", True)))
    c.full_su = int(bool(_get(cfg, "full_su", True)))
    c.save_ifeat = int(bool(_get(cfg, "save_ifeat", False)))
    for n in _F32_FIELDS:
        default = 1.0 if n in ("w_bias", "w_factor", "w_nbr") else 0.0
        v = _get(cfg, n, default)
        setattr(c, n, float(default if v is None else v))
    oh = _get(cfg, "ordinal_head", None)
    c.has_ordinal = int(oh is not None)
    if oh is not None:
        th = _get(oh, "th_init")
        for k in range(4):
            c.th_init[k] = float(th[k])
        c.th_gap = float(_get(oh, "th_gap"))
        c.lr_t = float(_get(oh, "lr_t"))
        c.reg_t = float(_get(oh, "reg_t"))
    c.nsvd_norm_exp = float(_get(cfg, "nsvd_norm_exp", 0.0) or 0.0)
    c.n_threads = int(_get(cfg, "n_threads", 0) or 0)
    se = _get(cfg, "seq_epochs", None) or [0, 0, 0, 0, 0]
    for k in range(5):
        c.seq_epochs[k] = int(se[k])
    return c


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc); returns the library path."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        vp, i64, i32, f32 = C.c_void_p, C.c_int64, C.c_int32, C.c_float
        L.or_new.restype = vp
        L.or_new.argtypes = [vp, vp, vp]
        L.or_error.restype = C.c_char_p
        L.or_error.argtypes = [vp]
        L.or_free.argtypes = [vp]
        L.or_fit_epoch.argtypes = [vp, vp, vp, i64]
        L.or_fit_epoch_order.argtypes = [vp, vp, vp, i64, vp]
        L.or4p_fit_epoch.argtypes = [vp, vp, vp, i64, C.c_int]
        L.or_fit_users_epoch.argtypes = [vp, vp, vp, i64]
        L.or_rebuild_ycache.argtypes = [vp, vp, vp]
        L.or_predict.argtypes = [vp, i64, vp, vp, vp, vp]
        L.or_predict_subscores.argtypes = [vp, i64, vp, vp, vp, vp, vp, vp, vp]
        L.or_predict_f64.argtypes = [vp, i64, vp, vp, vp, vp]
        L.or_calc_rmse.restype = C.c_double
        L.or_calc_rmse.argtypes = [vp, vp]
        L.or_param_len.restype = i64
        L.or_param_len.argtypes = [vp, C.c_int]
        L.or_get_param.argtypes = [vp, C.c_int, vp, i64]
        L.or_set_param.argtypes = [vp, C.c_int, vp, i64]
        L.or_day_range.restype = i32
        L.or_day_range.argtypes = [vp]
        L.or_n_ud.restype = i64
        L.or_n_ud.argtypes = [vp]
        L.or_sud_arrays.argtypes = [vp, vp, vp, vp]
        L.or_get_users.argtypes = [i64, C.c_int, C.c_uint64, i64, vp]
        L.or_normal_fill.argtypes = [C.c_uint64, f32, i64, vp]
        L.or_rng_words.argtypes = [C.c_uint64, C.c_int, i64, vp]
        L.or_chacha_block.argtypes = [vp, C.c_uint64, C.c_int, vp]
        L.or_zig_tables.argtypes = [vp, vp]
        L.or_dev.restype = f32
        L.or_dev.argtypes = [f32, i32, f32]
        L.or_freq_bin.restype = C.c_int
        L.or_freq_bin.argtypes = [C.c_uint32, C.c_int]
        L.or_set_plant.argtypes = [C.c_int]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


class OracleData:
    """Keeps numpy arrays alive behind the C dataset struct."""

    def __init__(self, ds):
        self.user_idxs = np.ascontiguousarray(ds.user_idxs, dtype=np.int32)
        self.item_idxs = np.ascontiguousarray(ds.item_idxs, dtype=np.int32)
        self.dates = np.ascontiguousarray(ds.dates, dtype=np.int16)
        res = getattr(ds, "residuals", None)
        self.residuals = None if res is None else np.ascontiguousarray(res, dtype=np.float32)
        raw = getattr(ds, "raw_ratings", None)
        self.raw_ratings = None if raw is None else np.ascontiguousarray(raw, dtype=np.int8)
        it = getattr(ds, "is_test", None)
        self.is_test = (np.zeros(len(self.user_idxs), np.int8) if it is None
                        else np.ascontiguousarray(it, dtype=np.int8))
        self.user_cnts = np.ascontiguousarray(ds.user_cnts, dtype=np.int32)
        self.n_users = int(ds.n_users)
        self.n_items = int(ds.n_items)
        s = _Dataset()
        s.n_ratings = len(self.user_idxs)
        s.n_users = self.n_users
        s.n_items = self.n_items
        s.user_idxs = _ptr(self.user_idxs)
        s.item_idxs = _ptr(self.item_idxs)
        s.dates = _ptr(self.dates)
        s.residuals = _ptr(self.residuals)
        s.raw_ratings = _ptr(self.raw_ratings)
        s.is_test = _ptr(self.is_test)
        s.user_cnts = _ptr(self.user_cnts)
        self.struct = s

    @property
    def ref(self):
        return C.byref(self.struct)


class OracleModel:
    """Serial TxModel restatement (and the Tsvdx4pModel epoch) on the CPU."""

    def __init__(self, cfg, tr, pr):
        self.cfg = cfg
        self._c = make_config(cfg)
        self.tr = tr if isinstance(tr, OracleData) else OracleData(tr)
        self.pr = pr if isinstance(pr, OracleData) else OracleData(pr)
        self._m = lib().or_new(C.byref(self._c), self.tr.ref, self.pr.ref)
        err = lib().or_error(self._m).decode()
        if err:
            lib().or_free(self._m)
            self._m = None
            raise RuntimeError(err)

    def __del__(self):
        if getattr(self, "_m", None):
            lib().or_free(self._m)
            self._m = None

    def fit_epoch(self, epoch: int, users=None):
        if users is None:
            lib().or_fit_epoch(self._m, self.tr.ref, self.pr.ref, int(epoch))
        else:
            u = np.ascontiguousarray(users, dtype=np.int64)
            lib().or_fit_epoch_order(self._m, self.tr.ref, self.pr.ref, int(epoch), u.ctypes.data)

    def fit_users_epoch(self, epoch: int):
        """The tsvdx4p user pass without its NSVD1 accumulators, then the ycache
        rebuild (users folded into fixed item tables)."""
        lib().or_fit_users_epoch(self._m, self.tr.ref, self.pr.ref, int(epoch))

    def fit_epoch_p(self, epoch: int, workers: int = 1):
        """Tsvdx4pModel::fit_epoch with `workers` OS threads."""
        lib().or4p_fit_epoch(self._m, self.tr.ref, self.pr.ref, int(epoch), int(workers))

    def rebuild_ycache(self):
        lib().or_rebuild_ycache(self._m, self.tr.ref, self.pr.ref)

    def predict(self, u, i, d):
        u, i, d = (np.ascontiguousarray(u, np.int32), np.ascontiguousarray(i, np.int32),
                   np.ascontiguousarray(d, np.int16))
        out = np.empty(len(u), np.float32)
        lib().or_predict(self._m, len(u), u.ctypes.data, i.ctypes.data, d.ctypes.data, out.ctypes.data)
        return out

    def predict_f64(self, u, i, d):
        u, i, d = (np.ascontiguousarray(u, np.int32), np.ascontiguousarray(i, np.int32),
                   np.ascontiguousarray(d, np.int16))
        out = np.empty(len(u), np.float64)
        lib().or_predict_f64(self._m, len(u), u.ctypes.data, i.ctypes.data, d.ctypes.data, out.ctypes.data)
        return out

    def predict_subscores(self, u, i, d):
        u, i, d = (np.ascontiguousarray(u, np.int32), np.ascontiguousarray(i, np.int32),
                   np.ascontiguousarray(d, np.int16))
        outs = [np.empty(len(u), np.float32) for _ in range(4)]
        lib().or_predict_subscores(self._m, len(u), u.ctypes.data, i.ctypes.data, d.ctypes.data,
                                   *[o.ctypes.data for o in outs])
        return outs if self.has_nbr else outs[:3]

    @property
    def has_nbr(self) -> bool:
        """The neighbourhood terms are built (lr_w > 0 or lr_c > 0)."""
        return (float(_get(self.cfg, "lr_w", 0.0) or 0.0) > 0.0) or (float(_get(self.cfg, "lr_c", 0.0) or 0.0) > 0.0)

    def calc_rmse(self, ds=None) -> float:
        d = self.pr if ds is None else (ds if isinstance(ds, OracleData) else OracleData(ds))
        return float(lib().or_calc_rmse(self._m, d.ref))

    def get(self, name: str) -> np.ndarray:
        pid = PARAM_ID[name]
        n = lib().or_param_len(self._m, pid)
        out = np.empty(max(n, 0), np.float32)
        if lib().or_get_param(self._m, pid, out.ctypes.data, n) != 0:
            raise KeyError(name)
        return out

    def set(self, name: str, value) -> None:
        pid = PARAM_ID[name]
        v = np.ascontiguousarray(value, dtype=np.float32).ravel()
        if lib().or_set_param(self._m, pid, v.ctypes.data, len(v)) != 0:
            raise ValueError(f"bad size for {name}")

    @property
    def day_range(self) -> int:
        return int(lib().or_day_range(self._m))

    def sparse_ud(self):
        n = lib().or_n_ud(self._m)
        starts = np.empty(self.tr.n_users + 1, np.int64)
        dates = np.empty(n, np.int16)
        cnts = np.empty(n, np.uint32)
        lib().or_sud_arrays(self._m, starts.ctypes.data, dates.ctypes.data, cnts.ctypes.data)
        return starts, dates, cnts


class planted:
    """Context manager that plants mistake `k` in the oracle (see or_set_plant
    in tx_oracle.c); for the tests that show an invariant catches it."""

    def __init__(self, k: int):
        self.k = int(k)

    def __enter__(self):
        lib().or_set_plant(self.k)
        return self

    def __exit__(self, *exc):
        lib().or_set_plant(0)
        return False


def get_users(n, shuffle, seed, epoch):
    out = np.empty(n, np.int64)
    lib().or_get_users(int(n), int(bool(shuffle)), int(seed), int(epoch), out.ctypes.data)
    return out


def normal_fill(seed, sigma, n):
    out = np.empty(n, np.float32)
    lib().or_normal_fill(int(seed), float(sigma), int(n), out.ctypes.data)
    return out


def rng_words(seed, n, rounds=12):
    out = np.empty(n, np.uint32)
    lib().or_rng_words(int(seed), int(rounds), int(n), out.ctypes.data)
    return out


def chacha_block(key_words, counter, rounds):
    key = np.ascontiguousarray(key_words, dtype=np.uint32)
    out = np.empty(16, np.uint32)
    lib().or_chacha_block(key.ctypes.data, int(counter), int(rounds), out.ctypes.data)
    return out


def zig_tables():
    x = np.empty(257, np.float64)
    f = np.empty(257, np.float64)
    lib().or_zig_tables(x.ctypes.data, f.ctypes.data)
    return x, f
