"""Python mirror of KnnfModel (src/knnf.rs), KnnsModel (src/knns.rs) and
Knn3Model (src/knn3.rs) over the GPU library's C ABI (include/tsvd_knn.h):
item-item kNN over a factor table (a tsvdx4 job's `.ifeat`) or co-occurrence
support with temporal decay and a top-k weighted mean of the user's
residuals, and item-item NNLS regression weights.  Used by the tests and the
bench."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import tx


@dataclass
class KnnfConfig:
    """KnnfConfig (src/knnf.rs:8-27) without the factor path; defaults of
    KnnfConfig::with_factors (src/knnf.rs:23-26)."""
    k: int = 25
    scaling: float = 1.0
    tau: float = 0.025


class _CKnnf(C.Structure):
    _fields_ = [("k", C.c_int32), ("scaling", C.c_float), ("tau", C.c_float)]


def _lib():
    L = tx.lib()
    if not getattr(L, "_knnf_bound", False):
        vp = C.c_void_p
        L.knnf_create.argtypes = [vp, vp, C.c_int32, C.c_int32, vp, C.POINTER(vp)]
        L.knnf_predict.argtypes = [vp, C.c_int64, vp, vp, vp, vp]
        L.knnf_similarity.argtypes = [vp, C.c_int64, vp, vp, vp]
        L.knnf_destroy.argtypes = [vp]
        for name in ("knnf_create", "knnf_predict", "knnf_similarity", "knnf_destroy"):
            getattr(L, name).restype = C.c_int
        L._knnf_bound = True
    return L


class KnnfModel:
    """KnnfModel::new_with_factors (src/knnf.rs:60-90) on the GPU; no SGD epochs
    (Regressor::n_epochs() is 0)."""

    n_epochs = 0

    def __init__(self, tr, factors: np.ndarray, cfg: KnnfConfig = KnnfConfig()):
        F = np.ascontiguousarray(factors, np.float32)
        if F.ndim != 2:
            raise ValueError("factors must be n_items x n_feat")
        self.cfg = cfg
        self._tr = tx._RatingsRef(tr)
        c = _CKnnf(int(cfg.k), float(cfg.scaling), float(cfg.tau))
        h = C.c_void_p()
        tx._check(_lib().knnf_create(C.byref(c), F.ctypes.data, F.shape[0], F.shape[1], C.byref(self._tr.s),
                                     C.byref(h)), "knnf_create")
        self._h = h

    def predict(self, users, items, days) -> np.ndarray:
        u = np.ascontiguousarray(users, np.int32)
        i = np.ascontiguousarray(items, np.int32)
        d = np.ascontiguousarray(days, np.int16)
        out = np.empty(len(u), np.float32)
        tx._check(_lib().knnf_predict(self._h, len(u), u.ctypes.data, i.ctypes.data, d.ctypes.data, out.ctypes.data),
                  "knnf_predict")
        return out

    def similarity(self, i, j) -> np.ndarray:
        a = np.ascontiguousarray(i, np.int32)
        b = np.ascontiguousarray(j, np.int32)
        out = np.empty(len(a), np.float32)
        tx._check(_lib().knnf_similarity(self._h, len(a), a.ctypes.data, b.ctypes.data, out.ctypes.data),
                  "knnf_similarity")
        return out

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib().knnf_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class KnnsConfig:
    """KnnsConfig (src/knns.rs:19-43) with SuppSource::Compute; defaults of
    KnnsConfig::default (src/knns.rs:35-43)."""
    k: int = 25
    shrinkage: float = 100.0
    scaling: float = 1.0
    tau: float = 0.025


class _CKnns(C.Structure):
    _fields_ = [("k", C.c_int32), ("shrinkage", C.c_float), ("scaling", C.c_float), ("tau", C.c_float)]


class KnnsModel(KnnfModel):
    """KnnsModel::new with SuppSource::Compute (src/knns.rs:78-100) on the GPU:
    co-occurrence support of the training and held-out items, shrunk-lift
    similarity; predictions as KnnfModel (the same rule)."""

    def __init__(self, tr, pr, cfg: KnnsConfig = KnnsConfig()):
        L = _lib()
        if not getattr(L, "_knns_bound", False):
            vp = C.c_void_p
            L.knns_create.argtypes = [vp, vp, vp, C.POINTER(vp)]
            L.knns_create.restype = C.c_int
            L._knns_bound = True
        self.cfg = cfg
        self._tr = tx._RatingsRef(tr)
        self._pr = tx._RatingsRef(pr, masked=True)
        c = _CKnns(int(cfg.k), float(cfg.shrinkage), float(cfg.scaling), float(cfg.tau))
        h = C.c_void_p()
        tx._check(L.knns_create(C.byref(c), C.byref(self._tr.s), C.byref(self._pr.s), C.byref(h)), "knns_create")
        self._h = h


@dataclass
class KnndConfig:
    """KnnConfig (src/knn.rs:19-57) in the `*__knn-d` form: SimType::Support
    over a loaded stat0, use_stat1; the job's values (src/bin/tsvdx4-new.rs:590-610)."""
    k: int = 15
    shrinkage: float = 100.0
    scaling: float = 3.5
    tau: float = 0.0


class KnndModel(KnnfModel):
    """KnnModel (src/knn.rs) on the GPU over stat tables given as n_items x
    n_items arrays: support similarity from stat0, and stat1 / (stat0 +
    shrinkage) averaged in place of the residual; the KnnfModel top-k rule."""

    def __init__(self, tr, stat0: np.ndarray, stat1: np.ndarray, cfg: KnndConfig = KnndConfig()):
        L = _lib()
        if not getattr(L, "_knnd_bound", False):
            vp = C.c_void_p
            L.knnd_create.argtypes = [vp, vp, vp, C.c_int32, vp, C.POINTER(vp)]
            L.knnd_create.restype = C.c_int
            L._knnd_bound = True
        S0 = np.ascontiguousarray(stat0, np.float32)
        S1 = np.ascontiguousarray(stat1, np.float32)
        if S0.ndim != 2 or S0.shape[0] != S0.shape[1] or S1.shape != S0.shape:
            raise ValueError("stat0 and stat1 must be n_items x n_items")
        self.cfg = cfg
        self._tr = tx._RatingsRef(tr)
        c = _CKnns(int(cfg.k), float(cfg.shrinkage), float(cfg.scaling), float(cfg.tau))
        h = C.c_void_p()
        tx._check(L.knnd_create(C.byref(c), S0.ctypes.data, S1.ctypes.data, S0.shape[0], C.byref(self._tr.s),
                                C.byref(h)), "knnd_create")
        self._h = h


@dataclass
class Knn3Config:
    """Knn3Config (src/knn3.rs:7-31); defaults of Knn3Config::default."""
    threshold: float = 0.25
    k_min: int = 10
    k_max: int = 60
    shrinkage: float = 20000.0
    reg: float = 0.01
    x: float = 0.8
    bl_reg_m: float = 25.0
    bl_reg_u: float = 10.0


class _CKnn3(C.Structure):
    _fields_ = [("threshold", C.c_float), ("k_min", C.c_int32), ("k_max", C.c_int32), ("shrinkage", C.c_float),
                ("reg", C.c_float), ("x", C.c_float), ("bl_reg_m", C.c_float), ("bl_reg_u", C.c_float)]


def _lib3():
    L = tx.lib()
    if not getattr(L, "_knn3_bound", False):
        vp = C.c_void_p
        L.knn3_create.argtypes = [vp, vp, C.POINTER(vp)]
        L.knn3_predict.argtypes = [vp, C.c_int64, vp, vp, vp, vp]
        L.knn3_weights.argtypes = [vp, vp, vp, vp]
        L.knn3_candidates.argtypes = [vp, vp, vp]
        L.knn3_baseline.argtypes = [vp, vp, vp, vp]
        L.knn3_destroy.argtypes = [vp]
        for name in ("knn3_create", "knn3_predict", "knn3_weights", "knn3_candidates", "knn3_baseline",
                     "knn3_destroy"):
            getattr(L, name).restype = C.c_int
        L._knn3_bound = True
    return L


class Knn3Model:
    """Knn3Model::new (src/knn3.rs:350-362) on the GPU: per item, regression
    weights on its most related items (NNLS over the item's users); no SGD
    epochs.  predict(u, i) sums weight x residual over the neighbours u rated
    (src/knn3.rs:379-389)."""

    n_epochs = 0

    def __init__(self, tr, cfg: Knn3Config = Knn3Config()):
        self.cfg = cfg
        self.n_items, self.n_users = int(tr.n_items), int(tr.n_users)
        self._tr = tx._RatingsRef(tr)
        c = _CKnn3(float(cfg.threshold), int(cfg.k_min), int(cfg.k_max), float(cfg.shrinkage), float(cfg.reg),
                   float(cfg.x), float(cfg.bl_reg_m), float(cfg.bl_reg_u))
        h = C.c_void_p()
        tx._check(_lib3().knn3_create(C.byref(c), C.byref(self._tr.s), C.byref(h)), "knn3_create")
        self._h = h

    def predict(self, users, items, days=None) -> np.ndarray:
        u = np.ascontiguousarray(users, np.int32)
        i = np.ascontiguousarray(items, np.int32)
        out = np.empty(len(u), np.float32)
        tx._check(_lib3().knn3_predict(self._h, len(u), u.ctypes.data, i.ctypes.data, None, out.ctypes.data),
                  "knn3_predict")
        return out

    def neighbors(self) -> list:
        """[(item, weight), ...] per item, in candidate order."""
        cnt = np.empty(self.n_items, np.int32)
        it = np.empty((self.n_items, 64), np.int32)
        w = np.empty((self.n_items, 64), np.float32)
        tx._check(_lib3().knn3_weights(self._h, cnt.ctypes.data, it.ctypes.data, w.ctypes.data), "knn3_weights")
        return [[(int(it[m, a]), float(w[m, a])) for a in range(cnt[m])] for m in range(self.n_items)]

    def candidates(self) -> list:
        cnt = np.empty(self.n_items, np.int32)
        it = np.empty((self.n_items, 64), np.int32)
        tx._check(_lib3().knn3_candidates(self._h, cnt.ctypes.data, it.ctypes.data), "knn3_candidates")
        return [it[m, :cnt[m]].copy() for m in range(self.n_items)]

    def baseline(self):
        mu = C.c_float()
        bm = np.empty(self.n_items, np.float32)
        bu = np.empty(self.n_users, np.float32)
        tx._check(_lib3().knn3_baseline(self._h, C.byref(mu), bm.ctypes.data, bu.ctypes.data), "knn3_baseline")
        return np.float32(mu.value), bm, bu

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib3().knn3_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
