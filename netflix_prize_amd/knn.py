"""Python mirror of KnnfModel (src/knnf.rs) over the GPU library's C ABI
(include/tsvd_knn.h): item-item kNN over a factor table (a tsvdx4 job's
`.ifeat`), cosine similarity, temporal decay, top-k weighted mean of the
user's residuals.  Used by the tests and the `*__knnf` jobs' Python path."""
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
