"""Fold-in encoder: warm-start factors for users folded into a trained model.

New users enter a trained tsvdx4 model through user-only SGD epochs with
every item table frozen (``TxModel.fit_users_epoch``; the reference's closest
analogue is the user pass of Tsvdx4pModel, src/tsvdx4p.rs:390-639).  The
reference starts such users from random factors (src/tx.rs:460-649).  This
encoder predicts a user's p_u (f values) and b_u from the items the user
rated instead:

    x_u = [ mean_j q_j , mean_j w_uj q_j , time features ]      (2f + 64)
    y_u = W2 gelu(W1 x_u + b1) + b2                             (p_u, b_u)

where q_j are the model's item factors over up to 512 of the user's training
ratings and w_uj = r_uj - mu - b_j the rating's residual against the model's
biases.  The network is trained on the p_u, b_u that plain SGD produced for
the training users.  All arithmetic runs in the kernels of libtsvd_enc.so
(BF16 MFMA GEMMs with FP32 accumulation and FP32 master weights, or the FP32
debug build); torch only provides device memory and the stream.
"""
from __future__ import annotations

import ctypes as C
import os
import time
from dataclasses import dataclass

import numpy as np
# torch first: its HIP runtime must be the one this process loads, before
# libtsvd_tx.so / libtsvd_enc.so bind theirs (same soname, so they then share it)
import torch  # noqa: F401

from . import tx

PKG = os.path.dirname(os.path.abspath(__file__))
ENC_LIB_PATH = os.path.join(PKG, "libtsvd_enc.so")

ENC_BF16, ENC_F32 = 0, 1
EPI_STORE, EPI_BIAS, EPI_BIAS_GELU, EPI_GELU_BWD = 0, 1, 2, 3
N_TIME = 64          # time-feature columns (8 used, zero padded to a multiple of 64)
MAX_ITEMS = 512      # rated items pooled per user
DAYS = 2243.0        # day scale of the Netflix date range


class EncoderError(RuntimeError):
    pass


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(ENC_LIB_PATH):
            raise EncoderError(f"{ENC_LIB_PATH} is missing: build it with `python -m netflix_prize_amd.build`")
        L = C.CDLL(ENC_LIB_PATH)
        vp, i32, i64, f32 = C.c_void_p, C.c_int, C.c_int64, C.c_float
        L.enc_gemm.argtypes = [i32, i32, i32, i32, vp, vp, i32, i32, i32, i32, i32, vp, vp, vp, vp, i32, vp]
        L.enc_loss_grad.argtypes = [i32, vp, vp, vp, vp, vp, i32, i32, i32, vp, vp]
        L.enc_colsum.argtypes = [vp, i32, i32, i32, vp, vp]
        L.enc_adam.argtypes = [i32, vp, vp, vp, vp, vp, i64, f32, f32, f32, f32, C.c_int32, vp]
        L.enc_cast.argtypes = [i32, vp, vp, i64, vp]
        L.enc_pool.argtypes = [vp, i32, vp, vp, vp, i32, i32, vp, i32, vp]
        L.enc_last_error.restype = C.c_char_p
        for n in ("enc_gemm", "enc_loss_grad", "enc_colsum", "enc_adam", "enc_cast", "enc_pool"):
            getattr(L, n).restype = C.c_int
        _lib = L
    return _lib


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise EncoderError(f"{what} failed ({rc}): {lib().enc_last_error().decode()}")


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise EncoderError("the fold-in encoder runs on the GPU (no HIP device visible)")
    return torch


def _p(t) -> int:
    return t.data_ptr()


def _stream() -> int:
    return _torch().cuda.current_stream().cuda_stream


def round64(n: int) -> int:
    return (n + 63) // 64 * 64


# --------------------------------------------------------------------------
# features
# --------------------------------------------------------------------------

def user_offsets(ds) -> np.ndarray:
    off = np.zeros(ds.n_users + 1, np.int64)
    np.cumsum(np.bincount(ds.user_idxs, minlength=ds.n_users), out=off[1:])
    return off


def time_features(ds, w: np.ndarray) -> np.ndarray:
    """Per-user count, residual mean/spread and date statistics (users sorted)."""
    U = ds.n_users
    off = user_offsets(ds)
    cnt = np.diff(off).astype(np.float64)
    has = cnt > 0
    starts = np.minimum(off[:-1], max(len(w) - 1, 0))
    days = ds.dates.astype(np.float64)

    def seg_sum(x):
        s = np.add.reduceat(x, starts) if len(x) else np.zeros(U)
        return np.where(has, s, 0.0)

    n = np.maximum(cnt, 1.0)
    mw = seg_sum(w.astype(np.float64)) / n
    sw = np.sqrt(np.maximum(seg_sum(w.astype(np.float64) ** 2) / n - mw * mw, 0.0))
    md = seg_sum(days) / n
    dmax = np.where(has, np.maximum.reduceat(days, starts) if len(days) else 0, 0.0)
    dmin = np.where(has, np.minimum.reduceat(days, starts) if len(days) else 0, 0.0)
    out = np.zeros((U, N_TIME), np.float32)
    out[:, 0] = np.log1p(cnt) / 8.0
    out[:, 1] = mw
    out[:, 2] = sw
    out[:, 3] = md / DAYS
    out[:, 4] = (dmax - dmin) / DAYS
    out[:, 5] = dmin / DAYS
    out[:, 6] = dmax / DAYS
    out[:, 7] = 1.0
    return out


def user_features(model: "tx.TxModel", ds, max_items: int = MAX_ITEMS) -> np.ndarray:
    """x_u for every user of ds (training ratings) against model's item tables."""
    torch = _torch()
    f = model.cfg.n_feat
    q = np.ascontiguousarray(model.get("ifeat").reshape(-1, f), np.float32)
    ib = model.get("ibias")
    mu = float(model.get("gbias")[0])
    w = (ds.residuals - mu - ib[ds.item_idxs]).astype(np.float32)
    off = user_offsets(ds)
    dev = torch.device("cuda:0")
    d_q = torch.from_numpy(q).to(dev)
    d_off = torch.from_numpy(off.astype(np.uint32).view(np.int32)).to(dev)
    d_items = torch.from_numpy(np.ascontiguousarray(ds.item_idxs, np.int32)).to(dev)
    d_w = torch.from_numpy(w).to(dev)
    out = torch.empty((ds.n_users, 2 * f), dtype=torch.float32, device=dev)
    _check(lib().enc_pool(_p(d_q), f, _p(d_off), _p(d_items), _p(d_w), ds.n_users, max_items, _p(out), 2 * f,
                          _stream()), "enc_pool")
    pooled = out.cpu().numpy()
    return np.concatenate([pooled, time_features(ds, w)], axis=1)


# --------------------------------------------------------------------------
# the network
# --------------------------------------------------------------------------

@dataclass
class EncoderConfig:
    hidden: int = 512
    dtype: str = "bf16"        # "bf16" (MFMA BF16, FP32 masters) or "f32" (debug build)
    lr: float = 1e-3
    beta1: float = 0.9
    beta2: float = 0.999
    eps: float = 1e-8
    batch: int = 4096
    epochs: int = 12
    bias_weight: float = 1.0   # loss weight of the b_u output relative to one p_u element
    seed: int = 0


class FoldInEncoder:
    """2-layer MLP d_in -> hidden (GELU) -> d_out on the GPU kernels."""

    def __init__(self, d_in: int, n_out: int, cfg: EncoderConfig | None = None):
        torch = _torch()
        self.cfg = cfg or EncoderConfig()
        self.d_in, self.n_out = round64(d_in), n_out
        self.hid, self.d_out = round64(self.cfg.hidden), round64(n_out)
        self.dt = ENC_BF16 if self.cfg.dtype == "bf16" else ENC_F32
        self.tdt = torch.int16 if self.dt == ENC_BF16 else torch.float32
        self.dev = torch.device("cuda:0")
        rng = np.random.default_rng(self.cfg.seed)
        W1 = np.zeros((self.d_in, self.hid), np.float32)
        W1[:d_in] = rng.normal(0.0, np.sqrt(2.0 / d_in), (d_in, self.hid))
        W2 = np.zeros((self.hid, self.d_out), np.float32)
        W2[:, :n_out] = rng.normal(0.0, np.sqrt(1.0 / self.hid), (self.hid, n_out))
        self.set_weights(W1, np.zeros(self.hid, np.float32), W2, np.zeros(self.d_out, np.float32))
        z = lambda a: torch.zeros_like(a)
        self.opt = [(z(a), z(a)) for a in self.masters]
        self.step_t = 0
        self.x_mean = self.x_std = self.t_mean = self.t_std = None
        self._bufs = {}

    # weights -----------------------------------------------------------------
    def set_weights(self, W1, b1, W2, b2) -> None:
        torch = _torch()
        up = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(self.dev)
        self.W1, self.b1, self.W2, self.b2 = up(W1), up(b1), up(W2), up(b2)
        self.masters = [self.W1, self.b1, self.W2, self.b2]
        self.W1c = torch.empty(self.W1.shape, dtype=self.tdt, device=self.dev)
        self.W2c = torch.empty(self.W2.shape, dtype=self.tdt, device=self.dev)
        self.cast(self.W1, self.W1c)
        self.cast(self.W2, self.W2c)

    def weights(self):
        return [a.cpu().numpy() for a in self.masters]

    def cast(self, src, dst) -> None:
        _check(lib().enc_cast(self.dt, _p(src), _p(dst), src.numel(), _stream()), "enc_cast")

    def to_device_x(self, X: np.ndarray):
        """FP32 host rows (already normalised, d_in columns) -> device in dtype."""
        torch = _torch()
        Xp = np.zeros((X.shape[0], self.d_in), np.float32)
        Xp[:, :X.shape[1]] = X
        xf = torch.from_numpy(Xp).to(self.dev)
        if self.dt == ENC_F32:
            return xf
        xt = torch.empty(xf.shape, dtype=self.tdt, device=self.dev)
        self.cast(xf, xt)
        return xt

    def _buf(self, name, shape, dtype):
        torch = _torch()
        b = self._bufs.get(name)
        if b is None or tuple(b.shape) != tuple(shape) or b.dtype != dtype:
            b = torch.empty(shape, dtype=dtype, device=self.dev)
            self._bufs[name] = b
        return b

    # passes ------------------------------------------------------------------
    def forward(self, X):
        """X: device rows in dtype, B x d_in (B a multiple of 64). Returns Z, H, Y."""
        torch = _torch()
        B = X.shape[0]
        Z = self._buf("Z", (B, self.hid), torch.float32)
        H = self._buf("H", (B, self.hid), self.tdt)
        Y = self._buf("Y", (B, self.d_out), torch.float32)
        L, st = lib(), _stream()
        _check(L.enc_gemm(self.dt, 0, 0, EPI_BIAS_GELU, _p(X), _p(self.W1c), B, self.hid, self.d_in, self.d_in,
                          self.hid, _p(Z), _p(H), _p(self.b1), None, self.hid, st), "enc_gemm fwd1")
        _check(L.enc_gemm(self.dt, 0, 0, EPI_BIAS, _p(H), _p(self.W2c), B, self.d_out, self.hid, self.hid,
                          self.d_out, _p(Y), None, _p(self.b2), None, self.d_out, st), "enc_gemm fwd2")
        return Z, H, Y

    def gradients(self, X, T, colw, loss):
        """FP32 gradients gW1, gb1, gW2, gb2; the batch loss is added to the
        device FP64 scalar `loss`."""
        torch = _torch()
        B = X.shape[0]
        Z, H, Y = self.forward(X)
        L, st = lib(), _stream()
        dyf = self._buf("dyf", (B, self.d_out), torch.float32)
        dyt = self._buf("dyt", (B, self.d_out), self.tdt)
        dzf = self._buf("dzf", (B, self.hid), torch.float32)
        dzt = self._buf("dzt", (B, self.hid), self.tdt)
        gW1 = self._buf("gW1", (self.d_in, self.hid), torch.float32)
        gb1 = self._buf("gb1", (self.hid,), torch.float32)
        gW2 = self._buf("gW2", (self.hid, self.d_out), torch.float32)
        gb2 = self._buf("gb2", (self.d_out,), torch.float32)
        _check(L.enc_loss_grad(self.dt, _p(Y), _p(T), _p(colw), _p(dyf), _p(dyt), B, self.d_out, self.d_out,
                               _p(loss), st), "enc_loss_grad")
        # dW2 = H^T dY, db2 = colsum dY
        _check(L.enc_gemm(self.dt, 1, 0, EPI_STORE, _p(H), _p(dyt), self.hid, self.d_out, B, self.hid, self.d_out,
                          _p(gW2), None, None, None, self.d_out, st), "enc_gemm dW2")
        _check(L.enc_colsum(_p(dyf), B, self.d_out, self.d_out, _p(gb2), st), "enc_colsum db2")
        # dZ = (dY W2^T) * gelu'(Z)
        _check(L.enc_gemm(self.dt, 0, 1, EPI_GELU_BWD, _p(dyt), _p(self.W2c), B, self.hid, self.d_out, self.d_out,
                          self.d_out, _p(dzf), _p(dzt), None, _p(Z), self.hid, st), "enc_gemm dZ")
        # dW1 = X^T dZ, db1 = colsum dZ
        _check(L.enc_gemm(self.dt, 1, 0, EPI_STORE, _p(X), _p(dzt), self.d_in, self.hid, B, self.d_in, self.hid,
                          _p(gW1), None, None, None, self.hid, st), "enc_gemm dW1")
        _check(L.enc_colsum(_p(dzf), B, self.hid, self.hid, _p(gb1), st), "enc_colsum db1")
        return gW1, gb1, gW2, gb2

    def adam_step(self, grads) -> None:
        c = self.cfg
        self.step_t += 1
        L, st = lib(), _stream()
        copies = [self.W1c, self.b1, self.W2c, self.b2]
        dts = [self.dt, ENC_F32, self.dt, ENC_F32]
        for w, g, (m, v), wc, dt in zip(self.masters, grads, self.opt, copies, dts):
            _check(L.enc_adam(dt, _p(w), _p(g), _p(m), _p(v), _p(wc), w.numel(), c.lr, c.beta1, c.beta2, c.eps,
                              self.step_t, st), "enc_adam")

    # training / inference on host arrays -------------------------------------
    def fit(self, X: np.ndarray, Tg: np.ndarray, log=None) -> list:
        """Train on rows X (n x d) -> targets Tg (n x n_out); returns the mean
        loss per epoch (normalised-target units)."""
        torch = _torch()
        c = self.cfg
        self.x_mean = X.mean(0)
        self.x_std = np.where(X.std(0) > 1e-6, X.std(0), 1.0)
        self.t_mean = Tg.mean(0)
        self.t_std = np.where(Tg.std(0) > 1e-6, Tg.std(0), 1.0)
        rng = np.random.default_rng(c.seed + 1)
        perm = rng.permutation(X.shape[0])
        nb = X.shape[0] // c.batch
        perm = perm[:nb * c.batch]
        Xd = self.to_device_x(((X[perm] - self.x_mean) / self.x_std).astype(np.float32))
        Tp = np.zeros((len(perm), self.d_out), np.float32)
        Tp[:, :self.n_out] = (Tg[perm] - self.t_mean) / self.t_std
        Td = torch.from_numpy(Tp).to(self.dev)
        colw = np.zeros(self.d_out, np.float32)
        colw[:self.n_out] = 1.0
        colw[self.n_out - 1] = c.bias_weight
        colw_d = torch.from_numpy(colw).to(self.dev)
        hist = []
        for ep in range(c.epochs):
            tot = torch.zeros(1, dtype=torch.float64, device=self.dev)
            for b in rng.permutation(nb):
                sl = slice(int(b) * c.batch, (int(b) + 1) * c.batch)
                self.adam_step(self.gradients(Xd[sl], Td[sl], colw_d, tot))
            hist.append(float(tot.item()) / max(nb, 1) / self.n_out)
            if log:
                log(f"encoder epoch {ep + 1}: loss {hist[-1]:.4f}")
        return hist

    def predict(self, X: np.ndarray) -> np.ndarray:
        """Outputs (n x n_out) in target units for host rows X."""
        n = X.shape[0]
        out = np.zeros((n, self.n_out), np.float32)
        bs = self.cfg.batch
        for s in range(0, n, bs):
            rows = X[s:s + bs]
            m = rows.shape[0]
            pad = np.zeros((round64(m), X.shape[1]), np.float32)
            pad[:m] = (rows - self.x_mean) / self.x_std
            _, _, Y = self.forward(self.to_device_x(pad))
            out[s:s + m] = Y[:m, :self.n_out].cpu().numpy() * self.t_std + self.t_mean
        return out


# --------------------------------------------------------------------------
# new-user evaluation
# --------------------------------------------------------------------------

ITEM_TABLES = ("ifeat", "ibias", "bit_bin", "ibias_freq", "ifeat_freq", "yfeat", "ybias", "gbias")


def split_users(tr, pr, n_first: int):
    """Users [0, n_first) and [n_first, U) as two datasets each (ids re-based)."""
    out = []
    for lo, hi in ((0, n_first), (n_first, tr.n_users)):
        parts = []
        for ds in (tr, pr):
            sel = (ds.user_idxs >= lo) & (ds.user_idxs < hi)
            parts.append(tx.Dataset.from_arrays(ds.user_idxs[sel] - lo, ds.item_idxs[sel], ds.raw_ratings[sel],
                                                ds.dates[sel], hi - lo, ds.n_items, is_test=ds.is_test[sel],
                                                name=ds.name, residuals=ds.residuals[sel]))
        out.append(tuple(parts))
    return out


def fold_in_model(base: "tx.TxModel", tr, pr, cfg) -> "tx.TxModel":
    """A model over new users (tr, pr) carrying base's item tables; its user
    parameters start from the reference's random initialisation."""
    m = tx.TxModel(tr, pr, cfg)
    for n in ITEM_TABLES:
        m.set(n, base.get(n))
    m.rebuild_ycache()
    return m


def fold_in_curve(m: "tx.TxModel", pr, epochs: int) -> tuple[list, list]:
    """Holdout RMSE after 0..epochs user-only epochs, and the GPU seconds of each."""
    rmse, secs = [m.calc_rmse(pr)], [0.0]
    for e in range(1, epochs + 1):
        m.fit_users_epoch(e)
        secs.append(m.last_epoch_stats()["ms_total"] / 1000.0)
        rmse.append(m.calc_rmse(pr))
    return rmse, secs


def sgd_targets(model: "tx.TxModel") -> np.ndarray:
    """The encoder's training targets: each user's p_u (f values) and b_u."""
    f = model.cfg.n_feat
    return np.concatenate([model.get("ufeat").reshape(-1, f), model.get("ubias")[:, None]], axis=1)


def warm_start(m: "tx.TxModel", enc: FoldInEncoder, X: np.ndarray) -> float:
    """Write the encoder's p_u, b_u into m; returns the seconds it took."""
    t0 = time.perf_counter()
    Y = enc.predict(X)
    f = Y.shape[1] - 1
    m.set("ufeat", Y[:, :f])
    m.set("ubias", Y[:, f])
    m.rebuild_ycache()
    return time.perf_counter() - t0


def evaluate_new_users(tr, pr, cfg, n_train: int, epochs: int, fold_epochs: int = 5, random_epochs: int = 20,
                       enc_cfg: EncoderConfig | None = None, log=None) -> dict:
    """Users [0, n_train) train the model (default parallel schedule, `epochs`
    epochs) and the encoder (on that model's p_u, b_u); users [n_train, U) are
    new: folded in with user-only epochs from the reference's random start
    (0..random_epochs epochs) and from the encoder's warm start
    (0..fold_epochs).  Holdout RMSE of the new users along both, and the time
    each start takes."""
    log = log or (lambda s: None)
    (tr_a, pr_a), (tr_b, pr_b) = split_users(tr, pr, n_train)
    t0 = time.perf_counter()
    base = tx.TxModel(tr_a, pr_a, cfg)
    gpu_s = 0.0
    for e in range(1, epochs + 1):
        base.fit_epoch(epoch=e)
        gpu_s += base.last_epoch_stats()["ms_total"] / 1000.0
    log(f"base model: {epochs} epochs on {n_train} users, {gpu_s:.2f} s GPU, holdout {base.calc_rmse(pr_a):.5f}")
    Xa = user_features(base, tr_a)
    Ta = sgd_targets(base)
    enc = FoldInEncoder(Xa.shape[1], Ta.shape[1], enc_cfg)
    t1 = time.perf_counter()
    hist = enc.fit(Xa, Ta, log=log)
    _torch().cuda.synchronize()
    fit_s = time.perf_counter() - t1
    rand = fold_in_model(base, tr_b, pr_b, cfg)
    r_rand, s_rand = fold_in_curve(rand, pr_b, max(random_epochs, fold_epochs))
    rand.close()
    warm = fold_in_model(base, tr_b, pr_b, cfg)
    t2 = time.perf_counter()
    Xb = user_features(base, tr_b)
    feat_s = time.perf_counter() - t2
    ws_s = warm_start(warm, enc, Xb)
    r_warm, s_warm = fold_in_curve(warm, pr_b, fold_epochs)
    warm.close()
    base.close()
    # user-only epochs the random start needs to reach the warm start's RMSE
    # before any epoch (within 1e-3), and the GPU seconds those epochs take
    target = r_warm[0] + 1e-3
    k_eq = next((k for k, r in enumerate(r_rand) if r <= target), None)
    return {
        "train_users": n_train, "new_users": tr_b.n_users, "base_epochs": epochs,
        "encoder": {"dtype": enc.cfg.dtype, "layers": [enc.d_in, enc.hid, enc.d_out], "outputs": enc.n_out,
                    "epochs": enc.cfg.epochs, "batch": enc.cfg.batch, "loss_first_last": [round(hist[0], 4), round(hist[-1], 4)],
                    "train_s": round(fit_s, 2)},
        "holdout_rmse_random_start": [round(x, 5) for x in r_rand],
        "holdout_rmse_encoder_start": [round(x, 5) for x in r_warm],
        "fold_epoch_s": [round(x, 4) for x in s_rand[1:]],
        "encoder_start_s": round(feat_s + ws_s, 3),
        "random_start_epochs_to_encoder_rmse": k_eq,
        "random_start_s_to_encoder_rmse": None if k_eq is None else round(sum(s_rand[1:k_eq + 1]), 4),
        "wall_s": round(time.perf_counter() - t0, 1),
    }
