"""Fold-in encoder: warm-start factors for users folded into a trained model.

New users enter a trained tsvdx4 model through user-only SGD epochs with
every item table frozen (``TxModel.fit_users_epoch``; the reference's closest
analogue is the user pass of Tsvdx4pModel, src/tsvdx4p.rs:390-639).  The
reference starts such users from random factors (src/tx.rs:460-649).  This
encoder predicts a user's p_u (f values) and b_u from the ratings the user
made instead.  Each rating k of item j is a token

    (q_j in BF16, one-hot of the rating class, d_k, w_k)

with d_k the rating day normalised by item j's day statistics and
w_k = r_k - mu - b_j the residual against the model's biases.  Mean set
pooling over up to 512 sampled tokens gives

    x_u = [ mean o_k[c] q_j (c = 1..5), mean d_k q_j, mean w_k q_j, scalars ]   (7f + 64)
    y_u = W2 gelu(W1 x_u + b1) + b2                                            (p_u, b_u)

(the scalars: count, class shares, mean/std of w and d, day statistics).
The network is trained on the p_u, b_u that plain SGD produced for the
training users.  Features, normalisation and the network run in the kernels
of libtsvd_enc.so (BF16 MFMA GEMMs with FP32 accumulation and FP32 master
weights, or the FP32 debug build) and stay on the device; the warm start
writes p_u and b_u into the model from device memory.  torch only provides
device memory and the stream.
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
N_SCALAR = 64        # per-user scalar columns (15 used, zero padded to a multiple of 64)
MAX_ITEMS = 512      # rated items pooled per user
STAT_GROUPS = 256    # row groups of enc_colstats (its work buffer holds 2 x 256 x N doubles)


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
        L.enc_gemm_tile.argtypes = [i32, i32, i32, i32, i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]
        L.enc_colsum.argtypes = [vp, i32, i32, i32, vp, vp]
        L.enc_adam.argtypes = [i32, vp, vp, vp, vp, vp, i64, f32, f32, f32, f32, C.c_int32, vp]
        L.enc_cast.argtypes = [i32, vp, vp, i64, vp]
        L.enc_item_day_stats.argtypes = [vp, vp, i64, i32, vp, vp, vp, vp]
        L.enc_user_features.argtypes = [vp, i32, i32, vp, f32, vp, vp, vp, vp, vp, vp, vp, i32, i32, vp, i64, vp]
        L.enc_affine_rows.argtypes = [i32, vp, i64, vp, i32, i32, vp, vp, vp, i64, i32, vp]
        L.enc_colstats.argtypes = [vp, i32, i32, i64, vp, vp, vp, vp]
        L.enc_train_step.argtypes = [C.POINTER(_Mlp), C.POINTER(_MlpWork), vp, vp, vp, i32, f32, f32, f32, f32,
                                     C.c_int32, vp, vp]
        L.enc_last_error.restype = C.c_char_p
        for n in ("enc_gemm", "enc_gemm_tile", "enc_train_step", "enc_loss_grad", "enc_colsum", "enc_adam", "enc_cast", "enc_item_day_stats",
                  "enc_user_features", "enc_affine_rows", "enc_colstats"):
            getattr(L, n).restype = C.c_int
        _lib = L
    return _lib


def gemm_plan(M: int, N: int, K: int, epi: int = EPI_STORE, dtype: int = ENC_BF16) -> tuple[int, int, int]:
    """(tile rows, tile columns, K splits) enc_gemm takes for an M x N x K
    product ((0, 0, 1): the generic kernel)."""
    bm, bn, sp = C.c_int(0), C.c_int(0), C.c_int(0)
    _check(lib().enc_gemm_tile(dtype, epi, M, N, K, C.byref(bm), C.byref(bn), C.byref(sp)), "enc_gemm_tile")
    return bm.value, bn.value, sp.value


def gemm_tile(M: int, N: int, K: int, dtype: int = ENC_BF16) -> tuple[int, int]:
    """The C tile enc_gemm's MFMA kernel takes for an M x N x K product
    ((0, 0): the generic kernel)."""
    return gemm_plan(M, N, K, EPI_STORE, dtype)[:2]


class _Mlp(C.Structure):
    _fields_ = [("dtype", C.c_int32), ("d_in", C.c_int32), ("hid", C.c_int32), ("d_out", C.c_int32),
                ("W1", C.c_void_p), ("b1", C.c_void_p), ("W2", C.c_void_p), ("b2", C.c_void_p),
                ("W1c", C.c_void_p), ("W2c", C.c_void_p), ("m", C.c_void_p * 4), ("v", C.c_void_p * 4)]


class _MlpWork(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("Z", "Y", "dyf", "dzf", "H", "dyt", "dzt", "gW1", "gb1", "gW2", "gb2")]


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


def n_features(f: int) -> int:
    """Columns of x_u for n_feat = f."""
    return 7 * f + N_SCALAR


def _up(a, dtype=None):
    """Host array -> device tensor (the features' inputs)."""
    torch = _torch()
    a = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype))
    return torch.from_numpy(a).to(torch.device("cuda:0"))


class ItemSide:
    """The item side the features read, on the device: the model's q as BF16
    (rows padded to the kernel's pitch), item biases, mu, and each item's day
    mean and 1 / max(std, 1) over the model's training ratings."""

    def __init__(self, model: "tx.TxModel", train_ds):
        torch = _torch()
        f = model.cfg.n_feat
        vec = 1
        while 64 * vec < f:
            vec *= 2
        self.f, self.ldq, self.n_items = f, 64 * vec, train_ds.n_items
        dev = torch.device("cuda:0")
        qp = np.zeros((self.n_items, self.ldq), np.float32)
        qp[:, :f] = model.get("ifeat").reshape(-1, f)
        qf = _up(qp)
        self.q = torch.empty((self.n_items, self.ldq), dtype=torch.int16, device=dev)
        _check(lib().enc_cast(ENC_BF16, _p(qf), _p(self.q), qf.numel(), _stream()), "enc_cast q")
        self.ibias = _up(model.get("ibias"), np.float32)
        self.mu = float(model.get("gbias")[0])
        items, days = _up(train_ds.item_idxs, np.int32), _up(train_ds.dates, np.int16)
        if train_ds.n_ratings and int(train_ds.item_idxs.max()) >= self.n_items:
            raise EncoderError("item index out of range")
        work = torch.empty(3 * self.n_items, dtype=torch.int64, device=dev)
        self.day_mean = torch.empty(self.n_items, dtype=torch.float32, device=dev)
        self.day_inv = torch.empty(self.n_items, dtype=torch.float32, device=dev)
        _check(lib().enc_item_day_stats(_p(items), _p(days), train_ds.n_ratings, self.n_items, _p(work),
                                        _p(self.day_mean), _p(self.day_inv), _stream()), "enc_item_day_stats")
        _torch().cuda.current_stream().synchronize()


def user_features(side: ItemSide, ds, max_items: int = MAX_ITEMS):
    """x_u for every user of ds (its training ratings) as a device FP32
    tensor of n_features(f) columns."""
    torch = _torch()
    off = user_offsets(ds)
    if off[-1] >= 2 ** 32:
        raise EncoderError("more than 2^32 ratings")
    d_off = _up(off.astype(np.uint32).view(np.int32))
    d_items = _up(ds.item_idxs, np.int32)
    d_r = _up(ds.raw_ratings, np.int8)
    d_t = _up(ds.residuals, np.float32)
    d_days = _up(ds.dates, np.int16)
    if ds.n_ratings and int(ds.item_idxs.max()) >= side.n_items:
        raise EncoderError("item index out of range")
    n = n_features(side.f)
    out = torch.empty((ds.n_users, n), dtype=torch.float32, device=torch.device("cuda:0"))
    _check(lib().enc_user_features(_p(side.q), side.f, side.ldq, _p(side.ibias), side.mu, _p(side.day_mean),
                                   _p(side.day_inv), _p(d_off), _p(d_items), _p(d_r), _p(d_t), _p(d_days),
                                   ds.n_users, max_items, _p(out), n, _stream()), "enc_user_features")
    return out


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
        self.x_mean_d = self.x_inv_d = self.y_shift_d = self.y_scale_d = self.x_cols = None
        self.t_mean = self.t_std = None
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

    def train_step(self, X, T, colw, loss) -> None:
        """gradients() and adam_step() in one library call (enc_train_step):
        the same launches in the same order on the same stream."""
        torch = _torch()
        c = self.cfg
        B = X.shape[0]
        # the tables' addresses are part of the key: set_weights() replaces them
        ptrs = tuple(_p(a) for a in (self.W1, self.b1, self.W2, self.b2, self.W1c, self.W2c)) + \
            tuple(_p(a) for mv in self.opt for a in mv)
        key = ("step", B, ptrs)
        if self._bufs.get(key) is None:
            for k in [k for k in self._bufs if isinstance(k, tuple) and k[0] == "step"]:
                del self._bufs[k]
            net = _Mlp(self.dt, self.d_in, self.hid, self.d_out, *ptrs[:6],
                       (C.c_void_p * 4)(*ptrs[6::2]), (C.c_void_p * 4)(*ptrs[7::2]))
            f32, dt = torch.float32, self.tdt
            bufs = {"Z": ((B, self.hid), f32), "Y": ((B, self.d_out), f32), "dyf": ((B, self.d_out), f32),
                    "dzf": ((B, self.hid), f32), "H": ((B, self.hid), dt), "dyt": ((B, self.d_out), dt),
                    "dzt": ((B, self.hid), dt), "gW1": ((self.d_in, self.hid), f32), "gb1": ((self.hid,), f32),
                    "gW2": ((self.hid, self.d_out), f32), "gb2": ((self.d_out,), f32)}
            # buffers of their own ("ts_" names: forward() may resize its "Z" etc.)
            work = _MlpWork(**{n: _p(self._buf("ts_" + n, sh, d)) for n, (sh, d) in bufs.items()})
            self._bufs[key] = (net, work)
        net, work = self._bufs[key]
        self.step_t += 1
        _check(lib().enc_train_step(C.byref(net), C.byref(work), _p(X), _p(T), _p(colw), B, c.lr, c.beta1, c.beta2,
                                    c.eps, self.step_t, _p(loss), _stream()), "enc_train_step")

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
    def _dev_f32(self, A):
        """A host array or device tensor -> contiguous device FP32 matrix."""
        torch = _torch()
        if isinstance(A, torch.Tensor):
            if A.device.type != "cuda" or A.dtype != torch.float32 or A.dim() != 2:
                raise EncoderError("expected a 2-D FP32 device tensor")
            return A.contiguous()
        return torch.from_numpy(np.ascontiguousarray(A, np.float32)).to(self.dev)

    def _colstats(self, A):
        """Column mean and 1 / std (1 where std <= 1e-6) of a device matrix."""
        torch = _torch()
        M, N = A.shape
        work = torch.empty(2 * STAT_GROUPS * N, dtype=torch.float64, device=self.dev)
        mean = torch.empty(N, dtype=torch.float32, device=self.dev)
        inv = torch.empty(N, dtype=torch.float32, device=self.dev)
        _check(lib().enc_colstats(_p(A), M, N, N, _p(work), _p(mean), _p(inv), _stream()), "enc_colstats")
        return mean, inv

    def _affine(self, dt, x, ldx, rows, n, d, shift, scale, out, ldo, n_out) -> None:
        _check(lib().enc_affine_rows(dt, x, ldx, None if rows is None else _p(rows), n, d,
                                     None if shift is None else _p(shift), None if scale is None else _p(scale),
                                     out, ldo, n_out, _stream()), "enc_affine_rows")

    # training / inference ----------------------------------------------------
    def fit(self, X, Tg, log=None) -> list:
        """Train on rows X (n x d) -> targets Tg (n x n_out), host arrays or
        device FP32 tensors; returns the mean loss per epoch
        (normalised-target units).  Inputs and targets are normalised per
        column on the device."""
        torch = _torch()
        c = self.cfg
        Xd, Td = self._dev_f32(X), self._dev_f32(Tg)
        n, dx = Xd.shape
        if dx > self.d_in or Td.shape[1] != self.n_out or Td.shape[0] != n:
            raise EncoderError("fit: shapes do not match the network")
        self.x_mean_d, self.x_inv_d = self._colstats(Xd)
        t_mean_d, t_inv_d = self._colstats(Td)
        self.x_cols = dx
        # de-normalisation y * std + mean as (y - shift) * scale
        t_mean, t_inv = t_mean_d.cpu().numpy(), t_inv_d.cpu().numpy()
        t_std = (1.0 / t_inv).astype(np.float32)
        self.y_shift_d = torch.from_numpy((-t_mean * t_inv).astype(np.float32)).to(self.dev)
        self.y_scale_d = torch.from_numpy(t_std).to(self.dev)
        self.t_mean, self.t_std = t_mean, t_std
        rng = np.random.default_rng(c.seed + 1)
        nb = n // c.batch
        perm = torch.from_numpy(rng.permutation(n)[:nb * c.batch].astype(np.int64)).to(self.dev)
        rows = nb * c.batch
        Xn = torch.empty((rows, self.d_in), dtype=self.tdt, device=self.dev)
        self._affine(self.dt, _p(Xd), dx, perm, rows, dx, self.x_mean_d, self.x_inv_d, _p(Xn), self.d_in, rows)
        Tn = torch.empty((rows, self.d_out), dtype=torch.float32, device=self.dev)
        self._affine(ENC_F32, _p(Td), self.n_out, perm, rows, self.n_out, t_mean_d, t_inv_d, _p(Tn), self.d_out,
                     rows)
        colw = np.zeros(self.d_out, np.float32)
        colw[:self.n_out] = 1.0
        colw[self.n_out - 1] = c.bias_weight
        colw_d = torch.from_numpy(colw).to(self.dev)
        hist = []
        for ep in range(c.epochs):
            tot = torch.zeros(1, dtype=torch.float64, device=self.dev)
            for b in rng.permutation(nb):
                sl = slice(int(b) * c.batch, (int(b) + 1) * c.batch)
                self.train_step(Xn[sl], Tn[sl], colw_d, tot)
            hist.append(float(tot.item()) / max(nb, 1) / self.n_out)
            if log:
                log(f"encoder epoch {ep + 1}: loss {hist[-1]:.4f}")
        return hist

    def predict_device(self, X, batch: int = 65536):
        """Outputs (n x n_out, target units) for rows X as a device FP32 tensor."""
        torch = _torch()
        Xd = self._dev_f32(X)
        n, dx = Xd.shape
        if dx != self.x_cols:
            raise EncoderError(f"predict: {dx} input columns, the network was fitted on {self.x_cols}")
        out = torch.empty((n, self.n_out), dtype=torch.float32, device=self.dev)
        for s in range(0, n, batch):
            m = min(batch, n - s)
            xb = self._buf("px", (round64(m), self.d_in), self.tdt)
            self._affine(self.dt, _p(Xd) + 4 * s * dx, dx, None, m, dx, self.x_mean_d, self.x_inv_d, _p(xb),
                         self.d_in, round64(m))
            _, _, Y = self.forward(xb)
            self._affine(ENC_F32, _p(Y), self.d_out, None, m, self.n_out, self.y_shift_d, self.y_scale_d,
                         _p(out) + 4 * s * self.n_out, self.n_out, m)
        return out

    def predict(self, X) -> np.ndarray:
        """Outputs (n x n_out) in target units, as a host array."""
        return self.predict_device(X).cpu().numpy()


BF16_DENSE_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 MFMA


def gemm_benchmark(d_in: int, n_out: int, hidden: int = 512, batch: int = 4096, reps: int = 50) -> list:
    """Times the five products of one training step of the d_in -> hidden ->
    n_out network (BF16, batch rows) on the current stream with events:
    forward 1 and 2, dW2, dZ, dW1, each with its epilogue (split-K reduce
    included).  Returns per product the shape, tile, splits, microseconds
    and TFLOP/s against the dense BF16 MFMA peak."""
    torch = _torch()
    L, st = lib(), _stream()
    di, hd, do = round64(d_in), round64(hidden), round64(n_out)
    g = torch.Generator(device="cuda").manual_seed(0)
    mk = lambda r, c: torch.randint(-8000, 8000, (r, c), dtype=torch.int16, device="cuda", generator=g)
    prods = [("forward 1 (X W1, bias + GELU)", batch, hd, di, 0, 0, EPI_BIAS_GELU),
             ("forward 2 (H W2, bias)", batch, do, hd, 0, 0, EPI_BIAS),
             ("dW2 = H^T dY", hd, do, batch, 1, 0, EPI_STORE),
             ("dZ = (dY W2^T) gelu'(Z)", batch, hd, do, 0, 1, EPI_GELU_BWD),
             ("dW1 = X^T dZ", di, hd, batch, 1, 0, EPI_STORE)]
    out = []
    for name, M, N, K, ta, tb, epi in prods:
        A = mk(K, M) if ta else mk(M, K)
        B = mk(N, K) if tb else mk(K, N)
        cf = torch.empty((M, N), dtype=torch.float32, device="cuda")
        ct = torch.empty((M, N), dtype=torch.int16, device="cuda")
        bias = torch.zeros(N, dtype=torch.float32, device="cuda")
        z = torch.zeros((M, N), dtype=torch.float32, device="cuda")
        run = lambda: _check(L.enc_gemm(ENC_BF16, ta, tb, epi, _p(A), _p(B), M, N, K, A.shape[1], B.shape[1], _p(cf),
                                        _p(ct), _p(bias), _p(z), N, st), name)
        for _ in range(3):
            run()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            run()
        e1.record()
        e1.synchronize()
        us = e0.elapsed_time(e1) / reps * 1e3
        tf = 2.0 * M * N * K / us / 1e6
        bm, bn, sp = gemm_plan(M, N, K, epi)
        out.append({"product": name, "M": M, "N": N, "K": K, "tile": [bm, bn], "k_splits": sp, "us": round(us, 2),
                    "tflops": round(tf, 1), "frac_of_bf16_peak": round(tf / BF16_DENSE_PEAK_TFLOPS, 4)})
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


def warm_start(m: "tx.TxModel", enc: FoldInEncoder, X) -> float:
    """Write the encoder's p_u, b_u for rows X (device features) into m from
    device memory; returns the seconds it took."""
    t0 = time.perf_counter()
    Y = enc.predict_device(X)
    n, f = Y.shape[0], enc.n_out - 1
    st = _stream()
    m.set_device("ufeat", _p(Y), n, enc.n_out, st)
    m.set_device("ubias", _p(Y) + 4 * f, n, enc.n_out, st)
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
    side = ItemSide(base, tr_a)
    Xa = user_features(side, tr_a)
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
    Xb = user_features(side, tr_b)
    _torch().cuda.current_stream().synchronize()
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
        "encoder_start_parts_s": {"features": round(feat_s, 4), "network_and_copy": round(ws_s, 4)},
        "random_start_epochs_to_encoder_rmse": k_eq,
        "random_start_s_to_encoder_rmse": None if k_eq is None else round(sum(s_rand[1:k_eq + 1]), 4),
        "wall_s": round(time.perf_counter() - t0, 1),
    }
