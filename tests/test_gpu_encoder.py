"""Fold-in encoder kernels (libtsvd_enc.so) against the FP64 oracle, and what
the trained encoder gains over random-start fold-in (needs an MI355X)."""
import numpy as np
import pytest

from netflix_prize_amd import encoder as E
from netflix_prize_amd import tx
from oracle import encoder_ref as R
from tests.fixtures import lm_config, synthetic

pytestmark = pytest.mark.gpu


def bf16_round(a):
    """FP32 -> nearest BF16 (ties to even), back in FP32."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def _dev(a, dtype):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()
    if dtype == E.ENC_F32:
        return t
    out = torch.empty(t.shape, dtype=torch.int16, device=t.device)
    E._check(E.lib().enc_cast(E.ENC_BF16, t.data_ptr(), out.data_ptr(), t.numel(), E._stream()), "cast")
    return out


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("dtype", [E.ENC_F32, E.ENC_BF16])
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_gemm_layouts(gpu, dtype, ta, tb):
    """Every operand layout against an FP64 product of the same (rounded) inputs;
    M, N, K not equal so a transposed operand cannot pass by accident."""
    import torch
    rng = np.random.default_rng(1)
    M, N, K = 128, 192, 96
    A = rng.normal(size=(K, M) if ta else (M, K)).astype(np.float32)
    B = rng.normal(size=(N, K) if tb else (K, N)).astype(np.float32)
    if dtype == E.ENC_BF16:
        A, B = bf16_round(A), bf16_round(B)
    ref = (A.T if ta else A).astype(np.float64) @ (B.T if tb else B).astype(np.float64)
    C = torch.empty((M, N), dtype=torch.float32, device="cuda")
    dA, dB = _dev(A, dtype), _dev(B, dtype)   # held until the product is read back
    E._check(E.lib().enc_gemm(dtype, ta, tb, E.EPI_STORE, dA.data_ptr(), dB.data_ptr(), M, N, K,
                              A.shape[1], B.shape[1], C.data_ptr(), None, None, None, N, E._stream()), "gemm")
    assert _rel(C.cpu().numpy(), ref) < (1e-6 if dtype == E.ENC_F32 else 1e-5)


# (M, N, K) -> (tile rows, tile columns, K splits) of a plain-store product:
# every tile, and K split over 8 and 4 workgroups (the fold-in network's dW2
# and dW1 shapes)
MFMA_SHAPES = {(8192, 1024, 128): (128, 128, 1), (4096, 1024, 192): (128, 64, 1), (2112, 2048, 64): (64, 128, 1),
               (4096, 512, 1856): (64, 64, 1), (512, 320, 4096): (64, 64, 8), (1856, 512, 4096): (64, 64, 4)}


@pytest.mark.parametrize("shape", list(MFMA_SHAPES))
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_mfma_gemm_tiles_and_layouts(gpu, shape, ta, tb):
    """The 16x16x32 MFMA kernel in each tile, K split and operand layout against
    an FP64 product of the same BF16 inputs."""
    import torch
    M, N, K = shape
    assert E.gemm_plan(M, N, K) == MFMA_SHAPES[shape]
    rng = np.random.default_rng(3)
    A = bf16_round(rng.normal(size=(K, M) if ta else (M, K)).astype(np.float32))
    B = bf16_round(rng.normal(size=(N, K) if tb else (K, N)).astype(np.float32))
    ref = (A.T if ta else A).astype(np.float64) @ (B.T if tb else B).astype(np.float64)
    C = torch.empty((M, N), dtype=torch.float32, device="cuda")
    dA, dB = _dev(A, E.ENC_BF16), _dev(B, E.ENC_BF16)
    E._check(E.lib().enc_gemm(E.ENC_BF16, ta, tb, E.EPI_STORE, dA.data_ptr(), dB.data_ptr(), M, N, K,
                              A.shape[1], B.shape[1], C.data_ptr(), None, None, None, N, E._stream()), "gemm")
    assert _rel(C.cpu().numpy(), ref) < 1e-5


def test_mfma_gemm_epilogues(gpu):
    import torch
    rng = np.random.default_rng(4)
    M, N, K = 4096, 1024, 128
    assert E.gemm_tile(M, N, K) == (128, 64)
    A = bf16_round(rng.normal(size=(M, K)).astype(np.float32))
    B = bf16_round(rng.normal(size=(K, N)).astype(np.float32) * 0.1)
    bias = rng.normal(size=N).astype(np.float32)
    Zin = rng.normal(size=(M, N)).astype(np.float32)
    acc = A.astype(np.float64) @ B.astype(np.float64)
    cf = torch.empty((M, N), dtype=torch.float32, device="cuda")
    ct = torch.empty((M, N), dtype=torch.int16, device="cuda")
    db, dz = torch.from_numpy(bias).cuda(), torch.from_numpy(Zin).cuda()
    dA, dB = _dev(A, E.ENC_BF16), _dev(B, E.ENC_BF16)
    L, st = E.lib(), E._stream()
    bf = lambda t: (t.cpu().numpy().astype(np.uint16).astype(np.uint32) << 16).view(np.float32)
    E._check(L.enc_gemm(E.ENC_BF16, 0, 0, E.EPI_BIAS, dA.data_ptr(), dB.data_ptr(), M, N, K, K, N, cf.data_ptr(),
                        None, db.data_ptr(), None, N, st), "gemm")
    np.testing.assert_allclose(cf.cpu().numpy(), acc + bias, atol=2e-5)
    E._check(L.enc_gemm(E.ENC_BF16, 0, 0, E.EPI_BIAS_GELU, dA.data_ptr(), dB.data_ptr(), M, N, K, K, N,
                        cf.data_ptr(), ct.data_ptr(), db.data_ptr(), None, N, st), "gemm")
    np.testing.assert_allclose(cf.cpu().numpy(), acc + bias, atol=2e-5)
    np.testing.assert_allclose(bf(ct), R.gelu(acc + bias), rtol=8e-3, atol=1e-6)
    E._check(L.enc_gemm(E.ENC_BF16, 0, 0, E.EPI_GELU_BWD, dA.data_ptr(), dB.data_ptr(), M, N, K, K, N,
                        cf.data_ptr(), ct.data_ptr(), None, dz.data_ptr(), N, st), "gemm")
    g = acc * R.gelu_grad(Zin)
    np.testing.assert_allclose(cf.cpu().numpy(), g, atol=2e-5)
    np.testing.assert_allclose(bf(ct), g, rtol=8e-3, atol=1e-6)


def test_colsum_matches_fp64(gpu):
    import torch
    rng = np.random.default_rng(5)
    for M, N in ((4096, 320), (4096, 512), (100, 70)):
        X = rng.normal(size=(M, N)).astype(np.float32)
        out = torch.empty(N, dtype=torch.float32, device="cuda")
        dx = torch.from_numpy(X).cuda()
        E._check(E.lib().enc_colsum(dx.data_ptr(), M, N, N, out.data_ptr(), E._stream()), "colsum")
        np.testing.assert_allclose(out.cpu().numpy(), X.astype(np.float64).sum(0), rtol=0, atol=1e-3)


@pytest.mark.parametrize("dtype", [E.ENC_F32, E.ENC_BF16])
def test_gemm_epilogues(gpu, dtype):
    import torch
    rng = np.random.default_rng(2)
    M, N, K = 64, 128, 64
    A = bf16_round(rng.normal(size=(M, K)).astype(np.float32))
    B = bf16_round(rng.normal(size=(K, N)).astype(np.float32) * 0.2)
    bias = rng.normal(size=N).astype(np.float32)
    Zin = rng.normal(size=(M, N)).astype(np.float32)
    acc = A.astype(np.float64) @ B.astype(np.float64)
    tdt = torch.float32 if dtype == E.ENC_F32 else torch.int16
    cf = torch.empty((M, N), dtype=torch.float32, device="cuda")
    ct = torch.empty((M, N), dtype=tdt, device="cuda")
    db, dz = torch.from_numpy(bias).cuda(), torch.from_numpy(Zin).cuda()
    dA, dB = _dev(A, dtype), _dev(B, dtype)
    L, st = E.lib(), E._stream()
    E._check(L.enc_gemm(dtype, 0, 0, E.EPI_BIAS_GELU, dA.data_ptr(), dB.data_ptr(), M, N, K, K, N, cf.data_ptr(),
                        ct.data_ptr(), db.data_ptr(), None, N, st), "gemm")
    np.testing.assert_allclose(cf.cpu().numpy(), acc + bias, atol=2e-5)
    h = ct.cpu().numpy() if dtype == E.ENC_F32 else (ct.cpu().numpy().astype(np.uint16).astype(np.uint32) << 16).view(np.float32)
    np.testing.assert_allclose(h, R.gelu(acc + bias), atol=1e-5 if dtype == E.ENC_F32 else 2e-2)
    E._check(L.enc_gemm(dtype, 0, 0, E.EPI_GELU_BWD, dA.data_ptr(), dB.data_ptr(), M, N, K, K, N, cf.data_ptr(),
                        ct.data_ptr(), None, dz.data_ptr(), N, st), "gemm")
    np.testing.assert_allclose(cf.cpu().numpy(), acc * R.gelu_grad(Zin), atol=2e-5)


@pytest.mark.parametrize("dtype,tol", [("f32", 1e-5), ("bf16", 2e-2)])
def test_gradients_match_fp64_oracle(gpu, dtype, tol):
    """Forward and backward kernels at the real layer widths (576 -> 512 ->
    320, 257 outputs) against the FP64 oracle: relative error ~1e-2 for BF16,
    ~1e-5 for the FP32 debug build."""
    import torch
    rng = np.random.default_rng(3)
    enc = E.FoldInEncoder(576, 257, E.EncoderConfig(dtype=dtype, seed=5))
    W1, b1, W2, b2 = enc.weights()
    b1 = rng.normal(size=b1.shape).astype(np.float32) * 0.1
    b2 = rng.normal(size=b2.shape).astype(np.float32) * 0.1
    enc.set_weights(W1, b1, W2, b2)
    B = 128
    X = rng.normal(size=(B, 576)).astype(np.float32)
    T = np.zeros((B, enc.d_out), np.float32)
    T[:, :257] = rng.normal(size=(B, 257))
    colw = np.zeros(enc.d_out, np.float32)
    colw[:257] = 1.0
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    g = enc.gradients(enc.to_device_x(X), torch.from_numpy(T).cuda(), torch.from_numpy(colw).cuda(), loss)
    rnd = bf16_round if dtype == "bf16" else (lambda a: a)
    val, ref = R.grads([rnd(W1), b1, rnd(W2), b2], rnd(X), T, colw)
    assert abs(loss.item() - val) / val < tol
    for k, (a, b) in enumerate(zip(g, ref)):
        assert _rel(a.cpu().numpy(), b) < tol, (k, _rel(a.cpu().numpy(), b))


def test_training_fits_a_teacher(gpu):
    """Adam on the GPU kernels recovers a random teacher network's mapping."""
    rng = np.random.default_rng(4)
    X = rng.normal(size=(8192, 96)).astype(np.float32)
    Wt = rng.normal(size=(96, 33)).astype(np.float32) / np.sqrt(96)
    T = np.tanh(X @ Wt).astype(np.float32)
    enc = E.FoldInEncoder(96, 33, E.EncoderConfig(hidden=128, batch=512, epochs=40, lr=3e-3))
    hist = enc.fit(X, T)
    assert hist[-1] < 0.1 * hist[0], hist
    P = enc.predict(X[:1000])
    assert np.mean((P - T[:1000]) ** 2) < 0.1 * np.var(T)


@pytest.mark.parametrize("dtype", ["bf16", "f32"])
def test_train_step_equals_separate_launches(gpu, dtype):
    """enc_train_step (one call per batch, used by fit) leaves every master,
    moment and copy bit-identical to gradients() + adam_step(), three steps in
    a row at the fold-in network's widths (the MFMA GEMMs and split-K among
    them)."""
    import torch
    rng = np.random.default_rng(6)
    d_in, n_out, B = 1856, 257, 1024
    X = rng.normal(size=(B, d_in)).astype(np.float32)
    T = rng.normal(size=(B, n_out)).astype(np.float32)
    nets = [E.FoldInEncoder(d_in, n_out, E.EncoderConfig(dtype=dtype)) for _ in range(2)]
    a, b = nets
    Xd = a.to_device_x(X)
    Tp = np.zeros((B, a.d_out), np.float32)
    Tp[:, :n_out] = T
    Td = torch.from_numpy(Tp).cuda()
    colw = torch.ones(a.d_out, dtype=torch.float32, device="cuda")
    la = torch.zeros(1, dtype=torch.float64, device="cuda")
    lb = torch.zeros(1, dtype=torch.float64, device="cuda")
    for _ in range(3):
        a.adam_step(a.gradients(Xd, Td, colw, la))
        b.train_step(Xd, Td, colw, lb)
    torch.cuda.synchronize()
    # the loss sums block partials with FP64 atomics (order free): equal to rounding
    assert la.item() == pytest.approx(lb.item(), rel=1e-12)
    for ta, tb in zip(a.masters + [a.W1c, a.W2c] + [t for mv in a.opt for t in mv],
                      b.masters + [b.W1c, b.W2c] + [t for mv in b.opt for t in mv]):
        assert torch.equal(ta, tb)


def test_warm_start_beats_random_start(gpu):
    """New users folded into a trained model (10k x 2k x 1M shape, f=32): users
    9000.. are held out of training; the encoder, trained on the SGD factors of
    users 0..8999, gives a lower holdout RMSE than the reference's random start
    before any user-only epoch and after one."""
    tr, pr = synthetic("small")
    (tr_a, pr_a), (tr_b, pr_b) = E.split_users(tr, pr, 9000)
    cfg = lm_config(n_feat=32, n_epochs=8, n_time_bins=30)
    base = tx.TxModel(tr_a, pr_a, cfg)
    for e in range(1, 9):
        base.fit_epoch(epoch=e)
    side = E.ItemSide(base, tr_a)
    Xa = E.user_features(side, tr_a)
    Ta = E.sgd_targets(base)
    enc = E.FoldInEncoder(Xa.shape[1], Ta.shape[1], E.EncoderConfig(batch=512, epochs=30))
    enc.fit(Xa, Ta)
    rand = E.fold_in_model(base, tr_b, pr_b, cfg)
    r_rand, _ = E.fold_in_curve(rand, pr_b, 2)
    warm = E.fold_in_model(base, tr_b, pr_b, cfg)
    E.warm_start(warm, enc, E.user_features(side, tr_b))
    r_warm, _ = E.fold_in_curve(warm, pr_b, 2)
    print("random", [round(x, 5) for x in r_rand], "encoder", [round(x, 5) for x in r_warm])
    assert r_warm[0] < r_rand[0] - 0.002
    assert r_warm[1] < r_rand[1]


def test_item_day_stats_match_oracle(gpu):
    tr, _ = synthetic("tiny_synth")
    cfg = lm_config(n_feat=8, n_time_bins=30)
    m = tx.TxModel(tr, synthetic("tiny_synth")[1], cfg)
    side = E.ItemSide(m, tr)
    mean, inv = R.item_day_stats(tr.item_idxs, tr.dates, tr.n_items)
    np.testing.assert_allclose(side.day_mean.cpu().numpy(), mean, rtol=1e-6)
    np.testing.assert_allclose(side.day_inv.cpu().numpy(), inv, rtol=1e-6)
    m.close()


@pytest.mark.parametrize("f", [32, 60, 256])
@pytest.mark.parametrize("max_items", [16, 512])
def test_user_features_match_oracle(gpu, f, max_items):
    """enc_user_features against the FP64 oracle on the 600 x 150 x 30k set
    (users up to 150 ratings, so max_items = 16 samples most of them), with
    the model's q rounded to BF16 as the kernel reads it and the device's
    own per-item day statistics."""
    tr, pr = synthetic("tiny_synth")
    cfg = lm_config(n_feat=f, n_time_bins=30)
    m = tx.TxModel(tr, pr, cfg)
    m.fit_epoch(epoch=1)
    side = E.ItemSide(m, tr)
    X = E.user_features(side, tr, max_items).cpu().numpy()
    q = bf16_round(m.get("ifeat").reshape(-1, f))
    ref = R.user_features(q, m.get("ibias"), float(m.get("gbias")[0]), side.day_mean.cpu().numpy(),
                          side.day_inv.cpu().numpy(), E.user_offsets(tr), tr.item_idxs, tr.raw_ratings, tr.residuals,
                          tr.dates, max_items)
    assert X.shape == ref.shape == (tr.n_users, E.n_features(f))
    err = np.abs(X - ref)
    print(f"f={f} max_items={max_items}: max |err| {err.max():.2e}")
    np.testing.assert_allclose(X, ref, rtol=1e-4, atol=2e-5)
    m.close()


def test_set_param_device_round_trip(gpu):
    """tx_set_param_device copies rows of a pitched device matrix into p_u and
    b_u exactly; tables with a skip flag are refused."""
    import torch
    tr, pr = synthetic("tiny_synth")
    f = 24
    m = tx.TxModel(tr, pr, lm_config(n_feat=f, n_time_bins=30))
    rng = np.random.default_rng(5)
    ld = f + 7
    A = rng.normal(size=(tr.n_users, ld)).astype(np.float32)
    d = torch.from_numpy(A).cuda()
    m.set_device("ufeat", d.data_ptr(), tr.n_users, ld)
    m.set_device("ubias", d.data_ptr() + 4 * f, tr.n_users, ld)
    np.testing.assert_array_equal(m.get("ufeat").reshape(-1, f), A[:, :f])
    np.testing.assert_array_equal(m.get("ubias"), A[:, f])
    with pytest.raises(tx.TxError):
        m.set_device("ybias", d.data_ptr(), tr.n_items, ld)
    with pytest.raises(tx.TxError):
        m.set_device("ufeat", d.data_ptr(), tr.n_users - 1, ld)
    m.close()
