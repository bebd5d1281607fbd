"""The fold-in encoder's FP64 oracle (oracle/encoder_ref.py), pinned on the
CPU: gelu against closed-form values, backward() against central finite
differences of loss(), the token features against hand-made users whose
pooled blocks and scalars are written out by hand."""
import numpy as np

from oracle import encoder_ref as R


def test_gelu_closed_form():
    # gelu(z) = z Phi(z); Phi(1) = 0.8413447460685429, Phi(-1) = 0.15865525393145707
    np.testing.assert_allclose(R.gelu([0.0, 1.0, -1.0, 8.0, -8.0]),
                               [0.0, 0.8413447460685429, -0.15865525393145707, 8.0, 0.0], atol=1e-12)
    z = np.linspace(-4, 4, 81)
    h = 1e-6
    fd = (R.gelu(z + h) - R.gelu(z - h)) / (2 * h)
    np.testing.assert_allclose(R.gelu_grad(z), fd, atol=1e-8)


def _problem(seed=0, m=7, d_in=6, hid=5, d_out=4):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(m, d_in))
    T = rng.normal(size=(m, d_out))
    colw = rng.uniform(0.5, 2.0, size=d_out)
    params = [rng.normal(size=(d_in, hid)) * 0.7, rng.normal(size=hid) * 0.3,
              rng.normal(size=(hid, d_out)) * 0.7, rng.normal(size=d_out) * 0.3]
    return X, T, colw, params


def test_backward_matches_finite_differences():
    X, T, colw, params = _problem()
    _, g = R.grads(params, X, T, colw)
    h = 1e-6
    for k, P in enumerate(params):
        fd = np.zeros_like(P)
        for idx in np.ndindex(P.shape):
            old = P[idx]
            P[idx] = old + h
            lp = R.loss(params, X, T, colw)
            P[idx] = old - h
            lm = R.loss(params, X, T, colw)
            P[idx] = old
            fd[idx] = (lp - lm) / (2 * h)
        np.testing.assert_allclose(g[k], fd, rtol=1e-6, atol=1e-8, err_msg=f"param {k}")


def test_loss_gradient_closed_form():
    # one linear layer's worth: with H fixed, dloss/db2 = 2/M sum_r colw (Y - T)
    X, T, colw, params = _problem(seed=3)
    Z, H, Y = R.forward(X, *params)
    val, dY = R.loss_and_grad_y(Y, T, colw)
    assert val == np.sum(colw * (Y - T) ** 2) / len(X)
    np.testing.assert_allclose(R.backward(X, params[2], Z, H, dY)[3], 2.0 / len(X) * (colw * (Y - T)).sum(0))


# ---------------------------------------------------------------------------
# features: hand-made users with the pooled blocks written out by hand
# ---------------------------------------------------------------------------

Q = np.array([[1.0, 0.0, 0.0], [0.0, 2.0, 0.0], [0.0, 0.0, 4.0], [1.0, 1.0, 1.0]])
IBIAS = np.array([0.1, 0.2, 0.3, 0.4])
MU = 3.0


def _day_stats():
    # item 0 on days 10, 20; item 1 on day 100; item 2 on days 0, 4, 8; item 3 never
    return R.item_day_stats([0, 0, 1, 2, 2, 2], [10, 20, 100, 0, 4, 8], 4)


def test_item_day_stats_closed_form():
    mean, inv = _day_stats()
    np.testing.assert_allclose(mean, [15.0, 100.0, 4.0, 0.0])
    np.testing.assert_allclose(inv, [1 / 5.0, 1.0, 1 / np.sqrt(32 / 3), 0.0])  # std < 1 -> 1; no ratings -> 0


def _users():
    # user 0: (item 0, class 5, day 20), (item 1, class 1, day 100), (item 2, class 5, day 8)
    # user 1: no ratings
    # user 2: five ratings of classes 2, 3, 4, 2, 3 on items 3, 0, 1, 2, 3
    off = np.array([0, 3, 3, 8])
    items = np.array([0, 1, 2, 3, 0, 1, 2, 3])
    ratings = np.array([5, 1, 5, 2, 3, 4, 2, 3])
    targets = ratings.astype(np.float64)
    days = np.array([20, 100, 8, 50, 60, 70, 80, 90])
    return off, items, ratings, targets, days


def test_user_features_closed_form():
    mean, inv = _day_stats()
    off, items, ratings, targets, days = _users()
    X = R.user_features(Q, IBIAS, MU, mean, inv, off, items, ratings, targets, days, max_items=512)
    f = 3
    blk = lambda u, b: X[u, b * f:(b + 1) * f]
    d = np.array([(20 - 15) / 5.0, 0.0, (8 - 4) / np.sqrt(32 / 3)])   # per-item normalised days
    w = np.array([5 - 3 - 0.1, 1 - 3 - 0.2, 5 - 3 - 0.3])             # residuals against mu + b_i
    np.testing.assert_allclose(blk(0, 0), [0, 2 / 3, 0])                 # class 1: q_1 / 3
    for b in (1, 2, 3):
        np.testing.assert_allclose(blk(0, b), 0.0)
    np.testing.assert_allclose(blk(0, 4), [1 / 3, 0, 4 / 3])             # class 5: (q_0 + q_2) / 3
    np.testing.assert_allclose(blk(0, 5), [d[0] / 3, 0, 4 * d[2] / 3])
    np.testing.assert_allclose(blk(0, 6), [w[0] / 3, 2 * w[1] / 3, 4 * w[2] / 3])
    sc = X[0, 7 * f:]
    np.testing.assert_allclose(sc[:15], [np.log(4) / 8, 1 / 3, 0, 0, 0, 2 / 3, w.mean(), w.std(), d.mean(), d.std(),
                                         128 / 3 / 2243, 92 / 2243, 8 / 2243, 100 / 2243, 1.0])
    np.testing.assert_allclose(sc[15:], 0.0)
    # no ratings: zeros and the constant column
    expect = np.zeros(7 * f + 64)
    expect[7 * f + 14] = 1.0
    np.testing.assert_array_equal(X[1], expect)


def test_user_features_sample_positions():
    """Two of user 2's five ratings are pooled: positions floor(t * 5 / 2) = 0, 2
    (item 3 in class 2, item 1 in class 4); the scalars still see all five."""
    mean, inv = _day_stats()
    off, items, ratings, targets, days = _users()
    X = R.user_features(Q, IBIAS, MU, mean, inv, off, items, ratings, targets, days, max_items=2)
    f = 3
    np.testing.assert_allclose(X[2, 1 * f:2 * f], Q[3] / 2)
    np.testing.assert_allclose(X[2, 2 * f:3 * f], 0.0)               # class 3 rated twice, never sampled
    np.testing.assert_allclose(X[2, 3 * f:4 * f], Q[1] / 2)
    np.testing.assert_allclose(X[2, 7 * f + 1:7 * f + 6], [0, 2 / 5, 2 / 5, 1 / 5, 0])
    # item 3 has no day statistics (inv 0): its normalised day is 0
    np.testing.assert_allclose(X[2, 5 * f:6 * f], ((70 - 100) * 1.0 * Q[1]) / 2)


def test_mfma_gemm_tile_rule():
    """enc_gemm_tile (host-only, no device work): the largest tile that divides
    M and N in at least 512 tiles, 64 x 64 otherwise, the generic kernel for
    FP32 and for K not a multiple of 64; the fold-in network's five products."""
    from netflix_prize_amd import encoder as E
    assert E.gemm_tile(8192, 1024, 128) == (128, 128)
    assert E.gemm_tile(4096, 1024, 128) == (128, 64)
    assert E.gemm_tile(2112, 2048, 64) == (64, 128)
    assert E.gemm_tile(4096, 512, 1856) == (64, 64)       # forward, layer 1: 512 tiles
    assert E.gemm_tile(4096, 320, 512) == (64, 64)        # forward, layer 2
    assert E.gemm_tile(512, 320, 4096) == (64, 64)        # dW2
    assert E.gemm_tile(4096, 512, 320) == (64, 64)        # dZ
    assert E.gemm_tile(1856, 512, 4096) == (64, 64)       # dW1
    assert E.gemm_tile(4096, 1024, 96) == (0, 0)
    assert E.gemm_tile(4096, 1024, 128, E.ENC_F32) == (0, 0)
    # K splits: plain-store products under 512 tiles only
    assert E.gemm_plan(1856, 512, 4096) == (64, 64, 4)       # dW1: 232 tiles
    assert E.gemm_plan(512, 320, 4096) == (64, 64, 8)        # dW2: 40 tiles
    assert E.gemm_plan(512, 320, 4096, E.EPI_BIAS) == (64, 64, 1)
    assert E.gemm_plan(4096, 512, 320) == (64, 64, 1)        # 512 tiles
    assert E.gemm_plan(512, 320, 512) == (64, 64, 2)         # 4 stages per split at most
