"""The fold-in encoder's FP64 oracle (oracle/encoder_ref.py), pinned on the
CPU: gelu against closed-form values, backward() against central finite
differences of loss()."""
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
