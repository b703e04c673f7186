"""CPU: the FP64 CNN oracle's backward pass pinned by central finite
differences on a tiny network."""
import numpy as np

from oracle import cnn_ref as R


def test_cnn_oracle_gradient_finite_differences():
    rng = np.random.default_rng(0)
    C, NB = 32, 1
    _, n = R.layout(C, NB)
    p = rng.normal(0, 0.05, n)
    x = rng.normal(0, 1, (2, 5, 6, 2))
    t = rng.normal(0, 1, (2, 5, 6, 2))
    loss, g = R.loss_and_grad(p, x, t, C, NB)
    idx = rng.choice(n, 40, replace=False)
    idx = np.concatenate([idx, [0, n - 1, n - 3]])   # head, tail bias, tail weight
    h = 1e-6
    for i in idx:
        e = np.zeros(n); e[i] = h
        lp = np.sum((R.forward(p + e, x, C, NB) - t) ** 2) / x.size
        lm = np.sum((R.forward(p - e, x, C, NB) - t) ** 2) / x.size
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(fd)) + 1e-8, (i, fd, g[i])


def test_cnn_oracle_conv_definition():
    """A single-tap weight reproduces a shift with zero padding."""
    x = np.arange(2 * 3 * 4 * 1, dtype=float).reshape(2, 3, 4, 1)
    w = np.zeros((1, 9, 1)); w[0, 5, 0] = 1.0     # tap 5: dy = 0, dx = +1
    y = R.conv(x, w, np.zeros(1))
    assert np.array_equal(y[:, :, :-1], x[:, :, 1:]) and np.all(y[:, :, -1] == 0)
