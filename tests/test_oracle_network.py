"""Pins the NumPy network oracle (oracle/network.py): its hand-derived
gradients against central finite differences, on the real architecture
at a small spatial size."""

import numpy as np

from dyfox_amd.warmstart.net import BUFFERS, SPEC, init_params, param_layout
from oracle import network as ON


def test_parameter_count_and_layout_agree():
    layout, n = param_layout(SPEC)
    olayout, on = ON.param_layout(SPEC)
    assert n == on == 246978
    assert layout == olayout


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    h0, w0, B = 8, 16, 2
    p = init_params(rng).astype(np.float64)
    layout, n = ON.param_layout(SPEC)
    # make the output layer non-negligible so every path carries gradient
    i, wo, (co, k), bo, _ = layout[-1]
    p[wo:wo + co * k] *= 10
    x = rng.normal(size=(B, h0, w0, 34))
    t = rng.normal(size=(B, h0, w0, 2))
    loss, g, _ = ON.loss_and_grad(SPEC, BUFFERS, p, x, t, h0, w0)
    picks = np.concatenate([rng.choice(n, 40, replace=False),
                            [l[3] for l in layout]])     # every bias
    eps = 1e-6
    for j in picks:
        pp = p.copy()
        pp[j] += eps
        lp, _, _ = ON.loss_and_grad(SPEC, BUFFERS, pp, x, t, h0, w0)
        pp[j] -= 2 * eps
        lm, _, _ = ON.loss_and_grad(SPEC, BUFFERS, pp, x, t, h0, w0)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[j]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (j, fd, g[j])


def test_im2col_and_col2im_are_adjoint():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(2, 5, 7, 3))
    y = rng.normal(size=(2, 5, 7, 27))
    lhs = np.sum(ON._im2col(x) * y)
    rhs = np.sum(x * ON._col2im(y, 3))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_conv3_equals_a_direct_zero_padded_convolution():
    """The column convention (ky*3 + kx) cin + ci against four plain loops."""
    rng = np.random.default_rng(4)
    B, H, W, cin, cout = 1, 4, 5, 2, 3
    x = rng.normal(size=(B, H, W, cin))
    Wt = rng.normal(size=(cout, 9 * cin))
    got = ON._im2col(x) @ Wt.T
    want = np.zeros((B, H, W, cout))
    for y in range(H):
        for xx in range(W):
            for ky in range(3):
                for kx in range(3):
                    yy, xs = y + ky - 1, xx + kx - 1
                    if 0 <= yy < H and 0 <= xs < W:
                        for ci in range(cin):
                            want[0, y, xx] += (Wt[:, (ky * 3 + kx) * cin + ci]
                                               * x[0, yy, xs, ci])
    assert np.abs(got - want).max() < 1e-12
