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
