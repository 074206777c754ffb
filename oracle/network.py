"""The warm-start network in NumPy FP64 (oracle; test use only).

Restates, layer for layer, the network of dyfox_amd/warmstart/net.py so
its device forward and backward passes can be checked: NHWC tensors,
3x3 convolutions with zero padding (weight (cout, 9 cin), column
(ky*3 + kx) cin + ci multiplying the input at (y + ky - 1, x + kx - 1)),
ReLU, 2x2 average pooling, nearest x2 upsampling into channel slices,
a 1x1 output convolution, and the mean-squared-error loss.  Gradients
are derived by hand and pinned against central finite differences in
tests/test_oracle_network.py.
"""

import numpy as np


def param_layout(spec):
    out, off = [], 0
    for i, op in enumerate(spec):
        if op[0] not in ("conv3", "conv1"):
            continue
        cin, cout = op[3], op[6]
        k = 9 * cin if op[0] == "conv3" else cin
        out.append((i, off, (cout, k), off + cout * k, cout))
        off += cout * k + cout
    return out, off


def _shape(buffers, name, B, h0, w0):
    lvl, c = buffers[name]
    return (B, h0 >> lvl, w0 >> lvl, c)


def _im2col(x):
    B, H, W, C = x.shape
    p = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    cols = [p[:, ky:ky + H, kx:kx + W, :] for ky in range(3) for kx in range(3)]
    return np.concatenate(cols, axis=-1)           # (B, H, W, 9C)


def _col2im(dcol, C):
    B, H, W, _ = dcol.shape
    dp = np.zeros((B, H + 2, W + 2, C))
    for tap in range(9):
        ky, kx = divmod(tap, 3)
        dp[:, ky:ky + H, kx:kx + W, :] += dcol[..., tap * C:(tap + 1) * C]
    return dp[:, 1:H + 1, 1:W + 1, :]


def forward(spec, buffers, params, x, h0, w0):
    """Output (B, h0, w0, 2) and the activations needed by backward."""
    layout, _ = param_layout(spec)
    B = x.shape[0]
    act = {n: np.zeros(_shape(buffers, n, B, h0, w0)) for n in buffers}
    act["in"] = np.asarray(x, np.float64)
    li = 0
    for op in spec:
        kind = op[0]
        if kind in ("conv3", "conv1"):
            _, src, soff, cin, dst, doff, cout, relu = op
            _, wo, (co, k), bo, _ = layout[li]
            Wt = params[wo:wo + co * k].reshape(co, k)
            b = params[bo:bo + co]
            xin = act[src][..., soff:soff + cin]
            z = (_im2col(xin) if kind == "conv3" else xin) @ Wt.T + b
            act[dst][..., doff:doff + cout] = np.maximum(z, 0) if relu else z
            li += 1
        elif kind == "pool":
            _, src, soff, c, dst = op
            a = act[src][..., soff:soff + c]
            act[dst] = 0.25 * (a[:, 0::2, 0::2] + a[:, 1::2, 0::2]
                               + a[:, 0::2, 1::2] + a[:, 1::2, 1::2])
        elif kind == "up":
            _, src, _, c, dst, doff = op
            act[dst][..., doff:doff + c] = np.repeat(
                np.repeat(act[src], 2, axis=1), 2, axis=2)
    return act["out"], act


def backward(spec, buffers, params, act, dout):
    """Gradient of sum(out * dout) with respect to the flat parameters."""
    layout, n = param_layout(spec)
    g = np.zeros(n)
    grad = {k: np.zeros_like(v) for k, v in act.items()}
    grad["out"] = np.asarray(dout, np.float64).copy()
    li = len(layout)
    for op in reversed(spec):
        kind = op[0]
        if kind in ("conv3", "conv1"):
            li -= 1
            _, src, soff, cin, dst, doff, cout, relu = op
            _, wo, (co, k), bo, _ = layout[li]
            Wt = params[wo:wo + co * k].reshape(co, k)
            dz = grad[dst][..., doff:doff + cout].copy()
            if relu:
                dz *= act[dst][..., doff:doff + cout] > 0
            xin = act[src][..., soff:soff + cin]
            xc = _im2col(xin) if kind == "conv3" else xin
            g[wo:wo + co * k] += np.einsum("bhwo,bhwk->ok", dz, xc).ravel()
            g[bo:bo + co] += dz.sum(axis=(0, 1, 2))
            dxc = dz @ Wt
            dx = _col2im(dxc, cin) if kind == "conv3" else dxc
            grad[src][..., soff:soff + cin] += dx
        elif kind == "pool":
            _, src, soff, c, dst = op
            d = np.repeat(np.repeat(grad[dst], 2, axis=1), 2, axis=2) * 0.25
            grad[src][..., soff:soff + c] += d
        elif kind == "up":
            _, src, _, c, dst, doff = op
            d = grad[dst][..., doff:doff + c]
            grad[src] += (d[:, 0::2, 0::2] + d[:, 1::2, 0::2]
                          + d[:, 0::2, 1::2] + d[:, 1::2, 1::2])
    return g


def loss_and_grad(spec, buffers, params, x, target, h0, w0):
    """Mean squared error and its gradient (flat parameters)."""
    out, act = forward(spec, buffers, params, x, h0, w0)
    d = out - target
    loss = float(np.mean(d * d))
    g = backward(spec, buffers, params, act, 2.0 * d / d.size)
    return loss, g, out
