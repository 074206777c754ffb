"""The warm-start network: a 4-level conv encoder-decoder on MFMA.

Input  (B, 64, 128, 34) bf16: cos and sin of H on the top 16 depth
       slices (32 channels) over the 64-plane x 128-pixel field, plus
       constant planes for the normalised s_dev and thickness.
Output (B, 64, 128, 2) fp32: the dislocation part of the shooting unknown
       v = D_g(0) (real, imaginary), scaled.

Encoder 24/32/48 channels with 2x2 average pooling, a 64-channel
bottleneck, a decoder with nearest x2 upsampling and skip concatenation,
two 3x3 convolutions (ReLU) per level and a 1x1 output: 246,978
parameters.  Weights are FP32 masters with BF16 copies; every product
runs in the hand-written kernels of libdyfox.so (BF16 operands on MFMA,
FP32 accumulation); Adam updates the masters.  The same network is
restated in NumPy FP64 in oracle/network.py for the gradient checks.
"""

import ctypes as C

import numpy as np

from .. import _native as N

H0, W0, IN_CH = 64, 128, 34

# (kind, src, src_off, cin, dst, dst_off, cout, relu) for conv3/conv1,
# (kind, src, src_off, c, dst) for pool, (kind, src, 0, c, dst, dst_off)
# for up; buffers (h, w, channels) relative to the input resolution.
SPEC = [
    ("conv3", "in", 0, 34, "e1a", 0, 24, 1),
    ("conv3", "e1a", 0, 24, "cat1", 0, 24, 1),
    ("pool", "cat1", 0, 24, "p1"),
    ("conv3", "p1", 0, 24, "e2a", 0, 32, 1),
    ("conv3", "e2a", 0, 32, "cat2", 0, 32, 1),
    ("pool", "cat2", 0, 32, "p2"),
    ("conv3", "p2", 0, 32, "e3a", 0, 48, 1),
    ("conv3", "e3a", 0, 48, "cat3", 0, 48, 1),
    ("pool", "cat3", 0, 48, "p3"),
    ("conv3", "p3", 0, 48, "ba", 0, 64, 1),
    ("conv3", "ba", 0, 64, "bb", 0, 64, 1),
    ("up", "bb", 0, 64, "cat3", 48),
    ("conv3", "cat3", 0, 112, "d3a", 0, 48, 1),
    ("conv3", "d3a", 0, 48, "d3b", 0, 48, 1),
    ("up", "d3b", 0, 48, "cat2", 32),
    ("conv3", "cat2", 0, 80, "d2a", 0, 32, 1),
    ("conv3", "d2a", 0, 32, "d2b", 0, 32, 1),
    ("up", "d2b", 0, 32, "cat1", 24),
    ("conv3", "cat1", 0, 56, "d1a", 0, 24, 1),
    ("conv3", "d1a", 0, 24, "d1b", 0, 24, 1),
    ("conv1", "d1b", 0, 24, "out", 0, 2, 0),
]

# buffer -> (downsampling level, channels)
BUFFERS = {"in": (0, 34), "e1a": (0, 24), "cat1": (0, 56), "p1": (1, 24),
           "e2a": (1, 32), "cat2": (1, 80), "p2": (2, 32), "e3a": (2, 48),
           "cat3": (2, 112), "p3": (3, 48), "ba": (3, 64), "bb": (3, 64),
           "d3a": (2, 48), "d3b": (2, 48), "d2a": (1, 32), "d2b": (1, 32),
           "d1a": (0, 24), "d1b": (0, 24), "out": (0, 2)}


def param_layout(spec=SPEC):
    """[(op index, w offset, w shape, b offset, cout)] and total size."""
    out, off = [], 0
    for i, op in enumerate(spec):
        if op[0] not in ("conv3", "conv1"):
            continue
        cin, cout = op[3], op[6]
        k = 9 * cin if op[0] == "conv3" else cin
        out.append((i, off, (cout, k), off + cout * k, cout))
        off += cout * k + cout
    return out, off


def init_params(rng, spec=SPEC):
    """He-normal weights (the output layer scaled by 0.1), zero biases;
    conv3 weights are (cout, 9 cin) with column (ky*3 + kx) cin + ci."""
    layout, n = param_layout(spec)
    p = np.zeros(n, np.float32)
    for li, (i, wo, (cout, k), bo, _) in enumerate(layout):
        std = np.sqrt(2.0 / k) * (0.1 if li == len(layout) - 1 else 1.0)
        p[wo:wo + cout * k] = rng.normal(0.0, std, cout * k)
    return p


def to_bf16_bits(a):
    """Round-to-nearest-even float32 -> bfloat16 bits (uint16)."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def from_bf16_bits(b):
    return (np.asarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32)


def _vp(a, byte_off=0):
    return C.c_void_p(a.ptr + byte_off)


class WarmStartNet:
    """Device-resident network, its optimiser state and work buffers."""

    def __init__(self, batch, params=None, spec=SPEC, h0=H0, w0=W0, seed=0):
        self.B, self.spec, self.h0, self.w0 = int(batch), spec, h0, w0
        self.layout, self.n_params = param_layout(spec)
        if params is None:
            params = init_params(np.random.default_rng(seed), spec)
        self.w = N.DeviceArray.from_host(np.asarray(params, np.float32))
        self.wb = N.DeviceArray.from_host(to_bf16_bits(params))
        self.g = N.DeviceArray.zeros((self.n_params,), np.float32)
        self.m = N.DeviceArray.zeros((self.n_params,), np.float32)
        self.v = N.DeviceArray.zeros((self.n_params,), np.float32)
        self.t = 0
        self.loss = N.DeviceArray.zeros((1,), np.float32)
        self.act, self.grad, self.cols = {}, {}, {}
        for name, (lvl, c) in BUFFERS.items():
            h, w = h0 >> lvl, w0 >> lvl
            n = self.B * h * w * c
            self.act[name] = N.DeviceArray.empty(
                (n,), np.float32 if name == "out" else np.uint16)
            self.grad[name] = N.DeviceArray.empty((n,), np.float32)
        self.dz = N.DeviceArray.empty((self.B * h0 * w0 * 64,), np.uint16)
        self.dcol = N.DeviceArray.empty((self.B * h0 * w0 * 9 * 56,),
                                        np.float32)
        for i, op in enumerate(spec):
            if op[0] == "conv3":
                lvl = BUFFERS[op[1]][0]
                P = self.B * (h0 >> lvl) * (w0 >> lvl)
                self.cols[i] = N.DeviceArray.empty((P * 9 * op[3],), np.uint16)

    # ---- helpers ----------------------------------------------------------
    def _dims(self, name):
        lvl, c = BUFFERS[name]
        return self.h0 >> lvl, self.w0 >> lvl, c

    def _wb(self, li):
        _, wo, _, _, _ = self.layout[li]
        return _vp(self.wb, 2 * wo)

    def params(self):
        return self.w.get()

    # ---- forward ------------------------------------------------------------
    def forward(self, x_bf16):
        """x: device uint16 array of B*64*128*34 bf16; returns the output
        buffer (device FP32, B*64*128*2)."""
        self._x = x_bf16
        li = 0
        for i, op in enumerate(self.spec):
            kind = op[0]
            src = x_bf16 if op[1] == "in" else self.act[op[1]]
            h, w, ctot = self._dims(op[1])
            if kind == "conv3":
                _, _, cin, dst, doff, cout, relu = op[1:]
                _, wo, (co, k), bo, _ = self.layout[li]
                dh, dw, dct = self._dims(dst)
                P = self.B * h * w
                col = self.cols[i]
                N.call("dyf_nn_im2col3", _vp(src), self.B, h, w, ctot, op[2],
                       cin, _vp(col), None)
                N.call("dyf_nn_gemm", _vp(col), 9 * cin, 1, self._wb(li),
                       9 * cin, 1, P, cout, 9 * cin, 0, relu,
                       _vp(self.w, 4 * bo), _vp(self.act[dst], 2 * doff), dct,
                       1, None)
                li += 1
            elif kind == "conv1":
                _, soff, cin, dst, doff, cout, relu = op[1:]
                _, wo, _, bo, _ = self.layout[li]
                P = self.B * h * w
                N.call("dyf_nn_gemm", _vp(src, 2 * soff), ctot, 1, self._wb(li),
                       cin, 1, P, cout, cin, 1, relu, _vp(self.w, 4 * bo),
                       _vp(self.act[dst], 4 * doff), cout, 1, None)
                li += 1
            elif kind == "pool":
                _, soff, c, dst = op[1:]
                N.call("dyf_nn_pool", _vp(src), self.B, h, w, ctot, soff, c,
                       _vp(self.act[dst]), None)
            elif kind == "up":
                _, _, c, dst, doff = op[1:]
                _, _, dct = self._dims(dst)
                N.call("dyf_nn_up", _vp(src), self.B, h, w, c,
                       _vp(self.act[dst]), dct, doff, None)
        return self.act["out"]

    # ---- backward -----------------------------------------------------------
    def backward(self, target_f32):
        """MSE against target (device FP32, same size as the output);
        fills self.g with the parameter gradients; returns the device loss."""
        N.call("dyf_nn_fill", _vp(self.g), self.n_params, 0.0, None)
        N.call("dyf_nn_fill", _vp(self.loss), 1, 0.0, None)
        for name, arr in self.grad.items():
            N.call("dyf_nn_fill", _vp(arr), arr.size, 0.0, None)
        n_out = self.act["out"].size
        N.call("dyf_nn_mse", _vp(self.act["out"]), _vp(target_f32), n_out,
               _vp(self.grad["out"]), _vp(self.loss), None)
        li = len(self.layout)
        for i in range(len(self.spec) - 1, -1, -1):
            op = self.spec[i]
            kind = op[0]
            src_name = op[1]
            src = self._x if src_name == "in" else self.act[src_name]
            h, w, ctot = self._dims(src_name)
            P = self.B * h * w
            if kind in ("conv3", "conv1"):
                li -= 1
                _, soff, cin, dst, doff, cout, relu = op[1:]
                _, wo, (co, k), bo, _ = self.layout[li]
                _, _, dct = self._dims(dst)
                N.call("dyf_nn_relu_grad", _vp(self.grad[dst]),
                       _vp(self.act[dst]) if dst != "out" else _vp(self.grad[dst]),
                       P, dct, doff, cout, relu, _vp(self.dz), None)
                # split the pixel sum: 1024 pixels per workgroup, at most
                # 256 splits (measured: 1.86 -> 0.65 ms per backward pass
                # for the 14 3x3 layers at batch 32, against 4096 / 64)
                splits = max(1, min(256, P // 1024))
                if kind == "conv3":
                    col = self.cols[i]
                    N.call("dyf_nn_gemm", _vp(self.dz), 1, cout, _vp(col), 1,
                           9 * cin, cout, 9 * cin, P, 2, 0, None,
                           _vp(self.g, 4 * wo), 9 * cin, splits, None)
                else:
                    N.call("dyf_nn_gemm", _vp(self.dz), 1, cout,
                           _vp(src, 2 * soff), 1, ctot, cout, cin, P, 2, 0,
                           None, _vp(self.g, 4 * wo), cin, splits, None)
                N.call("dyf_nn_bias_grad", _vp(self.dz), P, cout,
                       _vp(self.g, 4 * bo), None)
                if src_name == "in":
                    continue
                if kind == "conv3":
                    N.call("dyf_nn_gemm", _vp(self.dz), cout, 1, self._wb(li),
                           1, 9 * cin, P, 9 * cin, cout, 1, 0, None,
                           _vp(self.dcol), 9 * cin, 1, None)
                    N.call("dyf_nn_col2im3", _vp(self.dcol), self.B, h, w, cin,
                           _vp(self.grad[src_name]), ctot, soff, None)
                else:
                    N.call("dyf_nn_gemm", _vp(self.dz), cout, 1, self._wb(li),
                           1, cin, P, cin, cout, 2, 0, None,
                           _vp(self.grad[src_name], 4 * soff), ctot, 1, None)
            elif kind == "pool":
                _, soff, c, dst = op[1:]
                N.call("dyf_nn_pool_grad", _vp(self.grad[dst]), self.B, h, w,
                       c, _vp(self.grad[src_name]), ctot, soff, None)
            elif kind == "up":
                _, _, c, dst, doff = op[1:]
                _, _, dct = self._dims(dst)
                N.call("dyf_nn_up_grad", _vp(self.grad[dst]), self.B, h, w, c,
                       dct, doff, _vp(self.grad[src_name]), None)
        return self.loss

    def step(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        self.t += 1
        N.call("dyf_nn_adam", _vp(self.w), _vp(self.g), _vp(self.m),
               _vp(self.v), _vp(self.wb), self.n_params, float(lr),
               float(b1), float(b2), float(eps), self.t, None)
