"""CPU reference of the learned coarse solver's network (test infrastructure).

This module checks the library's device network; nothing in the product path
imports it.  It restates the same U-Net (4 levels, 16/32/48/64 channels, 3x3
kernels without biases, ReLU, 2x2 average pooling, nearest upsampling with
skip concatenation, a 1x1 output layer), the energy loss and the full backward
pass in numpy.

Two modes:
  * bf16=True follows the device's rounding points: the input, every layer
    output, every pooled value and every pre-activation gradient are rounded to
    BF16 (round to nearest even) and weights are used through BF16 copies;
    sums are accumulated in float64 here (float32 on the device), so results
    agree to a few BF16 ulps rather than bitwise;
  * bf16=False is plain float64, used to pin the backward pass against finite
    differences.

The network has no counterpart in the reference (its Poisson solve is PETSc's
KSP, lapl_lib.F90:50-54, 104); the architecture and the loss are this
library's.
"""
from __future__ import annotations

import numpy as np

# (mode, ca, cb, cout, level); mode 0 plain, 1 pooled input, 2 up-concat
LAYERS = [
    (0, 16, 0, 16, 0), (0, 16, 0, 16, 0),
    (1, 16, 0, 32, 1), (0, 32, 0, 32, 1),
    (1, 32, 0, 48, 2), (0, 48, 0, 48, 2),
    (1, 48, 0, 64, 3), (0, 64, 0, 64, 3),
    (2, 64, 48, 48, 2), (0, 48, 0, 48, 2),
    (2, 48, 32, 32, 1), (0, 32, 0, 32, 1),
    (2, 32, 16, 16, 0), (0, 16, 0, 16, 0),
]
SKIP = {8: 5, 10: 3, 12: 1}
IN_CH, PAD_IN, OUT_CH = 5, 16, 16


def cin(l):
    m, ca, cb, _, _ = LAYERS[l]
    return ca + cb if m == 2 else ca


def offsets():
    off, o = [], 0
    for l in range(len(LAYERS)):
        off.append(o)
        o += 9 * LAYERS[l][3] * cin(l)
    off.append(o)
    return off, o + OUT_CH


def bf16(x):
    """Round float32 values to BF16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


class UNetRef:
    def __init__(self, params: np.ndarray, use_bf16: bool = True):
        self.rb = use_bf16
        self.off, n = offsets()
        assert params.size == n
        self.p = np.asarray(params, dtype=np.float64)

    def _r(self, x):
        return bf16(x).astype(np.float64) if self.rb else x

    def weights(self, l):
        co, ci = LAYERS[l][3], cin(l)
        w = self.p[self.off[l]:self.off[l] + 9 * co * ci].reshape(9, co, ci)
        return self._r(w)

    def wout(self):
        return self._r(self.p[self.off[-1]:self.off[-1] + OUT_CH])

    @staticmethod
    def shift(x, t):
        """x(p + tap t) with zero outside: x [N, H, W, C]."""
        dy, dx = t // 3 - 1, t % 3 - 1
        out = np.zeros_like(x)
        H, W = x.shape[1], x.shape[2]
        ys, yd = (slice(dy, H), slice(0, H - dy)) if dy >= 0 else (slice(0, H + dy), slice(-dy, H))
        xs, xd = (slice(dx, W), slice(0, W - dx)) if dx >= 0 else (slice(0, W + dx), slice(-dx, W))
        out[:, yd, xd, :] = x[:, ys, xs, :]
        return out

    def pool(self, x):
        if self.rb:
            x32 = x.astype(np.float32)
            s = ((x32[:, 0::2, 0::2] + x32[:, 0::2, 1::2]) + x32[:, 1::2, 0::2]) + x32[:, 1::2, 1::2]
            return self._r(s * np.float32(0.25))
        return 0.25 * (x[:, 0::2, 0::2] + x[:, 0::2, 1::2] + x[:, 1::2, 0::2] + x[:, 1::2, 1::2])

    @staticmethod
    def up(x):
        return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)

    def layer_input(self, l, ys, x0):
        m = LAYERS[l][0]
        src = x0 if l == 0 else ys[l - 1]
        if m == 0:
            return src
        if m == 1:
            return self.pool(src)
        return np.concatenate([self.up(src), ys[SKIP[l]]], axis=3)

    def conv(self, x, w):
        """sum_t x(p + t) . w[t]^T : x [N,H,W,ci], w [9, co, ci]."""
        y = 0.0
        for t in range(9):
            y = y + self.shift(x, t) @ w[t].T
        return y

    def forward(self, in5):
        """in5 [N, H, W, 5] -> (e_hat [N, H, W], activations)."""
        N, H, W, _ = in5.shape
        x0 = np.zeros((N, H, W, PAD_IN))
        x0[..., :IN_CH] = in5
        x0 = self._r(x0)
        ys, xs = [], []
        for l in range(len(LAYERS)):
            xin = self.layer_input(l, ys, x0)
            xs.append(xin)
            ys.append(self._r(np.maximum(self.conv(xin, self.weights(l)), 0.0)))
        e = ys[-1] @ self.wout()
        return e, (x0, xs, ys)

    @staticmethod
    def apply_op(e, op):
        d, cx, cy = op[:, 0], op[:, 1], op[:, 2]
        ae = d * e
        ae[:, :, 1:] -= cx[:, :, :-1] * e[:, :, :-1]
        ae[:, :, :-1] -= cx[:, :, :-1] * e[:, :, 1:]
        ae[:, 1:, :] -= cy[:, :-1, :] * e[:, :-1, :]
        ae[:, :-1, :] -= cy[:, :-1, :] * e[:, 1:, :]
        return ae

    def loss(self, e, rhat, op):
        """Mean over the batch of (e.A e - 2 e.r_hat) / HW, and dL/de."""
        N, H, W = e.shape
        ae = self.apply_op(e, op)
        L = float(np.sum(e * ae - 2.0 * e * rhat) / (H * W) / N)
        return L, 2.0 * (ae - rhat) / (H * W * N)

    def backward(self, dout, acts):
        x0, xs, ys = acts
        g = np.zeros(self.p.size)
        y13 = ys[-1]
        g[self.off[-1]:] = np.einsum("nhw,nhwc->c", dout, y13)
        dpre = [None] * len(LAYERS)
        dpre[13] = self._r((y13 > 0) * (dout[..., None] * self.wout()[None, None, None, :]))
        dxcat = {}
        for l in range(len(LAYERS) - 1, -1, -1):
            m, ca, cb, co, lv = LAYERS[l]
            w = self.weights(l)
            dw = np.zeros((9, co, cin(l)))
            for t in range(9):
                dw[t] = np.einsum("nhwo,nhwi->oi", dpre[l], self.shift(xs[l], t))
            g[self.off[l]:self.off[l] + dw.size] = dw.ravel()
            if l == 0:
                break
            # input gradient: taps flipped, channel axes swapped
            dx = 0.0
            for t in range(9):
                dx = dx + self.shift(dpre[l], t) @ w[8 - t]
            if m == 0:
                dpre[l - 1] = self._r((ys[l - 1] > 0) * dx)
            elif m == 2:
                dxcat[lv] = dx
                s = ((dx[:, 0::2, 0::2, :ca] + dx[:, 0::2, 1::2, :ca]) + dx[:, 1::2, 0::2, :ca]) \
                    + dx[:, 1::2, 1::2, :ca]
                dpre[l - 1] = self._r((ys[l - 1] > 0) * s)
            else:
                lvs = lv - 1
                dl = 12 - 2 * lvs
                cat_ca = LAYERS[dl][1]
                s = dxcat[lvs][..., cat_ca:] + 0.25 * self.up(dx)
                dpre[l - 1] = self._r((ys[l - 1] > 0) * s)
        return g


def he_init(seed: int) -> np.ndarray:
    """A random parameter vector of the right shape (not the library's init)."""
    rng = np.random.default_rng(seed)
    off, n = offsets()
    p = np.zeros(n)
    for l in range(len(LAYERS)):
        co, ci = LAYERS[l][3], cin(l)
        fan = 9 * (IN_CH if l == 0 else ci)
        w = rng.normal(0, np.sqrt(2.0 / fan), size=(9, co, ci))
        if l == 0:
            w[..., IN_CH:] = 0.0
        p[off[l]:off[l] + w.size] = w.ravel()
    p[off[-1]:] = rng.uniform(-0.125, 0.125, OUT_CH)
    return p
