"""CPU restatement (numpy, FP64) of the concentration-reconstruction network
of libwacomm_hip.so (wc_unet_*), used ONLY as the checker of its GPU
forward and backward kernels.

TEST INFRASTRUCTURE: only tests/ import this module.  The product never does.

The network has no counterpart in the reference; this file restates the
architecture documented in include/wacomm_hip.h (section 4) from the
mathematics of convolution, pooling, upsampling and the mean squared error:

  in   = x as channel 0 of 32 channels
  enc1 = relu(conv(relu(conv(in))))             32     L0  -> skip s1
  enc2 = 2 convs on avgpool2(s1)                64     L1  -> skip s2
  enc3 = 2 convs on avgpool2(s2)                128    L2  -> skip s3
  bott = 2 convs on avgpool2(s3)                256    L3
  dec3 = 2 convs on concat(s3, up2(bott))       128    L2
  dec2 = 2 convs on concat(s2, up2(dec3))       64     L1
  dec1 = 2 convs on concat(s1, up2(dec2))       32     L0
  out  = x + sum_c wh[c] dec1[c] + bh
  loss = mean (out - t)^2

conv = 3x3, stride 1, zero padding 1, weights [cout][3][3][cin], bias; all
tensors NHWC.  With emulate_bf16=True the restatement rounds where the GPU
rounds to BF16 (weights; activations after each conv and pooling; the
masked output gradient before the weight-gradient and input-gradient
products), so the remaining GPU/oracle difference is FP32 vs FP64
accumulation.  With emulate_bf16=False it is exact FP64, which the
finite-difference test pins.  plant="wrong_taps" builds a deliberately
wrong backward (the input gradient scatters through the kernel taps in
reversed order) to show that the pin catches it.
"""
from __future__ import annotations

import numpy as np

BASE = 32
# (cout, c1, c2): the conv layers in parameter order
SPEC = [(BASE, BASE, 0), (BASE, BASE, 0), (2 * BASE, BASE, 0), (2 * BASE, 2 * BASE, 0),
        (4 * BASE, 2 * BASE, 0), (4 * BASE, 4 * BASE, 0), (8 * BASE, 4 * BASE, 0), (8 * BASE, 8 * BASE, 0),
        (4 * BASE, 4 * BASE, 8 * BASE), (4 * BASE, 4 * BASE, 0), (2 * BASE, 2 * BASE, 4 * BASE),
        (2 * BASE, 2 * BASE, 0), (BASE, BASE, 2 * BASE), (BASE, BASE, 0)]


def bf16(a):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def param_count():
    n = 0
    for co, c1, c2 in SPEC:
        n += co * 9 * (c1 + c2) + co
    return n + BASE + 1


def unpack(params):
    """[(W [co][3][3][ci], b [co]) per layer], head weights, head bias."""
    p = np.asarray(params, dtype=np.float64)
    out, off = [], 0
    for co, c1, c2 in SPEC:
        nw = co * 9 * (c1 + c2)
        W = p[off:off + nw].reshape(co, 3, 3, c1 + c2)
        off += nw
        b = p[off:off + co]
        off += co
        out.append((W, b))
    wh = p[off:off + BASE]
    bh = p[off + BASE]
    return out, wh, bh


def pack(layers, wh, bh):
    parts = []
    for W, b in layers:
        parts += [W.ravel(), b]
    parts += [wh, [bh]]
    return np.concatenate(parts)


class UNetOracle:
    def __init__(self, params, emulate_bf16=True, plant=None):
        self.layers, self.wh, self.bh = unpack(params)
        self.em = emulate_bf16
        self.plant = plant
        if self.em:
            self.layers = [(bf16(W), b) for W, b in self.layers]

    def _r(self, a):
        return bf16(a) if self.em else a

    # -- primitives -----------------------------------------------------------
    @staticmethod
    def _conv(x, W):
        n, h, w, _ = x.shape
        xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
        out = np.zeros((n, h, w, W.shape[0]))
        for ky in range(3):
            for kx in range(3):
                out += np.einsum("nhwc,oc->nhwo", xp[:, ky:ky + h, kx:kx + w, :], W[:, ky, kx, :], optimize=True)
        return out

    @staticmethod
    def _wgrad(g, x):
        n, h, w, _ = x.shape
        xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
        dW = np.zeros((g.shape[3], 3, 3, x.shape[3]))
        for ky in range(3):
            for kx in range(3):
                dW[:, ky, kx, :] = np.einsum("nhwo,nhwc->oc", g, xp[:, ky:ky + h, kx:kx + w, :], optimize=True)
        return dW

    def _dgrad(self, g, W):
        n, h, w, _ = g.shape
        gp = np.zeros((n, h + 2, w + 2, W.shape[3]))
        for ky in range(3):
            for kx in range(3):
                # scatter form of the adjoint: tap (ky, kx) carries g[y][x] back to
                # the input pixel it read, x[y + ky - 1][x + kx - 1]
                Wt = W[:, 2 - ky, 2 - kx, :] if self.plant == "wrong_taps" else W[:, ky, kx, :]
                gp[:, ky:ky + h, kx:kx + w, :] += np.einsum("nhwo,oc->nhwc", g, Wt, optimize=True)
        return gp[:, 1:-1, 1:-1, :]

    @staticmethod
    def _pool(x):
        return 0.25 * (x[:, 0::2, 0::2] + x[:, 1::2, 0::2] + x[:, 0::2, 1::2] + x[:, 1::2, 1::2])

    @staticmethod
    def _pool_bwd(g):
        return 0.25 * np.repeat(np.repeat(g, 2, axis=1), 2, axis=2)

    @staticmethod
    def _up(x):
        return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)

    @staticmethod
    def _up_bwd(g):
        return g[:, 0::2, 0::2] + g[:, 1::2, 0::2] + g[:, 0::2, 1::2] + g[:, 1::2, 1::2]

    def _layer(self, l, x1, x2=None):
        W, b = self.layers[l]
        x = x1 if x2 is None else np.concatenate([x1, x2], axis=3)
        return self._r(np.maximum(self._conv(x, W) + b, 0.0))

    # -- network --------------------------------------------------------------
    def forward(self, x):
        """x [n][h][w] (one input channel) or [n][C][h][w] (C input channels,
        channel 0 the residual's) -> out [n][h][w]; keeps the activations for
        backward."""
        x = np.asarray(x, dtype=np.float64)
        xs = x[:, None] if x.ndim == 3 else x
        n, cin, h, w = xs.shape
        x = xs[:, 0]
        a = {}
        a["x"] = x
        in0 = np.zeros((n, h, w, BASE))
        in0[..., :cin] = self._r(np.moveaxis(xs, 1, 3))
        a["in0"] = in0
        a["a1"] = self._layer(0, in0)
        a["s1"] = self._layer(1, a["a1"])
        a["p1"] = self._r(self._pool(a["s1"]))
        a["a2"] = self._layer(2, a["p1"])
        a["s2"] = self._layer(3, a["a2"])
        a["p2"] = self._r(self._pool(a["s2"]))
        a["a3"] = self._layer(4, a["p2"])
        a["s3"] = self._layer(5, a["a3"])
        a["p3"] = self._r(self._pool(a["s3"]))
        a["a4"] = self._layer(6, a["p3"])
        a["b4"] = self._layer(7, a["a4"])
        a["u3"] = self._up(a["b4"])
        a["d3a"] = self._layer(8, a["s3"], a["u3"])
        a["d3"] = self._layer(9, a["d3a"])
        a["u2"] = self._up(a["d3"])
        a["d2a"] = self._layer(10, a["s2"], a["u2"])
        a["d2"] = self._layer(11, a["d2a"])
        a["u1"] = self._up(a["d2"])
        a["d1a"] = self._layer(12, a["s1"], a["u1"])
        a["d1"] = self._layer(13, a["d1a"])
        a["out"] = x + a["d1"] @ self.wh + self.bh
        self.act = a
        return a["out"]

    def loss(self, x, t, w=None):
        """mean over pixels of w (out - t)^2 (w = 1 when not given)."""
        out = self.forward(x)
        w = 1.0 if w is None else np.asarray(w, dtype=np.float64)
        return float(np.mean(w * (out - np.asarray(t, dtype=np.float64)) ** 2))

    def _conv_bwd(self, l, gy, y, x1, x2=None, need_dx=True):
        W, _ = self.layers[l]
        gm = gy * (y > 0)
        gq = self._r(gm)
        x = x1 if x2 is None else np.concatenate([x1, x2], axis=3)
        dW = self._wgrad(gq, x)
        db = gq.sum(axis=(0, 1, 2))      # the device sums the rounded gradient it feeds the products
        if not need_dx:
            return dW, db, None, None
        gx = self._dgrad(gq, W)
        c1 = x1.shape[3]
        return dW, db, gx[..., :c1], (gx[..., c1:] if x2 is not None else None)

    def gradients(self, x, t, w=None):
        """(loss, flat gradient in the parameter order) of the loss above."""
        t = np.asarray(t, dtype=np.float64)
        w = 1.0 if w is None else np.asarray(w, dtype=np.float64)
        out = self.forward(x)
        a = self.act
        P = out.size
        loss = float(np.mean(w * (out - t) ** 2))
        g = 2.0 * w * (out - t) / P
        dwh = np.einsum("nhw,nhwc->c", g, a["d1"])
        dbh = g.sum()
        grads = [None] * len(SPEC)
        gd1 = g[..., None] * self.wh
        dW, db, gd1a, _ = self._conv_bwd(13, gd1, a["d1"], a["d1a"])
        grads[13] = (dW, db)
        dW, db, gs1, gu1 = self._conv_bwd(12, gd1a, a["d1a"], a["s1"], a["u1"])
        grads[12] = (dW, db)
        gd2 = self._up_bwd(gu1)
        dW, db, gd2a, _ = self._conv_bwd(11, gd2, a["d2"], a["d2a"])
        grads[11] = (dW, db)
        dW, db, gs2, gu2 = self._conv_bwd(10, gd2a, a["d2a"], a["s2"], a["u2"])
        grads[10] = (dW, db)
        gd3 = self._up_bwd(gu2)
        dW, db, gd3a, _ = self._conv_bwd(9, gd3, a["d3"], a["d3a"])
        grads[9] = (dW, db)
        dW, db, gs3, gu3 = self._conv_bwd(8, gd3a, a["d3a"], a["s3"], a["u3"])
        grads[8] = (dW, db)
        gb4 = self._up_bwd(gu3)
        dW, db, ga4, _ = self._conv_bwd(7, gb4, a["b4"], a["a4"])
        grads[7] = (dW, db)
        dW, db, gp3, _ = self._conv_bwd(6, ga4, a["a4"], a["p3"])
        grads[6] = (dW, db)
        gs3 = gs3 + self._pool_bwd(gp3)
        dW, db, ga3, _ = self._conv_bwd(5, gs3, a["s3"], a["a3"])
        grads[5] = (dW, db)
        dW, db, gp2, _ = self._conv_bwd(4, ga3, a["a3"], a["p2"])
        grads[4] = (dW, db)
        gs2 = gs2 + self._pool_bwd(gp2)
        dW, db, ga2, _ = self._conv_bwd(3, gs2, a["s2"], a["a2"])
        grads[3] = (dW, db)
        dW, db, gp1, _ = self._conv_bwd(2, ga2, a["a2"], a["p1"])
        grads[2] = (dW, db)
        gs1 = gs1 + self._pool_bwd(gp1)
        dW, db, ga1, _ = self._conv_bwd(1, gs1, a["s1"], a["a1"])
        grads[1] = (dW, db)
        dW, db, _, _ = self._conv_bwd(0, ga1, a["a1"], a["in0"], need_dx=False)
        grads[0] = (dW, db)
        return loss, pack(grads, dwh, dbh)


def he_params(seed=1, channels=1):
    """He-normal weights (fan-in of the real inputs: `channels` input
    channels), small random biases and head: a generic point for gradient
    checks."""
    rng = np.random.default_rng(seed)
    layers = []
    for l, (co, c1, c2) in enumerate(SPEC):
        cin = c1 + c2
        fan = 9 * (channels if l == 0 else cin)
        W = rng.normal(0.0, np.sqrt(2.0 / fan), (co, 3, 3, cin))
        if l == 0:
            W[..., channels:] = 0.0
        layers.append((W, rng.normal(0.0, 0.05, co)))
    return pack(layers, rng.normal(0.0, 0.1, BASE), 0.01)
