"""Pins of the network oracle (oracle/unet_oracle.py), which checks the GPU
training kernels.  The network has no counterpart in the reference, so the
oracle is pinned by the mathematics:
  - central finite differences of the FP64 loss against its analytic
    gradient, for parameters of every layer (relative error < 1e-6);
  - a planted mistake (input gradient through reversed taps) fails that pin;
  - the convolution against a direct loop on a tiny input, the pooling and
    upsampling adjoints (<Ax, y> = <x, A^T y>), and the parameter layout
    against the count the C ABI documents.
"""
import numpy as np
import pytest

import unet_oracle as uo


def finite_difference_check(plant=None, n_params=40, seed=3, weighted=False):
    rng = np.random.default_rng(seed)
    p = uo.he_params(seed)
    x = rng.uniform(0.0, 3.0, (1, 8, 8))
    t = x + rng.normal(0.0, 0.5, x.shape)
    w = rng.uniform(0.0, 4.0, x.shape) if weighted else None
    net = uo.UNetOracle(p, emulate_bf16=False, plant=plant)
    _, g = net.gradients(x, t, w)
    # parameters from every layer (weights and biases) and the head
    idx, off = [], 0
    for co, c1, c2 in uo.SPEC:
        nw = co * 9 * (c1 + c2)
        W = p[off:off + nw].reshape(co, 3, 3, c1 + c2)
        live = np.flatnonzero(W.ravel() != 0)            # enc1a: channel 0 only
        idx += list(off + rng.choice(live, 2, replace=False))
        idx.append(off + nw + int(rng.integers(co)))
        off += nw + co
    idx += [off + 3, off + uo.BASE]
    worst = 0.0
    for i in idx:
        eps = 1e-6 * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (uo.UNetOracle(pp, emulate_bf16=False).loss(x, t, w)
              - uo.UNetOracle(pm, emulate_bf16=False).loss(x, t, w)) / (2 * eps)
        scale = max(abs(fd), abs(g[i]), 1e-4 * np.max(np.abs(g)))
        worst = max(worst, abs(fd - g[i]) / scale)
    return worst


def test_finite_differences_fp64():
    assert finite_difference_check() < 1e-6


def test_finite_differences_weighted_loss():
    # per-pixel weights w: loss mean(w (out - t)^2)
    assert finite_difference_check(weighted=True, seed=5) < 1e-6
    # unit weights are the plain loss
    rng = np.random.default_rng(1)
    p = uo.he_params(2)
    x = rng.uniform(0.0, 3.0, (2, 8, 8))
    t = x + rng.normal(0.0, 0.5, x.shape)
    net = uo.UNetOracle(p, emulate_bf16=False)
    l0, g0 = net.gradients(x, t)
    l1, g1 = net.gradients(x, t, np.ones_like(x))
    assert l0 == l1 and np.array_equal(g0, g1)


def test_multichannel_input():
    """Several input channels ([n][C][h][w]): with the extra channels zero the
    output equals the one-channel network's; the loss gradient for first-layer
    weights on channels 1 and 2 (and every other sampled parameter) matches
    central finite differences, so the extra channels reach the network; the
    residual adds channel 0 only."""
    rng = np.random.default_rng(8)
    p = uo.he_params(4, channels=3)
    x1 = rng.uniform(0.0, 3.0, (1, 8, 8))
    x3 = np.concatenate([x1[:, None], np.zeros((1, 2, 8, 8))], axis=1)
    a = uo.UNetOracle(p, emulate_bf16=False).forward(x1)
    b = uo.UNetOracle(p, emulate_bf16=False).forward(x3)
    assert np.array_equal(a, b)
    x3 = np.concatenate([x1[:, None], rng.uniform(0.0, 3.0, (1, 2, 8, 8))], axis=1)
    c = uo.UNetOracle(p, emulate_bf16=False).forward(x3)
    assert np.abs(c - a).max() > 1e-3                      # channels 1, 2 change the output
    t = x1 + rng.normal(0.0, 0.5, x1.shape)
    _, g = uo.UNetOracle(p, emulate_bf16=False).gradients(x3, t)
    W0 = np.arange(32 * 9 * 32).reshape(32, 3, 3, 32)
    idx = [int(W0[5, 1, 1, 1]), int(W0[7, 0, 2, 2]), int(W0[3, 2, 0, 0]), 32 * 9 * 32 + 4]
    worst = 0.0
    for i in idx:
        eps = 1e-6 * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (uo.UNetOracle(pp, emulate_bf16=False).loss(x3, t) -
              uo.UNetOracle(pm, emulate_bf16=False).loss(x3, t)) / (2 * eps)
        worst = max(worst, abs(fd - g[i]) / max(abs(fd), abs(g[i]), 1e-4 * np.max(np.abs(g))))
    assert worst < 1e-6, worst
    assert np.all(g[int(W0[0, 0, 0, 3]):int(W0[0, 0, 0, 3]) + 1] == 0)      # channel 3 carries no input


def test_planted_wrong_taps_fails_the_pin():
    assert finite_difference_check(plant="wrong_taps") > 1e-2


def test_conv_matches_direct_loop():
    rng = np.random.default_rng(5)
    x = rng.normal(size=(2, 5, 6, 3))
    W = rng.normal(size=(4, 3, 3, 3))
    out = uo.UNetOracle._conv(x, W)
    ref = np.zeros((2, 5, 6, 4))
    for n in range(2):
        for y in range(5):
            for xx in range(6):
                for co in range(4):
                    s = 0.0
                    for ky in range(3):
                        for kx in range(3):
                            yy, xs = y + ky - 1, xx + kx - 1
                            if 0 <= yy < 5 and 0 <= xs < 6:
                                s += np.dot(x[n, yy, xs], W[co, ky, kx])
                    ref[n, y, xx, co] = s
    assert np.allclose(out, ref, rtol=1e-12, atol=1e-12)


def test_pool_and_upsample_adjoints():
    rng = np.random.default_rng(7)
    x = rng.normal(size=(2, 8, 6, 3))
    y = rng.normal(size=(2, 4, 3, 3))
    assert np.isclose(np.sum(uo.UNetOracle._pool(x) * y), np.sum(x * uo.UNetOracle._pool_bwd(y)))
    assert np.isclose(np.sum(uo.UNetOracle._up(y) * x), np.sum(y * uo.UNetOracle._up_bwd(x)))


def test_param_layout():
    p = uo.he_params(1)
    assert len(p) == uo.param_count()
    layers, wh, bh = uo.unpack(p)
    assert np.array_equal(uo.pack(layers, wh, bh), p)
    assert uo.param_count() == 1_955_233


def test_bf16_rounding():
    assert uo.bf16(1.0) == 1.0
    assert uo.bf16(1.0 + 2 ** -8) == 1.0                      # tie to even
    assert uo.bf16(1.0 + 3 * 2 ** -9) == 1.0 + 2 ** -7
    assert uo.bf16(-0.0) == 0.0
