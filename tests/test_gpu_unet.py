"""GPU tests of the concentration-reconstruction network (wc_unet_*): its
forward and backward kernels (BF16 MFMA, FP32 accumulation) against the
FP64 oracle (oracle/unet_oracle.py, pinned by finite differences), which
rounds to BF16 where the device does.  The remaining difference is FP32 vs
FP64 accumulation, which now and then moves a value across a BF16 rounding
boundary (one BF16 ulp is 0.4 %) or flips a ReLU mask; on the small coarse
levels of these images (60 and 15 pixels for 24 x 40) a few such flips move
a layer's gradient by percents.  Measured over 2x16x16, 1x24x40 and 4x32x32
with two kernels of different summation order: whole gradient 2e-4 .. 1.3e-2,
worst layer 1.5e-3 .. 4.4e-2, and both kernels equally far from the exact
FP64 oracle.  The bars are relative L2 errors: forward < 5e-3, loss < 1e-3,
whole gradient < 3e-2, every layer's weight and bias gradient < 1e-1.  A
layout or indexing mistake gives O(1) errors in the layer it touches.
"""
import numpy as np
import pytest

import unet_oracle as uo
import wacommplusplus_amd as wc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def net():
    n = wc.UNet(0, seed=11)
    yield n
    n.close()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def test_param_layout(net):
    assert net.n_params == uo.param_count()
    layers, wh, bh = uo.unpack(net.get_params())
    W0 = layers[0][0]
    assert np.all(W0[..., 1:] == 0) and np.any(W0[..., 0] != 0)      # enc1a sees channel 0 only
    assert all(np.all(b == 0) for _, b in layers)


# (1, 16, 256): rows as wide as the 256-cell training crops, whose stage
# leaves no LDS room for the weights (the register-weight tiled kernel).
# (1, 8, 528): rows too wide for any tiled kernel at the two finest levels,
# so the general convolution writes FP32 gradients that a second launch
# masks (and, for the pooled layer, pools back), and every weight gradient
# takes the split-K path with bias partials from the masked gradient.
@pytest.mark.parametrize("shape", [(2, 16, 16), (1, 24, 40), (4, 32, 32), (1, 16, 256), (1, 8, 528)])
def test_forward_and_gradients_match_oracle(net, shape):
    rng = np.random.default_rng(sum(shape))
    p = uo.he_params(seed=4).astype(np.float32)
    net.set_params(p)
    x = rng.uniform(0.0, 3.0, shape).astype(np.float32)
    t = (x + rng.normal(0.0, 0.3, shape)).astype(np.float32)
    ora = uo.UNetOracle(p.astype(np.float64), emulate_bf16=True)
    y, ms = net.forward(x)
    y_ref = ora.forward(x)
    print(f"{shape}: forward rel err {rel(y, y_ref):.2e} (device {ms:.3f} ms)", flush=True)
    assert rel(y, y_ref) < 5e-3
    loss, g = net.gradients(x, t)
    loss_ref, g_ref = ora.gradients(x, t)
    assert abs(loss - loss_ref) / loss_ref < 1e-3, (loss, loss_ref)
    off = 0
    worst = 0.0
    for l, (co, c1, c2) in enumerate(uo.SPEC):
        nw = co * 9 * (c1 + c2)
        ew = rel(g[off:off + nw], g_ref[off:off + nw])
        eb = rel(g[off + nw:off + nw + co], g_ref[off + nw:off + nw + co])
        worst = max(worst, ew, eb)
        assert ew < 1e-1 and eb < 1e-1, (l, ew, eb)
        off += nw + co
    eh = rel(g[off:], g_ref[off:])
    assert eh < 3e-2, eh
    whole = rel(g, g_ref)
    assert whole < 3e-2, whole
    print(f"{shape}: gradient rel err {whole:.2e}, worst layer {worst:.2e}, head {eh:.2e}", flush=True)


def test_weighted_loss_matches_oracle(net):
    """Per-pixel weighted loss mean w (y - t)^2 against the oracle, to the bars
    above; unit weights give the unweighted loss and gradient."""
    rng = np.random.default_rng(29)
    shape = (2, 16, 16)
    p = uo.he_params(seed=4).astype(np.float32)
    net.set_params(p)
    x = rng.uniform(0.0, 3.0, shape).astype(np.float32)
    t = (x + rng.normal(0.0, 0.3, shape)).astype(np.float32)
    w = rng.uniform(0.0, 4.0, shape).astype(np.float32)
    ora = uo.UNetOracle(p.astype(np.float64), emulate_bf16=True)
    loss, g = net.gradients(x, t, w)
    loss_ref, g_ref = ora.gradients(x, t, w.astype(np.float64))
    assert abs(loss - loss_ref) / loss_ref < 1e-3, (loss, loss_ref)
    assert rel(g, g_ref) < 3e-2, rel(g, g_ref)
    loss0, g0 = net.gradients(x, t)
    loss1, g1 = net.gradients(x, t, np.ones(shape, np.float32))
    assert abs(loss1 - loss0) <= 1e-6 * loss0 and rel(g1, g0) < 1e-6
    assert rel(g, g0) > 1e-2                      # the weights do change the gradient


def test_three_channel_input_matches_oracle():
    """A 3-channel network ([n][3][h][w] input) against the oracle, as above."""
    net3 = wc.UNet(0, seed=5, channels=3)
    try:
        rng = np.random.default_rng(17)
        p = uo.he_params(seed=6, channels=3).astype(np.float32)
        net3.set_params(p)
        x = rng.uniform(0.0, 3.0, (2, 3, 32, 32)).astype(np.float32)
        t = (x[:, 0] + rng.normal(0.0, 0.3, (2, 32, 32))).astype(np.float32)
        ora = uo.UNetOracle(p.astype(np.float64), emulate_bf16=True)
        y, _ = net3.forward(x)
        ey = rel(y, ora.forward(x))
        loss, g = net3.gradients(x, t)
        loss_ref, g_ref = ora.gradients(x, t)
        eg = rel(g, g_ref)
        print(f"3 channels: forward rel err {ey:.2e}, gradient {eg:.2e}", flush=True)
        assert ey < 5e-3 and abs(loss - loss_ref) / loss_ref < 1e-3 and eg < 3e-2
        W0 = g[:32 * 9 * 32].reshape(32, 3, 3, 32)
        assert np.all(W0[..., 3:] == 0) and np.any(W0[..., 2] != 0)
        with pytest.raises(ValueError):
            net3.forward(x[:, 0])
    finally:
        net3.close()


def blobs(rng, n, h, w, k=3):
    """Smooth positive fields (sums of Gaussians) and noisy copies of them."""
    yy, xx = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    clean = np.zeros((n, h, w))
    for q in range(n):
        for _ in range(k):
            cy, cx, s = rng.uniform(4, h - 4), rng.uniform(4, w - 4), rng.uniform(2, 5)
            clean[q] += rng.uniform(20, 200) * np.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2 * s * s))
    noisy = rng.poisson(clean / 10.0) * 10.0
    return np.sqrt(noisy).astype(np.float32), np.sqrt(clean).astype(np.float32)


def test_training_reduces_loss(net):
    net.set_params(uo.he_params(seed=7).astype(np.float32))
    rng = np.random.default_rng(9)
    x, t = blobs(rng, 64, 32, 32)
    losses = []
    for step in range(120):
        b = rng.choice(64, 8, replace=False)
        losses.append(net.train_step(x[b], t[b], 1e-3))
    first, last = np.mean(losses[:10]), np.mean(losses[-10:])
    print(f"training loss {first:.3f} -> {last:.3f}", flush=True)
    assert last < 0.5 * first
    # the trained network beats its input on held-out fields
    xh, th = blobs(rng, 16, 32, 32)
    y, _ = net.forward(xh)
    assert np.linalg.norm(y - th) < 0.8 * np.linalg.norm(xh - th)


def test_rejects_bad_shapes(net):
    with pytest.raises(wc.WcError):
        net.forward(np.zeros((1, 12, 16), np.float32))


def test_graph_replay_equals_direct_launches():
    """Training steps replayed from the captured graph equal the same steps
    launched one by one, bit for bit, on every replay: eight steps that
    switch batch shape and loss kind (each switch captures anew), with the
    parameters and the loss read back between replays.  Every sum over
    pixels runs in a fixed order where each level's width is a multiple of
    8 (the tiled weight gradient), so the two runs can be compared exactly."""
    rng = np.random.default_rng(21)
    p0 = uo.he_params(seed=5).astype(np.float32)
    batches = []
    for shape, weighted in (((4, 16, 64), False), ((4, 16, 64), False), ((2, 32, 128), False),
                            ((4, 16, 64), True), ((4, 16, 64), True), ((4, 16, 64), False),
                            ((2, 32, 128), True), ((4, 16, 64), False)):
        x = rng.uniform(0.0, 3.0, shape).astype(np.float32)
        t = (x + rng.normal(0.0, 0.3, shape)).astype(np.float32)
        w = rng.uniform(0.5, 2.0, shape).astype(np.float32) if weighted else None
        batches.append((x, t, w))
    runs = {}
    for graph in (True, False):
        n = wc.UNet(0, seed=3)
        n.set_graph(graph)
        n.set_params(p0)
        out = []
        for x, t, w in batches:
            loss = n.train_step(x, t, 1e-3, weight=w)
            out.append((loss, n.get_params()))
        lg, g = n.gradients(*batches[0][:2])
        out.append((lg, g))
        n.close()
        runs[graph] = out
    for q, ((la, pa), (lb, pb)) in enumerate(zip(runs[True], runs[False])):
        assert la == lb, (q, la, lb)
        assert np.array_equal(pa, pb), q
